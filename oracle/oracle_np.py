"""TEST INFRASTRUCTURE — independent NumPy restatement of the dense collocation assembly.

Only tests/ may import this module.  It is written separately from bie_oracle.cpp
(vectorised over targets, elements and quadrature points; Gauss-Legendre from
numpy.polynomial.legendre.leggauss) so that the two restatements check each other on
small meshes.  Reference functions restated (file:line):
  curveLinearQuadpoints        src/2d/quadrature.jl:5-11
  triangularQuadpoints         src/3d/quadrature.jl:70-82
  rotated_triangular_quadpoints src/3d/quadrature.jl:99-113
  rotated_midpoint_triangular_quadpoints src/3d/quadrature.jl:117-144
  create_rotated_element       src/3d/assembly_collocation.jl:122-136
  triangularLinear/Quadratic   src/3d/surface_functions.jl:191-233
  discontinuousTriangular{Constant,Linear,Quadratic} src/3d/surface_functions.jl:234-318
    (physics kinds 0x101, 0x103, 0x106 with the mesh's physics_beta)
  quadrilateralQuadpoints      src/3d/quadrature.jl:9-20, as create_shape_function calls it
                               (getQuadpoints(sf, m, n), assembly_collocation.jl:12-20)
  quadrilateral bases          src/3d/surface_functions.jl:1-149 (kinds 0x204 Linear4,
                               0x404 LinearSerendipity, 0x208 Quadratic, 0x209
                               QuadraticLagrange; discontinuous 0x301, 0x304, 0x309)
  generic assemble_parallel!   src/3d/assembly_collocation.jl:481-538 (one rule, no selection)
  interpolate_elements/jacobian! src/3d/assembly_collocation.jl:27-66, src/3d/jacobian.jl:7-87
  find_closest_corner          src/3d/assembly_collocation.jl:102-115
  greens3d!/freens3d!/freens3dk0! src/3d/kernels.jl:13-55
  computing_integrals!         src/3d/assembly_collocation.jl:213-252
  polar_quadrature / polar_gaussian src/3d/quadrature.jl:163-206
  create_polar_elements        src/3d/assembly_collocation_losses.jl:409-447
  connected_sources (discontinuous physics) src/3d/assembly_collocation_losses.jl:41-96
  sparse_assemble_parallel! (discontinuous branch) src/3d/assembly_collocation_losses.jl:200-405
"""
from __future__ import annotations

import numpy as np


def gl01(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


def tensor(n, m):
    nx, wx = gl01(n)
    ny, wy = gl01(m)
    u = np.kron(np.ones(m), nx)
    v = np.kron(ny, np.ones(n))
    w = np.kron(wy, wx)
    return u, v, w


QUAD_KINDS = (0x204, 0x404, 0x208, 0x209)


def is_quad(kind):
    return kind in QUAD_KINDS or kind in (0x301, 0x304, 0x309)


def quad_rule(n, m):
    """The quadrilateral rule of assemble_parallel!'s generic method: create_shape_function
    passes (m, n) to quadrilateralQuadpoints(n, m), so u runs over m Gauss-Legendre points on
    [-1, 1] and v over n."""
    xm, wm = np.polynomial.legendre.leggauss(m)
    xn, wn = np.polynomial.legendre.leggauss(n)
    u = np.kron(np.ones(n), xm)
    v = np.kron(xn, np.ones(m))
    w = np.kron(wn, wm)
    return u, v, w


def rule(c, n, m, quad=False):
    if quad:
        return quad_rule(n, m)
    if c == 1:
        u, v, w = tensor(n, m)
        return u * v, (1.0 - u) * v, w * v
    if c in (2, 3):
        u, v, w = tensor(n, m)
        xi, eta, W = u, v * (1.0 - u), w * np.abs(1.0 - u)
        return (xi, eta, W) if c == 2 else (eta, xi, W)
    mm = int(np.ceil(m / 2.0)) * 2
    u, v, w = tensor(n, mm)
    X = np.concatenate([0.5 * u, (u - 0.5) * v + 0.5])
    Y = np.concatenate([v * (1.0 - u), v * (1.0 - u)])
    W = np.concatenate([w * (0.5 * (1.0 - u)), w * (0.5 * v)])
    if c == 4:
        return X, Y, W
    if c == 5:
        return (1.0 - X) - Y, X, W
    return Y, (1.0 - X) - Y, W


def disc_basis(kind, u, v, beta):
    """Discontinuous triangle bases, N, dN/du, dN/dv with shape (Q, nodes): the linear or
    quadratic Lagrange basis in the three shrunk barycentrics
    lam = ((1-u-v, u, v) - beta) / (1 - 3 beta), whose nodes sit where lam is 0 or 1."""
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    z = np.zeros_like(u)
    if kind == 0x101:
        return np.ones((u.size, 1)), z[:, None], z[:, None]
    c = 1.0 / (1.0 - 3.0 * beta)
    lam = (np.stack([(1.0 - u) - v, u, v], axis=1) - beta) / (1.0 - 3.0 * beta)
    dlu = np.stack([-c + z, c + z, z], axis=1)    # d lam / du
    dlv = np.stack([-c + z, z, c + z], axis=1)
    if kind == 0x103:
        return lam, dlu, dlv
    pairs = [(1, 0), (1, 2), (2, 0)]               # edge nodes 4, 5, 6
    N = np.concatenate([lam * (2.0 * lam - 1.0)] +
                       [4.0 * lam[:, [a]] * lam[:, [b]] for a, b in pairs], axis=1)
    du = np.concatenate([(4.0 * lam - 1.0) * dlu] +
                        [4.0 * (dlu[:, [a]] * lam[:, [b]] + lam[:, [a]] * dlu[:, [b]])
                         for a, b in pairs], axis=1)
    dv = np.concatenate([(4.0 * lam - 1.0) * dlv] +
                        [4.0 * (dlv[:, [a]] * lam[:, [b]] + lam[:, [a]] * dlv[:, [b]])
                         for a, b in pairs], axis=1)
    return N, du, dv


def quad_basis(kind, u, v, beta=0.0):
    """Quadrilateral bases, N, dN/du, dN/dv with shape (Q, nodes), in the reference's
    formulas (surface_functions.jl:8-11, 22-49, 62-70, 81-90, 103, 123, 140)."""
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    o, z = np.ones_like(u), np.zeros_like(u)
    if kind == 0x301:
        return o[:, None], z[:, None], z[:, None]
    if kind in (0x304, 0x309):
        s = 1.0 / (1.0 - beta)
        N, du, dv = quad_basis(0x204 if kind == 0x304 else 0x209, u * s, v * s)
        return N, du * s, dv * s
    if kind == 0x204:
        N = [(1 - u) * (1 - v) / 4, (1 + u) * (1 - v) / 4, (1 - u) * (1 + v) / 4,
             (1 + u) * (1 + v) / 4]
        du = [-(1 - v) / 4, (1 - v) / 4, -(1 + v) / 4, (1 + v) / 4]
        dv = [-(1 - u) / 4, -(1 + u) / 4, (1 - u) / 4, (1 + u) / 4]
    elif kind == 0x404:
        N = [(1 - u) * (1 - v) / 4, (1 + u) * (1 - v) / 4, (1 + u) * (1 + v) / 4,
             (1 - u) * (1 + v) / 4]
        du = [-(1 - v) / 4, (1 - v) / 4, (1 + v) / 4, -(1 + v) / 4]
        dv = [-(1 - u) / 4, -(1 + u) / 4, (1 + u) / 4, (1 - u) / 4]
    elif kind == 0x208:
        N = [(1 - u) * (v - 1) * (u + v + 1) / 4, (1 + u) * (v - 1) * (v - u + 1) / 4,
             (1 + u) * (v + 1) * (u + v - 1) / 4, (u - 1) * (v + 1) * (u - v + 1) / 4,
             (1 - v) * (1 - u ** 2) / 2, (1 + u) * (1 - v ** 2) / 2,
             (1 + v) * (1 - u ** 2) / 2, (1 - u) * (1 - v ** 2) / 2]
        du = [-(v - 1) * (u + v + 1) / 4 + (1 - u) * (v - 1) / 4,
              (v - 1) * (v - u + 1) / 4 - (1 + u) * (v - 1) / 4,
              (v + 1) * (u + v - 1) / 4 + (1 + u) * (v + 1) / 4,
              (v + 1) * (u - v + 1) / 4 + (u - 1) * (v + 1) / 4,
              -(1 - v) * u, (1 - v ** 2) / 2, -(1 + v) * u, -(1 - v ** 2) / 2]
        dv = [(1 - u) * (u + v + 1) / 4 + (1 - u) * (v - 1) / 4,
              (1 + u) * (v - u + 1) / 4 + (1 + u) * (v - 1) / 4,
              (1 + u) * (u + v - 1) / 4 + (1 + u) * (v + 1) / 4,
              (u - 1) * (u - v + 1) / 4 - (u - 1) * (v + 1) / 4,
              (u ** 2 - 1) / 2, -(1 + u) * v, (1 - u ** 2) / 2, -(1 - u) * v]
    elif kind == 0x209:
        N = [u * (1 - u) * v * (1 - v) / 4, -u * (1 + u) * v * (1 - v) / 4,
             -u * (1 - u) * v * (1 + v) / 4, u * (1 + u) * v * (1 + v) / 4,
             -(1 + u) * (1 - u) * v * (1 - v) / 2, -u * (1 - u) * (1 + v) * (1 - v) / 2,
             (1 - u ** 2) * (1 - v ** 2), u * (1 + u) * (1 + v) * (1 - v) / 2,
             (1 + u) * (1 - u) * (1 + v) * v / 2]
        # derivatives of the products above (the reference takes them by ForwardDiff)
        du = [(1 - 2 * u) * v * (1 - v) / 4, -(1 + 2 * u) * v * (1 - v) / 4,
              -(1 - 2 * u) * v * (1 + v) / 4, (1 + 2 * u) * v * (1 + v) / 4,
              2 * u * v * (1 - v) / 2, -(1 - 2 * u) * (1 - v * v) / 2,
              -2 * u * (1 - v ** 2), (1 + 2 * u) * (1 - v * v) / 2, -2 * u * (1 + v) * v / 2]
        dv = [u * (1 - u) * (1 - 2 * v) / 4, -u * (1 + u) * (1 - 2 * v) / 4,
              -u * (1 - u) * (1 + 2 * v) / 4, u * (1 + u) * (1 + 2 * v) / 4,
              -(1 - u * u) * (1 - 2 * v) / 2, u * (1 - u) * 2 * v / 2,
              -(1 - u ** 2) * 2 * v, -u * (1 + u) * 2 * v / 2, (1 - u * u) * (1 + 2 * v) / 2]
    else:
        raise ValueError("not a quadrilateral kind")
    return (np.stack([x + z for x in N], axis=1), np.stack([x + z for x in du], axis=1),
            np.stack([x + z for x in dv], axis=1))


def basis(kind, u, v, beta=0.0):
    """Returns N, dN/du, dN/dv with shape (Q, nodes)."""
    if is_quad(kind):
        return quad_basis(kind, u, v, beta)
    if kind > 0xFF:
        return disc_basis(kind, u, v, beta)
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    o, z = np.ones_like(u), np.zeros_like(u)
    if kind == 3:
        N = np.stack([(1.0 - u) - v, u, v], axis=1)
        du = np.stack([-o, o, z], axis=1)
        dv = np.stack([-o, z, o], axis=1)
        return N, du, dv
    l = (1.0 - v) - u
    N = np.stack([l * ((1.0 - 2.0 * v) - 2.0 * u), u * (2.0 * u - 1.0), v * (2.0 * v - 1.0),
                  4.0 * u * l, 4.0 * u * v, 4.0 * v * l], axis=1)
    du = np.stack([-(2.0 * l - 1.0) - 2.0 * l, 4.0 * u - 1.0, z, 4.0 * l - 4.0 * u,
                   4.0 * v, -4.0 * v], axis=1)
    dv = np.stack([-(2.0 * l - 1.0) - 2.0 * l, z, 4.0 * v - 1.0, -4.0 * u, 4.0 * u,
                   4.0 * l - 4.0 * v], axis=1)
    return N, du, dv


# gauss_points_triangle (adaptive_integration.jl:13-64): rows (u, v, w), as listed there
_GPT = {
    1: [[1 / 3, 1 / 3, 0.5]],
    3: [[1 / 6, 1 / 6, 1 / 6], [2 / 3, 1 / 6, 1 / 6], [1 / 6, 2 / 3, 1 / 6]],
    4: [[1 / 3, 1 / 3, -9 / 32], [0.2, 0.6, 25 / 96], [0.2, 0.2, 25 / 96],
        [0.6, 0.2, 25 / 96]],
    7: [[1 / 3, 1 / 3, 0.1125],
        [0.797426985353087, 0.101286507323456, 0.0629695902724135],
        [0.101286507323456, 0.797426985353087, 0.0629695902724135],
        [0.101286507323456, 0.101286507323456, 0.0629695902724135],
        [0.059715871789770, 0.470142064105115, 0.066197076394253],
        [0.470142064105115, 0.059715871789770, 0.066197076394253],
        [0.470142064105115, 0.470142064105115, 0.066197076394253]],
    12: [[0.873821971016996, 0.063089014491502, 0.0254224531851035],
         [0.063089014491502, 0.873821971016996, 0.0254224531851035],
         [0.063089014491502, 0.063089014491502, 0.0254224531851035],
         [0.501426509658279, 0.249286745170910, 0.0583931378631895],
         [0.249286745170910, 0.501426509658279, 0.0583931378631895],
         [0.249286745170910, 0.249286745170910, 0.0583931378631895],
         [0.636502499121399, 0.310352451033785, 0.041425537809187],
         [0.636502499121399, 0.053145049844816, 0.041425537809187],
         [0.310352451033785, 0.636502499121399, 0.041425537809187],
         [0.310352451033785, 0.053145049844816, 0.041425537809187],
         [0.053145049844816, 0.636502499121399, 0.041425537809187],
         [0.053145049844816, 0.310352451033785, 0.041425537809187]],
    13: [[1 / 3, 1 / 3, -0.074785022233835],
         [0.479308067841923, 0.260345966079038, 0.087807628716602],
         [0.260345966079038, 0.479308067841923, 0.087807628716602],
         [0.260345966079038, 0.260345966079038, 0.087807628716602],
         [0.869739794195568, 0.065130102902216, 0.0266736178044195],
         [0.065130102902216, 0.869739794195568, 0.0266736178044195],
         [0.065130102902216, 0.065130102902216, 0.0266736178044195],
         [0.638444188569809, 0.312865496004875, 0.0385568804451285],
         [0.638444188569809, 0.048690315425316, 0.0385568804451285],
         [0.312865496004875, 0.638444188569809, 0.0385568804451285],
         [0.312865496004875, 0.048690315425316, 0.0385568804451285],
         [0.048690315425316, 0.638444188569809, 0.0385568804451285],
         [0.048690315425316, 0.312865496004875, 0.0385568804451285]],
}


def gauss_points_triangle(n):
    t = np.asarray(_GPT[n], dtype=np.float64)
    return t[:, 0].copy(), t[:, 1].copy(), t[:, 2].copy()


def julia_hypot3(a, b, c):
    """Julia Base n-ary hypot, elementwise (NaN/Inf handling as in select_exact.cpp)."""
    a, b, c = np.abs(a), np.abs(b), np.abs(c)
    with np.errstate(invalid="ignore", divide="ignore"):
        m = np.maximum(np.maximum(a, b), c)
        safe = np.where(m == 0, 1.0, m)
        s = (a / safe) ** 2
        s = s + (b / safe) ** 2
        s = s + (c / safe) ** 2
        out = m * np.sqrt(s)
    out = np.where(m == 0, 0.0, out)
    out = np.where(np.isinf(a) | np.isinf(b) | np.isinf(c), np.inf, out)
    return out


def n_rules(mesh):
    return 1 if is_quad(mesh.geometry_kind) else (6 if mesh.geometry_kind == 6 else 3)


def element_geometry(mesh, c, n, m):
    u, v, w = rule(c, n, m, is_quad(mesh.geometry_kind))
    N, du, dv = basis(mesh.geometry_kind, u, v)
    X = mesh.coordinates[:, mesh.topology]            # (3, ng, E)
    y = np.einsum("dke,qk->eqd", X, N)
    tu = np.einsum("dke,qk->eqd", X, du)
    tv = np.einsum("dke,qk->eqd", X, dv)
    cr = np.cross(tu, tv)
    J = julia_hypot3(cr[..., 0], cr[..., 1], cr[..., 2])
    nrm = -cr / J[..., None]
    return y, nrm, J * w[None, :]


def select(mesh, targets):
    """Rule index (1..R) per (target, element), first strict minimum."""
    P = mesh.sources[:, mesh.physics_topology]         # (3, np, E)
    d = julia_hypot3(P[0][None] - targets[0][:, None, None],
                     P[1][None] - targets[1][:, None, None],
                     P[2][None] - targets[2][:, None, None])   # (M, np, E)
    # first strict minimum: argmin returns the first occurrence of the minimum
    idx = np.argmin(np.where(np.isnan(d), np.inf, d), axis=1) + 1
    return np.minimum(idx, n_rules(mesh)).astype(np.int8)


def assemble(mesh, k, targets=None, n=4, m=4, fOn=True, gOn=True, cOn=True, rules=None):
    """F, G, C; rules (int8 M x E, 1-based) freezes the rule of every pair instead of
    selecting it."""
    targets = mesh.sources if targets is None else np.asarray(targets, dtype=np.float64)
    M, N, E = targets.shape[1], mesh.n_sources, mesh.n_elements
    if mesh.geometry_kind == 0x209 and n % 2 == 1 and m % 2 == 1:
        # assembly_collocation.jl:497-504: a Gauss node on the centre node
        raise ValueError("Gauss Node On Singularity.")
    R = n_rules(mesh)
    sel = select(mesh, targets) if rules is None else np.asarray(rules)
    F = np.zeros((M, N), dtype=np.complex128)
    G = np.zeros((M, N), dtype=np.complex128)
    C = np.zeros(M, dtype=np.complex128)
    k = complex(k)
    for c in range(1, R + 1):
        ii, ee = np.nonzero(sel == c)
        if ii.size == 0:
            continue
        y, nrm, jw = element_geometry(mesh, c, n, m)
        u, v, _ = rule(c, n, m, is_quad(mesh.geometry_kind))
        phi, _, _ = basis(mesh.physics_kind, u, v,
                          getattr(mesh, "physics_beta", 0.0))  # (Q, np)
        d = y[ee] - targets.T[ii][:, None, :]                 # (P, Q, 3)
        r = julia_hypot3(d[..., 0], d[..., 1], d[..., 2])
        nd = np.einsum("pqd,pqd->pq", nrm[ee], d)
        ekr = np.exp(1j * k * r)
        g = ekr / (4 * np.pi * r) * jw[ee]
        f = ekr * (1j * k * r - 1.0) * nd / (4 * np.pi * r ** 3) * jw[ee]
        k0 = -nd / (4 * np.pi * r ** 3) * jw[ee]
        cols = mesh.physics_topology[:, ee].T                 # (P, np)
        if gOn:
            np.add.at(G, (ii[:, None], cols), g @ phi)
        if fOn:
            np.add.at(F, (ii[:, None], cols), f @ phi)
        if cOn:
            np.add.at(C, ii, k0.sum(axis=1))
    return F, G, C


def partial_assemble(mesh, k, sources, element_lists, column_lists, n=3, fOn=True, gOn=True):
    """partial_assemble_parallel! (fast_multipole_method.jl:11-81) as values on the pattern
    (column_lists[i] the stored columns of row i, ascending; element_lists[i] its connected
    elements): every (i, e) integrated with gauss_points_triangle(n), the kernels and the
    physics basis of the mesh at those points, added into row i's entries of e's nodes."""
    u, v, w = gauss_points_triangle(n)
    N, du, dv = basis(mesh.geometry_kind, u, v)
    phi, _, _ = basis(mesh.physics_kind, u, v, getattr(mesh, "physics_beta", 0.0))
    X = mesh.coordinates[:, mesh.topology]                 # (3, ng, E)
    y = np.einsum("dke,qk->eqd", X, N)
    tu = np.einsum("dke,qk->eqd", X, du)
    tv = np.einsum("dke,qk->eqd", X, dv)
    cr = np.cross(tu, tv)
    J = julia_hypot3(cr[..., 0], cr[..., 1], cr[..., 2])
    nrm = -cr / J[..., None]
    jw = J * w[None, :]
    k = complex(k)
    Fv, Gv = [], []
    for i, (els, cols) in enumerate(zip(element_lists, column_lists)):
        x = sources[:, i]
        f_row = np.zeros(len(cols), dtype=np.complex128)
        g_row = np.zeros(len(cols), dtype=np.complex128)
        where = {int(c): t for t, c in enumerate(cols)}
        for e in els:
            d = y[e] - x[None, :]
            r = julia_hypot3(d[:, 0], d[:, 1], d[:, 2])
            nd = np.einsum("qd,qd->q", nrm[e], d)
            ekr = np.exp(1j * k * r)
            g = ekr / (4 * np.pi * r) * jw[e]
            f = ekr * (1j * k * r - 1.0) * nd / (4 * np.pi * r ** 3) * jw[e]
            idx = [where[int(c)] for c in mesh.physics_topology[:, e]]
            if fOn:
                np.add.at(f_row, idx, f @ phi)
            if gOn:
                np.add.at(g_row, idx, g @ phi)
        Fv.append(f_row)
        Gv.append(g_row)
    return np.concatenate(Fv), np.concatenate(Gv)


# ---- near field of discontinuous physics ---------------------------------------------------

DTRI1, DTRI3, DTRI6 = 0x101, 0x103, 0x106


def polar_gaussian(n):
    """polar_gaussian(n): n rounded up to even radii and as many angles; Gauss-Legendre
    angles on [0, pi/2], radii on [0, 1 / (cos + sin)], weights w_angle w_radius r."""
    nr = int(np.ceil(n / 2.0)) * 2
    ga, wa = np.polynomial.legendre.leggauss(nr)
    gr, wr = np.polynomial.legendre.leggauss(nr)
    th = np.pi / 4 + np.pi / 4 * ga
    thw = np.pi / 4 * wa
    rmax = 1.0 / (np.cos(th) + np.sin(th))
    rad = rmax[:, None] / 2 + rmax[:, None] / 2 * gr[None, :]     # (angle, radius)
    radw = rmax[:, None] / 2 * wr[None, :] * rad
    ang = np.repeat(th, nr)
    rad = rad.ravel()
    return rad * np.cos(ang), rad * np.sin(ang), np.repeat(thw, nr) * radw.ravel()


def disc_nodes(kind, beta):
    """get_nodal_nodes_u/_v of the discontinuous triangles (SurfaceFunction.jl:429-438)."""
    b = float(beta)
    if kind == DTRI1:
        return np.array([1 / 3]), np.array([1 / 3])
    if kind == DTRI3:
        return np.array([b, 1 - 2 * b, b]), np.array([b, b, 1 - 2 * b])
    h = (1 - b) / 2
    return np.array([b, 1 - 2 * b, b, h, h, b]), np.array([b, b, 1 - 2 * b, b, h, h])


def create_polar_elements(kind, beta, n, c):
    """(u, v, w) of create_polar_elements(physics_function, n, c) (c 1-based): the polar
    rule mapped onto the three sub-triangles (P, (0,1), (0,0)), (P, (1,0), (0,1)),
    (P, (0,0), (1,0)) around the physics node P, weights times |det| of each map."""
    X, Y, W = polar_gaussian(n)
    un, vn = disc_nodes(kind, beta)
    P = np.array([un[c - 1], vn[c - 1]])
    us, vs, ws = [], [], []
    for A, B in (((0.0, 1.0), (0.0, 0.0)), ((1.0, 0.0), (0.0, 1.0)), ((0.0, 0.0), (1.0, 0.0))):
        A, B = np.array(A), np.array(B)
        T = np.stack([P, A, B], axis=1)                     # 2 x 3
        G = T @ np.stack([1 - X - Y, X, Y])                 # 2 x Q
        Du, Dv = A - P, B - P
        ws.append(abs(Du[0] * Dv[1] - Du[1] * Dv[0]) * W)
        us.append(G[0])
        vs.append(G[1])
    return np.concatenate(us), np.concatenate(vs), np.concatenate(ws)


def connected_sources_disc(mesh, depth=1):
    """connected_sources for discontinuous physics, per ELEMENT (row r uses the lists of
    element r // n): (element lists, source lists), 0-based, sorted."""
    E = mesh.n_elements
    ptop = mesh.physics_topology
    if depth == 1:
        return [np.array([e]) for e in range(E)], [np.unique(ptop[:, e]) for e in range(E)]
    if depth != 2:
        raise ValueError("depth must be 1 or 2")
    corners = mesh.topology[:3]
    touch = {}
    for e in range(E):
        for node in corners[:, e]:
            touch.setdefault(int(node), []).append(e)
    els, srcs = [], []
    for e in range(E):
        el = np.unique(np.concatenate([touch[int(nd)] for nd in corners[:, e]]))
        els.append(el)
        srcs.append(np.unique(ptop[:, el]))
    return els, srcs


def sparse_assemble_disc(mesh, k, sources=None, depth=1, Ngauss=40, fOn=True, gOn=True):
    """sparse_assemble_parallel! for discontinuous triangular physics as (I, J, F, G):
    rows in order, columns ascending (0-based), complex values.  Each (row, element) pair
    takes the polar rule of the physics node whose index is that of the element's
    geometry node closest to the source (the reference's filler rules: nodes 1, 2, 3 for
    the missing ones, node 1 for constant physics; no finite distance: the sixth slot)."""
    kind, beta = mesh.physics_kind, float(getattr(mesh, "physics_beta", 0.0))
    if kind == DTRI6:
        Ngauss = 40
    npk, E, N = mesh.physics_topology.shape[0], mesh.n_elements, mesh.n_sources
    if N != E * npk:
        raise ValueError("one row per element node")
    src = mesh.sources if sources is None else np.asarray(sources, dtype=np.float64)
    rules = []
    for slot in range(6):
        node = 1 if kind == DTRI1 else (slot % 3 + 1 if kind == DTRI3 else slot + 1)
        u, v, w = create_polar_elements(kind, beta, Ngauss, node)
        Ng, du, dv = basis(mesh.geometry_kind, u, v)
        phi, _, _ = basis(kind, u, v, beta)
        rules.append((w, Ng, du, dv, phi))
    els, cols = connected_sources_disc(mesh, depth)
    k = complex(k)
    I, Jc, Fv, Gv = [], [], [], []
    for r in range(N):
        e0 = r // npk
        x = src[:, r]
        c_row = cols[e0]
        where = {int(c): t for t, c in enumerate(c_row)}
        f_row = np.zeros(len(c_row), dtype=np.complex128)
        g_row = np.zeros(len(c_row), dtype=np.complex128)
        for e in els[e0]:
            X = mesh.coordinates[:, mesh.topology[:, e]]            # 3 x ng
            d = julia_hypot3(X[0] - x[0], X[1] - x[1], X[2] - x[2])
            d = np.where(np.isnan(d), np.inf, d)
            slot = int(np.argmin(d)) if np.isfinite(d.min()) else 5
            w, Ng, du, dv, phi = rules[slot]
            y, tu, tv = Ng @ X.T, du @ X.T, dv @ X.T                 # Q x 3
            cr = np.cross(tu, tv)
            J = julia_hypot3(cr[:, 0], cr[:, 1], cr[:, 2])
            nrm = -cr / J[:, None]
            jw = J * w
            dd = y - x[None, :]
            rr = julia_hypot3(dd[:, 0], dd[:, 1], dd[:, 2])
            nd = np.einsum("qd,qd->q", nrm, dd)
            ekr = np.exp(1j * k * rr)
            idx = [where[int(c)] for c in mesh.physics_topology[:, e]]
            if gOn:
                np.add.at(g_row, idx, (ekr / (4 * np.pi * rr) * jw) @ phi)
            if fOn:
                f = ekr * (1j * k * rr - 1.0) * nd / (4 * np.pi * rr ** 3) * jw
                np.add.at(f_row, idx, f @ phi)
        I.append(np.full(len(c_row), r, dtype=np.int64))
        Jc.append(np.asarray(c_row, dtype=np.int64))
        Fv.append(f_row)
        Gv.append(g_row)
    return np.concatenate(I), np.concatenate(Jc), np.concatenate(Fv), np.concatenate(Gv)
