"""Pins of the oracle's near field of discontinuous triangular physics (the reference's
sparse_assemble_parallel! branch for it, src/3d/assembly_collocation_losses.jl:233-257,
with connected_sources :41-96 and create_polar_elements :409-447) against mathematics,
and the host-side checks of the C ABI for it.

* polar_gaussian (src/3d/quadrature.jl:163-206): weights sum to the triangle's area 1/2,
  low moments integrate exactly and the integral of 1/rho from the vertex matches its
  closed form sqrt(2) log(1 + sqrt(2));
* create_polar_elements: for any node inside the triangle the three sub-triangles tile it
  (area, moments), and the integral of 1/|p - P| matches the closed form of the flat
  triangle seen from an interior point (a wrong sub-triangle, Jacobian or pole fails it);
* connected_sources: depth 1 the element itself, depth 2 the nonzeros of C C^T (C the
  element x corner incidence), rows element by element;
* assembled values at k = 0 on flat elements: each row's single-layer sum over its own
  element equals the closed form of the integral of 1/(4 pi r) from the node, a linear
  density p(y) = a + b.y gives the closed form of the integral of p/(4 pi r), and the
  double layer vanishes; at depth 2 the double-layer sum over each other element equals
  its solid angle / (4 pi) (van Oosterom-Strackee).
None of these restates the oracle's formula.
"""
import ctypes
from math import factorial

import numpy as np
import pytest
import scipy.sparse as sps

import oracle_np
from bie_amd import _lib
from bie_amd import mesh as M

ORDERS = {M.DTRI1: "disctriconstant", M.DTRI3: "disctrilinear", M.DTRI6: "disctriquadratic"}


def _edge_terms(x, p, q):
    """For the edge p -> q of a flat triangle around x (in its plane): the distance h of
    x to the edge's line, the outward unit normal nu and the arc parameters of p and q."""
    t = (q - p) / np.linalg.norm(q - p)
    foot = p + np.dot(x - p, t) * t
    nu = foot - x
    h = np.linalg.norm(nu)
    return h, nu / h, np.dot(p - foot, t), np.dot(q - foot, t)


def _flat_integrals(x, P):
    """Closed forms over the flat triangle P (3 x 3 columns) from x inside it:
    int 1/|y - x| dA = sum_e h_e [asinh(s/h_e)] and int (y - x)/|y - x| dA =
    sum_e nu_e [(s sqrt(h^2 + s^2) + h^2 asinh(s/h)) / 2] (divergence theorem in the
    plane: div (y - x)/r = 1/r, grad r = (y - x)/r)."""
    inv_r, vec = 0.0, np.zeros(3)
    for a, b in ((0, 1), (1, 2), (2, 0)):
        h, nu, s0, s1 = _edge_terms(x, P[:, a], P[:, b])
        inv_r += h * (np.arcsinh(s1 / h) - np.arcsinh(s0 / h))

        def Fr(s):
            return 0.5 * (s * np.sqrt(h * h + s * s) + h * h * np.arcsinh(s / h))
        vec += nu * (Fr(s1) - Fr(s0))
    return inv_r, vec


def _solid_angle(x, P1, P2, P3):
    a, b, c = P1 - x, P2 - x, P3 - x
    la, lb, lc = np.linalg.norm(a), np.linalg.norm(b), np.linalg.norm(c)
    num = np.dot(a, np.cross(b, c))
    den = la * lb * lc + np.dot(a, b) * lc + np.dot(a, c) * lb + np.dot(b, c) * la
    return 2.0 * np.arctan2(num, den)


def _flat_quadratic(f):
    """Quadratic geometry whose edge nodes are the midpoints of straight edges (flat
    six-node elements), so the closed forms of flat triangles apply."""
    mesh = M.icosphere(f, "quadratic")
    X = mesh.coordinates.copy()
    T = mesh.topology
    for s, (a, b) in zip((3, 4, 5), ((0, 1), (1, 2), (2, 0))):
        X[:, T[s]] = 0.5 * (X[:, T[a]] + X[:, T[b]])
    return M.with_coordinates(mesh, X)


# ------------------------------------------------------------------ polar rules ---
@pytest.mark.parametrize("n", [4, 7, 12, 40])
def test_polar_gaussian(n):
    x, y, w = oracle_np.polar_gaussian(n)
    nr = 2 * ((n + 1) // 2)
    assert len(w) == nr * nr and np.all(w > 0)
    assert np.all(x >= 0) and np.all(y >= 0) and np.all(x + y <= 1 + 1e-15)
    tol = {4: 1e-3, 7: 1e-6, 12: 1e-9, 40: 1e-14}[n]
    for a in range(3):
        for b in range(3 - a):
            exact = factorial(a) * factorial(b) / factorial(a + b + 2)
            assert abs(np.sum(w * x ** a * y ** b) - exact) < tol, (a, b)
    # 1/rho from the singular vertex: the radial rule is exact, only the angle is quadrature
    exact = np.sqrt(2.0) * np.log(1.0 + np.sqrt(2.0))
    assert abs(np.sum(w / np.hypot(x, y)) - exact) < tol


@pytest.mark.parametrize("kind", [M.DTRI1, M.DTRI3, M.DTRI6])
@pytest.mark.parametrize("beta", [0.1127, 0.2113])
def test_polar_elements_tile_the_triangle(kind, beta):
    un, vn = oracle_np.disc_nodes(kind, beta)
    for c in range(1, len(un) + 1):
        u, v, w = oracle_np.create_polar_elements(kind, beta, 40, c)
        assert len(w) == 3 * 1600 and np.all(w >= 0)
        assert np.all(u >= -1e-15) and np.all(v >= -1e-15) and np.all(u + v <= 1 + 1e-15)
        for a in range(3):
            for b in range(3 - a):
                exact = factorial(a) * factorial(b) / factorial(a + b + 2)
                assert abs(np.sum(w * u ** a * v ** b) - exact) < 1e-14, (c, a, b)
        # 1/|p - P| from the pole P, against the closed form of the flat triangle
        P = np.array([un[c - 1], vn[c - 1], 0.0])
        V = np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, 0.0]])
        exact, _ = _flat_integrals(P, V)
        got = np.sum(w / np.hypot(u - P[0], v - P[1]))
        # the angular rule's own error: ~1e-12, up to 4e-8 for a node at beta = 0.11 next
        # to an edge (the flat sub-triangle on that edge); a wrong sub-triangle, Jacobian or
        # pole is an error of 1e-3 or more
        assert abs(got / exact - 1) < 1e-6, c


# ------------------------------------------------------------- connected_sources ---
@pytest.mark.parametrize("gorder", ["linear", "quadratic"])
@pytest.mark.parametrize("kind", [M.DTRI1, M.DTRI3, M.DTRI6])
def test_connected_sources_match_incidence(gorder, kind):
    mesh = M.with_discontinuous_physics(M.icosphere(2, gorder), ORDERS[kind])
    E, npk = mesh.n_elements, M.physics_nodes(kind)
    corners = mesh.topology[:3]
    C = sps.csr_matrix((np.ones(3 * E), (np.repeat(np.arange(E), 3), corners.ravel(order="F"))))
    A = (C @ C.T).tocsr()
    for depth in (1, 2):
        els, srcs = oracle_np.connected_sources_disc(mesh, depth)
        for e in range(E):
            want = np.array([e]) if depth == 1 else np.sort(A[e].indices)
            assert np.array_equal(els[e], want)
            assert np.array_equal(srcs[e], np.sort(mesh.physics_topology[:, want].ravel()))
        # the library's pattern has one row per element node with its element's list
        hm = _lib.HostMesh(mesh)
        nnz = ctypes.c_int64()
        _lib.check(_lib.lib().bie_sparse_count(ctypes.byref(hm.struct), depth,
                                                ctypes.byref(nnz)))
        assert nnz.value == npk * sum(len(s) for s in srcs)


def test_library_rejects_rows_that_are_not_element_nodes():
    mesh = M.with_discontinuous_physics(M.icosphere(1, "linear"), "disctrilinear")
    hm = _lib.HostMesh(mesh)
    nnz = ctypes.c_int64()
    hm.struct.n_sources = mesh.n_sources - 1   # n_sources != n_elements * 3
    # (the topology then names a source beyond the count: rejected either way)
    assert _lib.lib().bie_sparse_count(ctypes.byref(hm.struct), 1, ctypes.byref(nnz)) == \
        _lib.BIE_EINVAL


# ------------------------------------------------------------ values at k = 0 ---
@pytest.mark.parametrize("case", ["lin_d1", "lin_d3", "lin_d6", "flatquad_d3", "flatquad_d6"])
def test_own_element_closed_forms(case):
    geo, kind = {"lin_d1": ("lin", M.DTRI1), "lin_d3": ("lin", M.DTRI3),
                 "lin_d6": ("lin", M.DTRI6), "flatquad_d3": ("fq", M.DTRI3),
                 "flatquad_d6": ("fq", M.DTRI6)}[case]
    base = M.ellipsoid(1, axes=(1.0, 0.8, 1.3)) if geo == "lin" else _flat_quadratic(1)
    mesh = M.with_discontinuous_physics(base, ORDERS[kind])
    npk = M.physics_nodes(kind)
    I, J, F, G = oracle_np.sparse_assemble_disc(mesh, 0.0, depth=1)
    rng = np.random.default_rng(3)
    a, b = 0.7, rng.standard_normal(3)
    X, T = mesh.coordinates, mesh.topology
    checked = 0
    for r in range(mesh.n_sources):
        e, j = divmod(r, npk)
        x = mesh.sources[:, r]
        # the pole is physics node (closest geometry node) - with the reference's fillers
        # (nodes 1, 2, 3 for slots 4, 5, 6; node 1 for constant physics); where that is
        # another node than the row's own (edge nodes of quadratic physics on linear
        # geometry, some nodes of linear physics on quadratic geometry) the rule is not
        # adapted to the singularity and only the reference's own inaccuracy is defined
        slot = int(np.argmin(np.linalg.norm(X[:, T[:, e]] - x[:, None], axis=0)))
        pole = 0 if kind == M.DTRI1 else (slot % 3 if kind == M.DTRI3 else slot)
        if pole != (0 if kind == M.DTRI1 else j):
            continue
        checked += 1
        inv_r, vec = _flat_integrals(x, X[:, T[:3, e]])
        sel = I == r
        cols, g = J[sel], G[sel]
        assert np.array_equal(cols, np.sort(mesh.physics_topology[:, e]))
        assert abs(g.sum().real * 4 * np.pi / inv_r - 1) < 1e-6
        if kind != M.DTRI1:  # linear densities are represented exactly
            p = a + b @ mesh.sources[:, cols]
            exact = ((a + b @ x) * inv_r + b @ vec) / (4 * np.pi)
            assert abs(np.sum(g * p).real - exact) < 1e-6 * abs(exact)
    assert checked >= mesh.n_sources // 3
    assert np.max(np.abs(G.imag)) == 0.0
    # in-plane source: the double layer is the rounding of n.d only
    assert np.max(np.abs(F)) < 1e-9 * np.max(np.abs(G))


@pytest.mark.parametrize("kind", [M.DTRI1, M.DTRI3])
def test_neighbour_double_layer_is_solid_angle(kind):
    mesh = M.with_discontinuous_physics(M.ellipsoid(1, axes=(1.0, 0.8, 1.3)), ORDERS[kind])
    npk = M.physics_nodes(kind)
    I, J, F, G = oracle_np.sparse_assemble_disc(mesh, 0.0, depth=2, gOn=False)
    X, T = mesh.coordinates, mesh.topology
    els, _ = oracle_np.connected_sources_disc(mesh, 2)
    worst = 0.0
    for r in range(0, mesh.n_sources, 3):
        e0 = r // npk
        sel = I == r
        cols, f = J[sel], F[sel]
        for e in els[e0]:
            if e == e0:
                continue
            P = X[:, T[:, e]]
            exact = _solid_angle(mesh.sources[:, r], P[:, 0], P[:, 1], P[:, 2]) / (4 * np.pi)
            got = f[np.isin(cols, mesh.physics_topology[:, e])].sum().real
            worst = max(worst, abs(got - exact))
    assert worst < 1e-9
    assert np.max(np.abs(G)) == 0.0
