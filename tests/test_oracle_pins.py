"""Pins of the CPU oracle against mathematics and the reference's own test tolerances.

Each check is chosen so that a plausible restatement mistake (a dropped term, a wrong
sign, a transposed rule) fails it; none retypes the oracle's formula.
"""
import math
import os

import numpy as np
import pytest

import analytic
import oracle
import oracle_np
from bie_amd import mesh as M

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


# ------------------------------------------------------------------------ quadrature ---
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 8, 12, 20])
def test_gauss_legendre_matches_numpy(n):
    x, w = oracle.gauss_legendre(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.max(np.abs(x - xr)) < 4e-16
    assert np.max(np.abs(w - wr)) < 4e-15


def _tri_monomial(a, b):
    # integral of u^a v^b over the reference triangle = a! b! / (a+b+2)!
    return math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)


@pytest.mark.parametrize("c", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("n,m", [(2, 2), (3, 2), (3, 3), (4, 4)])
def test_rule_weights_and_exactness(c, n, m):
    u, v, w = oracle.rule(c, n, m)
    assert abs(w.sum() - 0.5) < 1e-15          # surface_functions_test.jl:16,37
    assert np.all(u >= 0) and np.all(v >= 0) and np.all(u + v <= 1 + 1e-15)
    # Duffy rules integrate polynomials of total degree <= min(n,m)-... exactly; check a
    # safe degree for every rule family.
    deg = min(2 * n - 2, 2 * m - 2 if c <= 3 else 2 * m - 2)
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            assert abs(np.sum(w * u ** a * v ** b) - _tri_monomial(a, b)) < 1e-14


def test_rule_collapse_points():
    # Rule c is the one that integrates a 1/r singularity sitting on node c best:
    # corner rules collapse at nodes 1, 2, 3, mid rules at the edge mid points 4, 5, 6.
    # A rule wired to the wrong node fails this by orders of magnitude.
    targets = {1: (0, 0), 2: (1, 0), 3: (0, 1), 4: (0.5, 0), 5: (0.5, 0.5), 6: (0, 0.5)}
    for c, (u0, v0) in targets.items():
        def integral(cc, n, m):
            u, v, w = oracle.rule(cc, n, m)
            return np.sum(w / np.hypot(u - u0, v - v0))
        ref = integral(c, 40, 40)
        errs = [abs(integral(cc, 4, 4) - ref) / ref for cc in range(1, 7)]
        assert errs[c - 1] < 3e-3
        assert all(errs[c - 1] * 5 < e for i, e in enumerate(errs) if i != c - 1)


def test_rule_sizes_and_mid_rounding():
    assert oracle.rule(1, 4, 4)[0].size == 16
    assert oracle.rule(4, 4, 4)[0].size == 32
    assert oracle.rule(4, 3, 3)[0].size == 2 * 3 * 4   # m rounded up to even
    assert oracle.rule(5, 3, 2)[0].size == 12


@pytest.mark.parametrize("c", [1, 2, 3, 4, 5, 6])
def test_rules_match_numpy_restatement(c):
    for n, m in [(2, 2), (3, 2), (4, 4), (3, 5)]:
        a = oracle.rule(c, n, m)
        b = oracle_np.rule(c, n, m)
        for x, y in zip(a, b):
            assert np.max(np.abs(x - y)) < 1e-15


# ------------------------------------------------------------------------- bases ------
@pytest.mark.parametrize("kind", [3, 6])
def test_basis_partition_nodal_and_derivatives(kind):
    rng = np.random.default_rng(0)
    u = rng.random(50) * 0.5
    v = rng.random(50) * 0.5
    N, du, dv = oracle.basis(kind, u, v)
    assert np.max(np.abs(N.sum(axis=1) - 1)) < 1e-14
    assert np.max(np.abs(du.sum(axis=1))) < 1e-13
    h = 1e-6
    Np, _, _ = oracle.basis(kind, u + h, v)
    Nm, _, _ = oracle.basis(kind, u - h, v)
    assert np.max(np.abs((Np - Nm) / (2 * h) - du)) < 1e-8
    Np, _, _ = oracle.basis(kind, u, v + h)
    Nm, _, _ = oracle.basis(kind, u, v - h)
    assert np.max(np.abs((Np - Nm) / (2 * h) - dv)) < 1e-8
    nodes_u = [0.0, 1.0, 0.0, 0.5, 0.5, 0.0][:kind]   # SurfaceFunction.jl:426-429
    nodes_v = [0.0, 0.0, 1.0, 0.0, 0.5, 0.5][:kind]
    Nn, _, _ = oracle.basis(kind, np.array(nodes_u), np.array(nodes_v))
    assert np.max(np.abs(Nn - np.eye(kind))) < 1e-15


# ---------------------------------------------------------------------- geometry ------
def test_flat_triangle_area_and_normal():
    coords = np.array([[0.0, 2.0, 0.0], [0.0, 0.0, 3.0], [1.0, 1.0, 1.0]])
    m = M.Mesh3d(coords, np.array([[0], [1], [2]]), coords, np.array([[0], [1], [2]]), 3, 3)
    for c in (1, 2, 3):
        y, nr, jw = oracle.element_geometry(m, c, 3, 3)
        assert abs(jw.sum() - 3.0) < 1e-14                   # area = |2 x 3| / 2
        assert np.allclose(nr, [0.0, 0.0, -1.0], atol=1e-15)  # negated (jacobian.jl:65)
        assert np.allclose(y[..., 2], 1.0)


@pytest.mark.parametrize("go,atol", [("linear", 1e-1), ("quadratic", 1e-5)])
def test_sphere_area_reference_tolerance(go, atol):
    # fmm_test.jl:11-27: sum of J*w with n=m=2 over sphere_1m_extremely_fine ~ 4 pi
    po = "linear"
    mesh = M.load_comsol(os.path.join(GOLD, "sphere_1m_extremely_fine"), go, po)
    total = 0.0
    for c in ((1,) if go == "linear" else (1,)):
        _, _, jw = oracle.element_geometry(mesh, c, 2, 2)
        total += jw.sum()
    assert abs(total - 4 * np.pi) < atol


def test_normals_point_into_body():
    mesh = M.icosphere(4, "quadratic")
    y, nr, _ = oracle.element_geometry(mesh, 2, 3, 3)
    assert np.all(np.einsum("eqd,eqd->eq", y, nr) < 0)


# ---------------------------------------------------------------------- selection -----
def test_julia_hypot_semantics():
    assert oracle.julia_hypot3(3.0, 4.0, 12.0) == 13.0
    assert oracle.julia_hypot3(0.0, 0.0, 0.0) == 0.0
    assert oracle.julia_hypot3(np.inf, np.nan, 1.0) == np.inf
    assert np.isnan(oracle.julia_hypot3(np.nan, 1.0, 1.0))
    rng = np.random.default_rng(1)
    for x in rng.standard_normal((200, 3)) * 10.0 ** rng.integers(-200, 200, (200, 1)):
        h = oracle.julia_hypot3(*x)
        s = np.max(np.abs(x))
        ref = s * np.sqrt(np.sum((x / s) ** 2))
        assert abs(h - ref) <= 4 * np.spacing(ref)


def test_selection_first_index_wins_on_ties():
    # target equidistant from all three nodes -> rule 1
    coords = np.array([[1.0, -0.5, -0.5], [0.0, np.sqrt(3) / 2, -np.sqrt(3) / 2], [0, 0, 0]])
    m = M.Mesh3d(coords, np.array([[0], [1], [2]]), coords, np.array([[0], [1], [2]]), 3, 3)
    sel = oracle.select(m, np.array([[0.0], [0.0], [5.0]]))
    assert sel[0, 0] == 1
    # exactly equidistant from nodes 2 and 3 only -> rule 2
    sel = oracle.select(m, np.array([[-2.0], [0.0], [0.0]]))
    assert sel[0, 0] == 2


def test_selection_matches_brute_force_and_numpy():
    mesh = M.icosphere(3, "quadratic")
    s1 = oracle.select(mesh)
    s2 = oracle_np.select(mesh, mesh.sources)
    assert np.array_equal(s1, s2)
    # collocation at an element's own node picks that node's rule
    for e in range(mesh.n_elements):
        for s, node in enumerate(mesh.physics_topology[:, e]):
            assert s1[node, e] == s + 1


# ---------------------------------------------------------------------- assembly ------
@pytest.mark.parametrize("meshname", ["lin2", "quad2", "quad2_linphys", "comsol_xc"])
@pytest.mark.parametrize("k", [0.0, 1.0, 5 + 0.3j])
@pytest.mark.parametrize("nm", [(2, 2), (3, 2), (4, 4)])
def test_cpp_oracle_matches_numpy_restatement(meshname, k, nm):
    mesh = {"lin2": lambda: M.icosphere(2, "linear"),
            "quad2": lambda: M.icosphere(2, "quadratic"),
            "quad2_linphys": lambda: M.icosphere(2, "quadratic", physics="linear"),
            "comsol_xc": lambda: M.load_comsol(os.path.join(GOLD, "sphere_1m_extremely_coarse"))}[meshname]()
    F, G, C = oracle.assemble(mesh, k, n=nm[0], m=nm[1])
    F2, G2, C2 = oracle_np.assemble(mesh, k, n=nm[0], m=nm[1])
    assert rel(F, F2) < 1e-13 and rel(G, G2) < 1e-13 and rel(C, C2) < 1e-13


def test_flags_leave_outputs_zero():
    mesh = M.icosphere(2, "linear")
    F, G, C = oracle.assemble(mesh, 1.0, n=2, m=2, fOn=False, cOn=False)
    assert not F.any() and not C.any() and G.any()


def test_row_sum_identity_k0():
    # C_i = sum_j F^{(k=0)}_ij (physics bases partition unity)
    for mesh in (M.icosphere(4, "linear"), M.icosphere(3, "quadratic")):
        F, _, C = oracle.assemble(mesh, 0.0, n=3, m=3)
        assert np.max(np.abs(F.sum(axis=1) - C)) < 1e-13


def test_free_term_solid_angle_and_sphere_row_sums():
    mesh = M.icosphere(4, "quadratic")
    for k in (1.0, 3.0, 5 + 0.3j):
        F, G, C = oracle.assemble(mesh, k, n=4, m=4)
        assert np.max(np.abs(C - 0.5)) < 2e-3
        rg = analytic.sphere_row_sum_G(k, 1.0)
        rf = analytic.sphere_row_sum_F(k, 1.0)
        assert np.max(np.abs(G.sum(axis=1) - rg)) / abs(rg) < 3e-3
        assert np.max(np.abs(F.sum(axis=1) - rf)) / abs(rf) < 5e-3


def test_taylor_derivatives_match_finite_differences():
    mesh = M.icosphere(2, "quadratic")
    k, h = 2.0 + 0.1j, 1e-5
    F1, G1 = oracle.assemble_taylor(mesh, k, 1, n=3, m=2)
    Fp, Gp, _ = oracle.assemble(mesh, k + h, n=3, m=2)
    Fm, Gm, _ = oracle.assemble(mesh, k - h, n=3, m=2)
    assert rel(F1, (Fp - Fm) / (2 * h)) < 1e-8
    assert rel(G1, (Gp - Gm) / (2 * h)) < 1e-8
    F2, G2 = oracle.assemble_taylor(mesh, k, 2, n=3, m=2)
    F1p, G1p = oracle.assemble_taylor(mesh, k + h, 1, n=3, m=2)
    F1m, G1m = oracle.assemble_taylor(mesh, k - h, 1, n=3, m=2)
    assert rel(F2, (F1p - F1m) / (2 * h)) < 1e-7
    F0, G0 = oracle.assemble_taylor(mesh, k, 0, n=3, m=2)
    Fa, Ga, _ = oracle.assemble(mesh, k, n=3, m=2)
    assert rel(F0, Fa) < 1e-15 and rel(G0, Ga) < 1e-15


# ------------------------------------------------------------ system-level pins -------
@pytest.mark.parametrize("go,po,tol", [("quadratic", "geometry", 2.8e-5),
                                       ("linear", "linear", 0.037),
                                       ("quadratic", "linear", 0.048)])
def test_rigid_sphere_reference_tolerance_table(go, po, tol):
    # test/3d/assembly_test.jl:12-44 on sphere_1m with the default n=m=4
    k = 2 * np.pi * 100.0 / 340.0
    mesh = M.load_comsol(os.path.join(GOLD, "sphere_1m"), go, po)
    F, _, C = oracle.assemble(mesh, k, gOn=False)
    pI = analytic.incoming_wave([np.pi / 2, 0.0], 1.0, mesh.sources, k)
    p = np.linalg.solve(F + np.diag(1.0 - C), pI)
    th = np.arccos(np.clip(mesh.sources[0], -1, 1))
    pa = analytic.plane_wave_scattering_sphere(k, 1.0, 1.0, th, 1e-6)
    assert rel(p, pa) < tol


def test_pulsating_sphere_config1_exterior():
    # 642-node linear sphere, k = 1, (n,m) = (2,2), rho0 = 1.21, c = 343, v_n = 1
    mesh = M.icosphere(8, "linear")
    Z0 = 1.21 * 343.0
    F, G, C = oracle.assemble(mesh, 1.0, n=2, m=2)
    p = np.linalg.solve(F + np.diag(1 - C), -1j * 1.0 * Z0 * (G @ np.ones(mesh.n_sources)))
    pa = analytic.pulsating_sphere_exterior(1.0, 1.0, 1.0, Z0, 1.0)
    assert np.max(np.abs(p - pa)) / abs(pa) < 5e-3


def test_pulsating_sphere_interior_reference_workflow():
    # examples/3d_pulsating_sphere.jl:11-62 on a quadratic mesh, n=m=2
    mesh = M.load_comsol(os.path.join(GOLD, "sphere_1m"))
    Z0 = 1.21 * 343.0
    k = 2 * np.pi * 54.59 / 343.0
    F, G, C = oracle.assemble(mesh, k, n=2, m=2)
    vs = 1j * Z0 * k * np.ones(mesh.n_sources)
    p = np.linalg.solve(np.diag(1.0 - C) - F, G @ vs)
    pa = analytic.pulsating_sphere_interior(k, 1.0, 1.0, Z0, 1.0)
    assert np.max(np.abs(p - pa)) / abs(pa) < 5e-3
    z = np.linspace(0.1, 0.9, 20)
    X = np.vstack([0 * z, 0 * z, z])
    Fp, Gp, _ = oracle.assemble(mesh, k, targets=X, n=2, m=2)
    pf = Fp @ p + Gp @ vs
    assert np.max(np.abs(pf - analytic.pulsating_sphere_interior(k, 1.0, z, Z0, 1.0))) / abs(pa) < 1e-2


def test_scaled_hypot_is_the_plainly_rounded_sum_of_squares():
    """orc_hypot3_scaled (the other n-ary hypot of Julia releases) equals
    sqrt((x*x + y*y) + z*z) with every operation rounded once, computed by numpy, bit for
    bit away from overflow and underflow; inf beats NaN, zero stays zero, and huge or tiny
    arguments neither overflow nor underflow (math.hypot within 2 ulp)."""
    import math
    L = oracle.lib()
    rng = np.random.default_rng(3)
    X = rng.standard_normal((20000, 3)) * np.exp(rng.uniform(-30, 30, (20000, 1)))
    ref = np.sqrt((X[:, 0] * X[:, 0] + X[:, 1] * X[:, 1]) + X[:, 2] * X[:, 2])
    got = np.array([L.orc_hypot3_scaled(*x) for x in X])
    assert np.array_equal(got, ref)
    assert L.orc_hypot3_scaled(np.inf, np.nan, 1.0) == np.inf
    assert np.isnan(L.orc_hypot3_scaled(np.nan, 1.0, 2.0))
    assert L.orc_hypot3_scaled(0.0, -0.0, 0.0) == 0.0
    for s in (1e300, 1e-300, 5e-324 * 7):
        x = (3.0 * s, 4.0 * s, 12.0 * s)
        h = L.orc_hypot3_scaled(*x)
        assert np.isfinite(h) and h > 0 and abs(h - math.hypot(*x)) <= 2 * math.ulp(math.hypot(*x))
    # the two formulas round differently on some arguments (here: in the last place)
    assert L.orc_hypot3_scaled(0.6732655185893088, 0.3428080423874833, 0.13687617154257523) == \
        0.7678143647933542
    assert L.orc_julia_hypot3(0.6732655185893088, 0.3428080423874833, 0.13687617154257523) == \
        0.7678143647933543


def test_scaled_hypot_selection_by_brute_force():
    """oracle.select(hypot="scaled") against a numpy brute force of find_closest_corner with
    that distance, on a tie-heavy mesh and targets."""
    mesh = M.icosphere(4, "linear")
    rng = np.random.default_rng(9)
    c = mesh.coordinates
    i, j = rng.integers(0, c.shape[1], (2, 150))
    T = np.hstack([0.5 * (c[:, i] + c[:, j]), c[:, :80]])
    got = oracle.select(mesh, T, hypot="scaled")
    P = mesh.sources[:, mesh.physics_topology]           # 3 x 3 x E
    D = T[:, :, None, None] - P[:, None, :, :]            # 3 x M x 3 x E
    d = np.sqrt((D[0] * D[0] + D[1] * D[1]) + D[2] * D[2])  # M x 3 x E
    best = np.full(d[:, 0].shape, np.inf)
    idx = np.zeros(best.shape, dtype=np.int8)
    for s_ in range(3):                                  # first strict minimum
        better = d[:, s_] < best
        best = np.where(better, d[:, s_], best)
        idx = np.where(better, s_ + 1, idx)
    assert np.array_equal(got, idx)
    assert not np.array_equal(got, oracle.select(mesh, T))   # ties do fall differently
