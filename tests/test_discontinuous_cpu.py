"""Discontinuous triangular physics (DiscontinuousTriangularConstant / Linear / Quadratic,
src/3d/surface_functions.jl:234-318, nodes src/3d/SurfaceFunction.jl:429-438, sources
compute_sources src/3d/mesh.jl:116-148, beta get_beta_tri_* src/3d/mesh.jl:182-214):
the NumPy oracle's bases and the mesh builder pinned by what the mathematics fixes, and the
host-side checks of the C ABI.

Pins (each chosen so that a plausible slip fails one):
* every basis is the identity at its own nodes (node order vs basis order, edge-node pairs);
* partition of unity and reproduction of u, v (and u^2, uv, v^2 for the quadratic one):
  a wrong beta scaling or sign fails it;
* beta of "legendre" is a root of the Legendre polynomial (P2 / P3) mapped to the
  triangle, "equidistant" the closed forms, a number is halved;
* assembly: with the rule of every pair frozen, discontinuous physics reproduces the
  continuous assembly on any density the discontinuous basis represents exactly
  (constants; fields linear in x on linear geometry, quadratic ones for the quadratic
  basis): G_disc p(sources_disc) = G_cont p(sources_cont), the same for F.  A wrong node
  position (compute_sources), column order or basis fails it.
* rule selection over the element's own nodes: brute force (nearest node, first minimum,
  indices beyond the rule count map to the last rule, assembly_collocation.jl:316-344).
"""
import ctypes

import numpy as np
import pytest

import oracle
import oracle_np
from bie_amd import _lib
from bie_amd import mesh as M

KINDS = [M.DTRI1, M.DTRI3, M.DTRI6]
ORDERS = {M.DTRI1: "disctriconstant", M.DTRI3: "disctrilinear", M.DTRI6: "disctriquadratic"}


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("beta", [0.1127, 0.2113, 0.25])
def test_basis_is_identity_at_nodes(kind, beta):
    u, v = M.discontinuous_nodes(kind, beta)
    N, _, _ = oracle_np.basis(kind, u, v, beta)
    assert np.allclose(N, np.eye(M.physics_nodes(kind)), atol=1e-14)


@pytest.mark.parametrize("kind", [M.DTRI3, M.DTRI6])
def test_basis_reproduces_polynomials(kind):
    rng = np.random.default_rng(5)
    beta = 0.17
    u = rng.random(50) * 0.9
    v = rng.random(50) * (1 - u)
    N, du, dv = oracle_np.basis(kind, u, v, beta)
    un, vn = M.discontinuous_nodes(kind, beta)
    assert np.allclose(N.sum(axis=1), 1.0, atol=1e-14)
    assert np.allclose(N @ un, u, atol=1e-14) and np.allclose(N @ vn, v, atol=1e-14)
    assert np.allclose(du.sum(axis=1), 0.0, atol=1e-12)
    assert np.allclose(du @ un, 1.0, atol=1e-12) and np.allclose(dv @ vn, 1.0, atol=1e-12)
    if kind == M.DTRI6:
        for f in (lambda a, b: a * a, lambda a, b: a * b, lambda a, b: b * b):
            assert np.allclose(N @ f(un, vn), f(u, v), atol=1e-13)
    # derivatives against central differences
    h = 1e-6
    Np, _, _ = oracle_np.basis(kind, u + h, v, beta)
    Nm, _, _ = oracle_np.basis(kind, u - h, v, beta)
    assert np.allclose((Np - Nm) / (2 * h), du, atol=1e-7)


def test_beta_values():
    def p2(x):
        return 0.5 * (3 * x * x - 1)

    def p3(x):
        return 0.5 * (5 * x ** 3 - 3 * x)

    bl = M.discontinuous_beta("disctrilinear")
    bq = M.discontinuous_beta("disctriquadratic")
    # half of 1 + x1, x1 the smallest Gauss-Legendre node on [-1, 1]
    assert abs(p2(2 * bl - 1)) < 1e-15 and 2 * bl - 1 < 0
    assert abs(p3(2 * bq - 1)) < 1e-15 and 2 * bq - 1 < -0.5
    assert M.discontinuous_beta("disctrilinear", "equidistant") == 0.25
    assert M.discontinuous_beta("disctriquadratic", "equidistant") == 1 / 6
    assert M.discontinuous_beta("disctrilinear", 0.3) == 0.15
    assert M.discontinuous_beta("disctriconstant") == 0.0


@pytest.mark.parametrize("order", ["linear", "quadratic"])
@pytest.mark.parametrize("kind", KINDS)
def test_sources_are_geometry_at_nodes(order, kind):
    base = M.icosphere(2, order)
    mesh = M.with_discontinuous_physics(base, ORDERS[kind])
    npk = M.physics_nodes(kind)
    assert mesh.n_sources == npk * base.n_elements
    assert np.array_equal(mesh.physics_topology.ravel(order="F"), np.arange(mesh.n_sources))
    u, v = M.discontinuous_nodes(kind, mesh.physics_beta)
    N, _, _ = oracle_np.basis(base.geometry_kind, u, v)
    for e in (0, 7, base.n_elements - 1):
        X = base.coordinates[:, base.topology[:, e]]
        assert np.allclose(mesh.sources[:, mesh.physics_topology[:, e]], X @ N.T, atol=1e-15)
    # interior nodes of a sphere mesh lie inside the unit ball, off the corners
    r = np.linalg.norm(mesh.sources, axis=0)
    assert np.all(r <= 1 + 1e-12)


def _targets():
    a = M.fibonacci_sphere(40, 1.6)
    b = M.fibonacci_sphere(20, 0.45)
    return np.concatenate([a, b], axis=1)


@pytest.mark.parametrize("gorder,kind,field", [
    ("linear", M.DTRI1, "const"), ("linear", M.DTRI3, "linear"), ("linear", M.DTRI6, "linear"),
    ("quadratic", M.DTRI1, "const"), ("quadratic", M.DTRI3, "const"),
    ("quadratic", M.DTRI6, "linear")])
def test_assembly_reproduces_continuous(gorder, kind, field):
    cont = M.perturb_radial(M.icosphere(2, gorder), 0.03, 4)
    disc = M.with_discontinuous_physics(cont, ORDERS[kind])
    T = _targets()
    rules = oracle_np.select(cont, T)
    k = 2.5
    Fc, Gc, Cc = oracle_np.assemble(cont, k, T, n=3, m=3, rules=rules)
    Fd, Gd, Cd = oracle_np.assemble(disc, k, T, n=3, m=3, rules=rules)
    a, b = 0.7, np.array([0.3, -1.1, 0.5])

    def p(X):
        return a + (b @ X if field == "linear" else np.zeros(X.shape[1]))

    for Xc, Xd in ((Gc, Gd), (Fc, Fd)):
        ref = Xc @ p(cont.sources)
        got = Xd @ p(disc.sources)
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert np.allclose(Cd, Cc, rtol=0, atol=1e-14)  # C does not see the physics


def test_selection_brute_force():
    # quadratic discontinuous physics on linear geometry: nodes 4-6 fall back to rule 3
    mesh = M.with_discontinuous_physics(M.icosphere(1, "linear"), "disctriquadratic")
    T = np.concatenate([mesh.sources[:, ::7], M.fibonacci_sphere(9, 1.2)], axis=1)
    got = oracle_np.select(mesh, T)
    for i in range(T.shape[1]):
        for e in range(mesh.n_elements):
            P = mesh.sources[:, mesh.physics_topology[:, e]]
            # the distances as Julia's hypot rounds them (the C restatement): symmetric
            # meshes have near ties that another formula decides the other way
            d = [oracle.julia_hypot3(*(P[:, s] - T[:, i])) for s in range(6)]
            best = int(np.argmin(d)) + 1
            assert got[i, e] == min(best, 3)
    one = M.with_discontinuous_physics(M.icosphere(1, "quadratic"), "disctriconstant")
    assert np.all(oracle_np.select(one, T) == 1)


def test_library_rejects_bad_discontinuous_meshes():
    L = _lib.lib()
    mesh = M.with_discontinuous_physics(M.icosphere(1, "linear"), "disctrilinear")
    hm = _lib.HostMesh(mesh)
    nnz = ctypes.c_int64()
    # the near field of discontinuous physics: one row per element node, its own element
    assert L.bie_sparse_count(ctypes.byref(hm.struct), 1, ctypes.byref(nnz)) == _lib.BIE_OK
    assert nnz.value == 3 * mesh.n_sources
    # beta = 1/3 collapses the nodes
    hm.struct.physics_beta = 1.0 / 3.0
    assert L.bie_sparse_count(ctypes.byref(hm.struct), 1, ctypes.byref(nnz)) == _lib.BIE_EINVAL
    assert b"physics_beta" in L.bie_last_error()
    hm.struct.physics_beta = float("nan")
    assert L.bie_sparse_count(ctypes.byref(hm.struct), 1, ctypes.byref(nnz)) == _lib.BIE_EINVAL
    # an unknown kind
    hm.struct.physics_beta = mesh.physics_beta
    hm.struct.physics_kind = 0x102
    assert L.bie_sparse_count(ctypes.byref(hm.struct), 1, ctypes.byref(nnz)) == \
        _lib.BIE_EUNSUPPORTED
    assert L.bie_version() >= 3


def test_loaders_build_discontinuous_meshes():
    import os
    f = os.path.join(M.golden_dir(), "sphere_1m")   # the reference's test mesh (840 elements)
    for order in ("linear", "quadratic"):
        for phys, kind in (("disctriconstant", M.DTRI1), ("disctrilinear", M.DTRI3),
                           ("disctriquadratic", M.DTRI6)):
            mesh = M.load_comsol(f, order, phys)
            assert mesh.physics_kind == kind
            # meshing_test.jl:21-27: 840 elements, one source per physics node
            assert mesh.topology.shape == (3 if order == "linear" else 6, 840)
            assert mesh.sources.shape == (3, len(np.unique(mesh.physics_topology)))
            assert mesh.topology.shape[1] == mesh.physics_topology.shape[1]


@pytest.mark.parametrize("go,po,tol", [("linear", "disctrilinear", 0.0044),
                                       ("quadratic", "disctriquadratic", 4.9e-5)])
def test_rigid_sphere_reference_tolerance_table(go, po, tol):
    """test/3d/assembly_test.jl:12-44 for discontinuous physics (the oracle, assembled in
    row blocks): the rigid-sphere solution on the reference's sphere_1m at n = m = 4 within
    the reference's own tolerance."""
    import os

    import analytic
    k = 2 * np.pi * 100.0 / 340.0
    mesh = M.load_comsol(os.path.join(M.golden_dir(), "sphere_1m"), go, po)
    N = mesh.n_sources
    F = np.zeros((N, N), dtype=np.complex128)
    C = np.zeros(N, dtype=np.complex128)
    for r0 in range(0, N, 420):
        T = mesh.sources[:, r0:r0 + 420]
        F[r0:r0 + T.shape[1]], _, C[r0:r0 + T.shape[1]] = oracle_np.assemble(mesh, k, T,
                                                                             gOn=False)
    pI = analytic.incoming_wave([np.pi / 2, 0.0], 1.0, mesh.sources, k)
    p = np.linalg.solve(F + np.diag(1.0 - C), pI)
    th = np.arccos(np.clip(mesh.sources[0], -1, 1))
    pa = analytic.plane_wave_scattering_sphere(k, 1.0, 1.0, th, 1e-6)
    assert np.linalg.norm(p - pa) / np.linalg.norm(pa) < tol
