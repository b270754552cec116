"""Quadrilateral elements (the reference's generic collocation path,
assemble_parallel!(mesh, k, sources, ::SurfaceFunction), src/3d/assembly_collocation.jl:481-538):
the NumPy oracle's rule, bases and assembly pinned by what the mathematics and the reference's
own tests fix, the mesh builders, and the host-side checks of the C ABI.

Pins:
* the rule: m Gauss-Legendre points in u and n in v (create_shape_function passes (m, n) to
  quadrilateralQuadpoints(n, m)); weights sum to the area 4; exact for u^a v^b up to
  a = 2m - 1, b = 2n - 1;
* every basis is the identity at its own nodes (get_nodal_nodes_u/_v,
  SurfaceFunction.jl:439-470, restated separately in bie_amd.mesh), sums to one, and its
  derivatives match central differences;
* geometry: a flat square's area exact, the sphere's area converging to 4 pi;
* the system: the reference's rigid-sphere test (test/3d/assembly_test.jl:46-78) on its own
  quad_sphere mesh, BEM against the analytic solution within the reference's tolerance for
  each geometry and physics order.
"""
import ctypes
import os

import numpy as np
import pytest

import analytic
import oracle_np
from bie_amd import _lib
from bie_amd import mesh as M

GEOM = [M.QUAD4, M.QUAD4S, M.QUAD8, M.QUAD9]
# get_nodal_nodes_u/_v for the continuous quadrilaterals
NODES = {
    M.QUAD4: ([-1, 1, -1, 1], [-1, -1, 1, 1]),
    M.QUAD4S: ([-1, 1, 1, -1], [-1, -1, 1, 1]),
    M.QUAD8: ([-1, 1, 1, -1, 0, 1, 0, -1], [-1, -1, 1, 1, -1, 0, 1, 0]),
    M.QUAD9: ([-1, 1, -1, 1, 0, -1, 0, 1, 0], [-1, -1, 1, 1, -1, 0, 0, 0, 1]),
}


@pytest.mark.parametrize("n,m", [(1, 1), (2, 3), (4, 4), (5, 2)])
def test_rule(n, m):
    u, v, w = oracle_np.quad_rule(n, m)
    assert len(np.unique(u)) == m and len(np.unique(v)) == n
    assert abs(w.sum() - 4.0) < 1e-14
    for a in range(2 * m):
        for b in range(2 * n):
            exact = (2 / (a + 1) if a % 2 == 0 else 0) * (2 / (b + 1) if b % 2 == 0 else 0)
            assert abs(np.sum(w * u ** a * v ** b) - exact) < 1e-13


@pytest.mark.parametrize("kind", GEOM)
def test_continuous_bases(kind):
    un, vn = NODES[kind]
    N, _, _ = oracle_np.basis(kind, un, vn)
    assert np.allclose(N, np.eye(len(un)), atol=1e-15)
    # the product mesh builder's own basis agrees
    assert np.allclose(M.geometry_basis(kind, un, vn), np.eye(len(un)), atol=1e-15)
    rng = np.random.default_rng(3)
    u, v = rng.uniform(-1, 1, 40), rng.uniform(-1, 1, 40)
    N, du, dv = oracle_np.basis(kind, u, v)
    assert np.allclose(N.sum(axis=1), 1.0, atol=1e-14)
    assert np.allclose(M.geometry_basis(kind, u, v).T, N, atol=1e-15)
    h = 1e-6
    for d, (a, b) in ((du, (h, 0)), (dv, (0, h))):
        Np, _, _ = oracle_np.basis(kind, u + a, v + b)
        Nm, _, _ = oracle_np.basis(kind, u - a, v - b)
        assert np.allclose((Np - Nm) / (2 * h), d, atol=1e-8)
    # bilinear / quadratic reproduction of u and v
    assert np.allclose(N @ np.asarray(un, float), u, atol=1e-14)
    assert np.allclose(N @ np.asarray(vn, float), v, atol=1e-14)


@pytest.mark.parametrize("kind", [M.DQUAD1, M.DQUAD4, M.DQUAD9])
@pytest.mark.parametrize("beta", [0.2254, 0.4226])
def test_discontinuous_bases(kind, beta):
    un, vn = M.discontinuous_nodes(kind, beta)
    N, du, _ = oracle_np.basis(kind, un, vn, beta)
    assert np.allclose(N, np.eye(len(un)), atol=1e-14)
    rng = np.random.default_rng(4)
    u, v = rng.uniform(-1, 1, 30), rng.uniform(-1, 1, 30)
    N, du, _ = oracle_np.basis(kind, u, v, beta)
    assert np.allclose(N.sum(axis=1), 1.0, atol=1e-14)
    h = 1e-6
    Np, _, _ = oracle_np.basis(kind, u + h, v, beta)
    Nm, _, _ = oracle_np.basis(kind, u - h, v, beta)
    assert np.allclose((Np - Nm) / (2 * h), du, atol=1e-8)


def test_beta_of_quadrilaterals():
    # get_beta_quad_linear / _quadratic (mesh.jl:193-214): 1 + the smallest Gauss node
    bl = M.discontinuous_beta("discquadlinear")
    bq = M.discontinuous_beta("discquadquadratic")
    assert abs(bl - (1 - 1 / np.sqrt(3))) < 1e-15 and abs(bq - (1 - np.sqrt(0.6))) < 1e-15
    assert M.discontinuous_beta("discquadlinear", "equidistant") == 0.5
    assert M.discontinuous_beta("discquadlinear", 0.3) == 0.3   # not halved on quads


def test_geometry_area():
    # flat unit square [0, 1]^2 as one QUAD4 element
    X4 = np.array([[0, 1, 0, 1], [0, 0, 1, 1], [0, 0, 0, 0]], dtype=float)
    sq = M.Mesh3d(X4, np.arange(4)[:, None], X4.copy(), np.arange(4)[:, None], M.QUAD4,
                  M.QUAD4)
    _, nrm, jw = oracle_np.element_geometry(sq, 1, 3, 2)
    assert abs(jw.sum() - 1.0) < 1e-15
    assert np.allclose(nrm[0], [0, 0, -1])        # the reference's normal: -(t_u x t_v)
    areas = []
    for f in (2, 4, 8):
        mesh = M.cubed_sphere(f, "quadratic")
        _, _, jw = oracle_np.element_geometry(mesh, 1, 4, 4)
        areas.append(abs(jw.sum() - 4 * np.pi))
    # quadratic elements: the area error falls as h^4 (x16 per halving)
    assert areas[0] / areas[1] > 10 and areas[1] / areas[2] > 10 and areas[2] < 5e-4


def test_selection_is_rule_one():
    mesh = M.cubed_sphere(2, "quadratic")
    T = np.concatenate([mesh.sources, M.fibonacci_sphere(7, 1.3)], axis=1)
    assert np.all(oracle_np.select(mesh, T) == 1)


def test_gauss_node_on_singularity():
    mesh = M.cubed_sphere(1, "quadratic")
    with pytest.raises(ValueError, match="Singularity"):
        oracle_np.assemble(mesh, 1.0, n=3, m=3)
    oracle_np.assemble(mesh, 1.0, n=3, m=2)   # one even order is fine


def test_loader_sizes():
    # meshing_test.jl:52-87 on the reference's quad_sphere: 315 elements
    f = os.path.join(M.golden_dir(), "quad_sphere")
    for go in ("linear", "quadratic"):
        for po in ("linear", "geometry", "discquadconstant", "discquadlinear",
                   "discquadquadratic"):
            if go == "linear" and po == "geometry":
                continue
            mesh = M.load_comsol_quad(f, go, po)
            assert mesh.topology.shape == (4 if go == "linear" else 9, 315)
            assert mesh.sources.shape == (3, len(np.unique(mesh.physics_topology)))
            assert mesh.topology.shape[1] == mesh.physics_topology.shape[1]


@pytest.mark.parametrize("go,po,tol", [("linear", "linear", 0.012),
                                       ("quadratic", "geometry", 2.8e-5),
                                       ("linear", "discquadlinear", 0.009),
                                       ("quadratic", "discquadquadratic", 5.9e-5)])
def test_rigid_sphere_reference_tolerance_table(go, po, tol):
    """test/3d/assembly_test.jl:46-78 with the oracle (n = m = 4 defaults, gOn = false)."""
    k = 2 * np.pi * 100.0 / 340.0
    mesh = M.load_comsol_quad(os.path.join(M.golden_dir(), "quad_sphere"), go, po)
    N = mesh.n_sources
    F = np.zeros((N, N), dtype=np.complex128)
    C = np.zeros(N, dtype=np.complex128)
    for r0 in range(0, N, 500):
        T = mesh.sources[:, r0:r0 + 500]
        F[r0:r0 + T.shape[1]], _, C[r0:r0 + T.shape[1]] = oracle_np.assemble(mesh, k, T,
                                                                             gOn=False)
    pI = analytic.incoming_wave([np.pi / 2, 0.0], 1.0, mesh.sources, k)
    p = np.linalg.solve(F + np.diag(1.0 - C), pI)
    th = np.arccos(np.clip(mesh.sources[0], -1, 1))
    pa = analytic.plane_wave_scattering_sphere(k, 1.0, 1.0, th, 1e-6)
    assert np.linalg.norm(p - pa) / np.linalg.norm(pa) < tol


def test_library_rejects_bad_quadrilateral_meshes():
    L = _lib.lib()
    mesh = M.cubed_sphere(2, "quadratic")
    hm = _lib.HostMesh(mesh)
    nnz = ctypes.c_int64()
    # the near field of quadrilaterals is not covered
    assert L.bie_sparse_count(ctypes.byref(hm.struct), 1, ctypes.byref(nnz)) == \
        _lib.BIE_EUNSUPPORTED
    # triangle physics on quadrilateral geometry
    hm.struct.physics_kind = M.TRI3
    assert L.bie_sparse_count(ctypes.byref(hm.struct), 1, ctypes.byref(nnz)) == \
        _lib.BIE_EUNSUPPORTED
    # beta = 1 collapses the discontinuous nodes
    hm.struct.physics_kind = M.DQUAD4
    hm.struct.physics_beta = 1.0
    assert L.bie_sparse_count(ctypes.byref(hm.struct), 1, ctypes.byref(nnz)) == _lib.BIE_EINVAL
