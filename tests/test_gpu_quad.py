"""Quadrilateral elements on the GPU (BIE_QUAD4 / QUAD4S / QUAD8 / QUAD9 geometry with
continuous or BIE_DQUAD* physics, through the C ABI) against the NumPy oracle (oracle_np,
pinned in test_quad_cpu.py by the reference's own rigid-sphere test).

The reference's generic method (assemble_parallel!(mesh, k, sources, ::SurfaceFunction),
src/3d/assembly_collocation.jl:481-538) integrates every (target, element) pair with one
tensor Gauss-Legendre rule, m points in u and n in v.  Bar: G and C within 1e-12 relative
Frobenius error; F too, except on an element's own nodes, where the plain rule leaves the
rounding noise of n.d amplified by 1/r^3 (gpu_helpers.check_F).
"""
import os

import numpy as np
import pytest

import oracle_np
from bie_amd import _lib
from bie_amd import mesh as M
from bie_amd.assembly import Batch, DeviceMesh, assemble_parallel
from gpu_helpers import TOL, DevBuf, check_F, rel, sync

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def quad8_from_quad9(mesh9):
    """The 8-node serendipity element on the nodes of a 9-node one (centre dropped): corners
    counter-clockwise, then the edge mid points (0,-1), (1,0), (0,1), (-1,0)."""
    t = mesh9.topology[[0, 1, 3, 2, 4, 7, 8, 5], :]
    used = np.unique(t)
    remap = -np.ones(mesh9.n_coordinates, dtype=np.int64)
    remap[used] = np.arange(used.size)
    X = mesh9.coordinates[:, used]
    return M.Mesh3d(X, remap[t], X.copy(), remap[t], M.QUAD8, M.QUAD8)


def quad4s_from_quad4(mesh4):
    t = mesh4.topology[[0, 1, 3, 2], :]
    return M.Mesh3d(mesh4.coordinates, t, mesh4.sources, t.copy(), M.QUAD4S, M.QUAD4S)


def make(name):
    base4 = M.perturb_radial(M.cubed_sphere(4, "linear"), 0.02, 9)
    base9 = M.perturb_radial(M.cubed_sphere(3, "quadratic"), 0.02, 9)
    return {
        "quad4": lambda: base4,
        "quad4s": lambda: quad4s_from_quad4(base4),
        "quad9": lambda: base9,
        "quad8": lambda: quad8_from_quad9(base9),
        "quad9_linphys": lambda: M.cubed_sphere(3, "quadratic", physics="linear"),
        "quad4_dconst": lambda: M.with_discontinuous_physics(base4, "discquadconstant"),
        "quad4_dlin": lambda: M.with_discontinuous_physics(base4, "discquadlinear"),
        "quad9_dquad": lambda: M.with_discontinuous_physics(base9, "discquadquadratic"),
        "quad9_dlin": lambda: M.with_discontinuous_physics(base9, "discquadlinear"),
    }[name]()


NAMES = ["quad4", "quad4s", "quad9", "quad8", "quad9_linphys", "quad4_dconst", "quad4_dlin",
         "quad9_dquad", "quad9_dlin"]


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("k", [0.0, 3.0, 2.0 + 0.5j])
def test_collocation_matches_oracle(name, k):
    mesh = make(name)
    print(f"[case] quad collocation {name} k={k} N={mesh.n_sources}", flush=True)
    F, G, C = assemble_parallel(mesh, k)           # the generic method's n = m = 4
    Fo, Go, Co = oracle_np.assemble(mesh, k, n=4, m=4)
    assert rel(G, Go) <= TOL and rel(C, Co) <= TOL
    check_F(F, Fo, mesh, name)


@pytest.mark.parametrize("name", ["quad4", "quad9", "quad8", "quad9_dquad"])
@pytest.mark.parametrize("nm", [(2, 2), (3, 2), (2, 3), (4, 4), (5, 4)])
def test_rule_orders_and_instances(name, nm):
    """(2,2), (3,2), (4,4): the unrolled real-k instances (rules of 4, 6, 16 points) with
    and without G; (2,3) the generic staged one; (5,4): 20 points, not staged."""
    n, m = nm
    mesh = make(name)
    N = mesh.n_sources
    print(f"[case] quad device {name} nm={nm} N={N}", flush=True)
    dm = DeviceMesh(mesh)
    F, G, C = DevBuf(16 * N * N), DevBuf(16 * N * N), DevBuf(16 * N)
    Fo, Go, Co = oracle_np.assemble(mesh, 5.0, n=n, m=m)
    for gOn in (True, False):
        dm.assemble(5.0, F.ptr, G.ptr, C.ptr, n=n, m=m, gOn=gOn)
        sync()
        f = F.download(np.zeros((N, N), dtype=np.complex128, order="F"))
        check_F(f, Fo, mesh, f"{name} gOn={gOn}")
        assert rel(C.download(np.zeros(N, dtype=np.complex128)), Co) <= TOL
        if gOn:
            assert rel(G.download(np.zeros((N, N), dtype=np.complex128, order="F")), Go) <= TOL
    dm.close()


@pytest.mark.parametrize("name", ["quad4", "quad9", "quad9_dquad"])
def test_field_points_match_oracle(name):
    mesh = make(name)
    T = np.concatenate([M.fibonacci_sphere(60, 1.6), M.fibonacci_sphere(25, 0.4)], axis=1)
    print(f"[case] quad field {name}", flush=True)
    F, G, C = assemble_parallel(mesh, 3.5, T, n=4, m=4)
    Fo, Go, Co = oracle_np.assemble(mesh, 3.5, T, n=4, m=4)
    assert rel(F, Fo) <= TOL and rel(G, Go) <= TOL
    assert np.max(np.abs(C - Co)) <= TOL * max(np.max(np.abs(Co)), 1.0)


def test_gauss_node_on_singularity_is_refused():
    mesh = make("quad9")
    with pytest.raises(ValueError, match="Singularity"):
        assemble_parallel(mesh, 1.0, n=3, m=3)
    F, _, _ = assemble_parallel(mesh, 1.0, n=3, m=2)
    assert np.all(np.isfinite(F))


def test_other_entry_points_refuse_quadrilaterals():
    mesh = make("quad4")
    dm = DeviceMesh(mesh)
    N = mesh.n_sources
    kb = DevBuf(16)
    kb.upload(np.array([1.0, 2.0]))
    F, G, C = DevBuf(32 * N * N), DevBuf(32 * N * N), DevBuf(16 * N)
    with pytest.raises(Exception) as ei:
        dm.assemble_sweep(kb.ptr, 2, F.ptr, G.ptr, N * N, C.ptr)
    assert "triangular" in str(ei.value)
    with pytest.raises(Exception):
        Batch([dm], n=2, m=2)
    dm.close()


@pytest.mark.parametrize("go,po,tol", [
    ("linear", "linear", 0.012), ("linear", "discquadconstant", 0.014),
    ("linear", "discquadlinear", 0.009), ("linear", "discquadquadratic", 0.008),
    ("quadratic", "linear", 0.004), ("quadratic", "geometry", 2.8e-5),
    ("quadratic", "discquadconstant", 0.0019), ("quadratic", "discquadlinear", 0.003),
    ("quadratic", "discquadquadratic", 5.9e-5)])
def test_rigid_sphere_reference_tolerance_table(go, po, tol):
    """test/3d/assembly_test.jl:46-78: plane-wave scattering by the rigid unit sphere on the
    reference's quad_sphere mesh (n = m = 4, gOn = false), BEM against the analytic solution
    within the reference's tolerance for each geometry and physics order."""
    import analytic
    k = 2 * np.pi * 100.0 / 340.0
    mesh = M.load_comsol_quad(os.path.join(GOLD, "quad_sphere"), go, po)
    print(f"[case] quad rigid sphere {go} {po} N={mesh.n_sources}", flush=True)
    F, _, C = assemble_parallel(mesh, k, gOn=False)
    pI = analytic.incoming_wave([np.pi / 2, 0.0], 1.0, mesh.sources, k)
    p = np.linalg.solve(F + np.diag(1.0 - C), pI)
    th = np.arccos(np.clip(mesh.sources[0], -1, 1))
    pa = analytic.plane_wave_scattering_sphere(k, 1.0, 1.0, th, 1e-6)
    err = rel(p, pa)
    print(f"  relative error {err:.3e} (reference tolerance {tol})", flush=True)
    assert err < tol
