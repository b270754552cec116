"""Discontinuous triangular physics on the GPU (BIE_DTRI1 / BIE_DTRI3 / BIE_DTRI6 through the
C ABI) against the NumPy oracle (oracle_np, whose discontinuous bases and sources are
pinned in test_discontinuous_cpu.py).

The reference assembles these meshes with the same collocation code as continuous physics
(assemble_parallel!, src/3d/assembly_collocation.jl:270-479): the rule of a pair is chosen
by the nearest of the element's own physics nodes (for quadratic physics on linear
geometry nodes 4-6 fall back to rule 3), and each element adds into its own columns.
Bar: F, G, C within 1e-12 relative Frobenius error of the oracle.
"""
import numpy as np
import pytest

import oracle_np
from bie_amd import _lib
from bie_amd import mesh as M
from bie_amd.assembly import Batch, DeviceMesh, assemble_parallel
from gpu_helpers import DevBuf, check_F, own_blocks, rel, sync

pytestmark = pytest.mark.gpu
TOL = 1e-12
ORDERS = ["disctriconstant", "disctrilinear", "disctriquadratic"]


def disc_mesh(gorder, physics, f=3, beta_type="legendre"):
    base = M.perturb_radial(M.icosphere(f, gorder), 0.02, 7)
    return M.with_discontinuous_physics(base, physics, beta_type)


@pytest.mark.parametrize("gorder", ["linear", "quadratic"])
@pytest.mark.parametrize("physics", ORDERS)
@pytest.mark.parametrize("k", [0.0, 3.0, 2.0 + 0.5j])
def test_collocation_matches_oracle(gorder, physics, k):
    # (On linear geometry at (n, m) = (3, 3) the Legendre node offset of quadratic physics,
    # beta = (1 - sqrt(3/5)) / 2, is itself a 3-point Gauss node on [0, 1]: the edge nodes,
    # which fall back to rule 3, then coincide with one of its points and r = 0 there - the
    # reference's 1/r gives NaN.  Linear geometry is checked at (3, 2).)
    n, m = (3, 2) if gorder == "linear" else (3, 3)
    mesh = disc_mesh(gorder, physics)
    print(f"[case] disc collocation {gorder} {physics} k={k} N={mesh.n_sources}", flush=True)
    F, G, C = assemble_parallel(mesh, k, n=n, m=m)
    Fo, Go, Co = oracle_np.assemble(mesh, k, n=n, m=m)
    assert rel(G, Go) <= TOL and rel(C, Co) <= TOL
    check_F(F, Fo, mesh, "collocation")


@pytest.mark.parametrize("gorder", ["linear", "quadratic"])
@pytest.mark.parametrize("physics", ORDERS)
def test_field_points_match_oracle(gorder, physics):
    mesh = disc_mesh(gorder, physics, f=2)
    T = np.concatenate([M.fibonacci_sphere(70, 1.7), M.fibonacci_sphere(30, 0.5)], axis=1)
    print(f"[case] disc field {gorder} {physics}", flush=True)
    F, G, C = assemble_parallel(mesh, 4.0, T, n=2, m=2)
    Fo, Go, Co = oracle_np.assemble(mesh, 4.0, T, n=2, m=2)
    assert rel(F, Fo) <= TOL and rel(G, Go) <= TOL
    # off the surface C is a cancelling sum (near 0 or 1 in magnitude): compare against
    # the larger of its size and 1
    assert np.max(np.abs(C - Co)) <= TOL * max(np.max(np.abs(Co)), 1.0)


@pytest.mark.parametrize("physics", ORDERS)
@pytest.mark.parametrize("nm", [(2, 2), (3, 2), (4, 4)])
def test_device_fast_instances(physics, nm):
    """Real k below the fast-sincos bound takes the unrolled instances (rules of 4, 6 and 16
    points); F-only launches take the G-less ones."""
    n, m = nm
    mesh = disc_mesh("quadratic", physics, f=3)
    N = mesh.n_sources
    print(f"[case] disc device {physics} nm={nm} N={N}", flush=True)
    dm = DeviceMesh(mesh)
    F, G, C = DevBuf(16 * N * N), DevBuf(16 * N * N), DevBuf(16 * N)
    Fo, Go, Co = oracle_np.assemble(mesh, 6.0, n=n, m=m)
    for gOn in (True, False):
        dm.assemble(6.0, F.ptr, G.ptr, C.ptr, n=n, m=m, gOn=gOn)
        sync()
        f = F.download(np.zeros((N, N), dtype=np.complex128, order="F"))
        c = C.download(np.zeros(N, dtype=np.complex128))
        check_F(f, Fo, mesh, f"gOn={gOn}")
        assert rel(c, Co) <= TOL
        if gOn:
            g = G.download(np.zeros((N, N), dtype=np.complex128, order="F"))
            assert rel(g, Go) <= TOL
    dm.close()


def test_beta_is_part_of_the_cached_mesh():
    """The host drop-in keeps the last mesh on the device: a mesh that differs only in beta
    (same arrays otherwise) must not hit that cache."""
    a = disc_mesh("linear", "disctrilinear", f=2)
    b = M.Mesh3d(a.coordinates, a.topology, a.sources, a.physics_topology, a.geometry_kind,
                 a.physics_kind, 0.25)
    for mesh in (a, b, a):
        F, G, _ = assemble_parallel(mesh, 2.0, n=2, m=2)
        Fo, Go, _ = oracle_np.assemble(mesh, 2.0, n=2, m=2)
        check_F(F, Fo, mesh, f"beta={mesh.physics_beta}")
        assert rel(G, Go) <= TOL


@pytest.mark.parametrize("physics", ["disctriconstant", "disctrilinear"])
def test_batch_matches_single(physics):
    meshes = [disc_mesh("linear", physics, f=f) for f in (2, 3, 4)]
    ks = [1.5, 4.0, 2.5]
    dms = [DeviceMesh(x) for x in meshes]
    print(f"[case] disc batch {physics}", flush=True)
    bufs = [(DevBuf(16 * x.n_sources ** 2), DevBuf(16 * x.n_sources ** 2),
             DevBuf(16 * x.n_sources)) for x in meshes]
    b = Batch(dms, n=3, m=3)
    b.assemble(ks, [f.ptr for f, _, _ in bufs], [g.ptr for _, g, _ in bufs],
               [c.ptr for _, _, c in bufs])
    sync()
    for x, k, (Fb, Gb, Cb) in zip(meshes, ks, bufs):
        N = x.n_sources
        Fo, Go, Co = oracle_np.assemble(x, k, n=3, m=3)
        check_F(Fb.download(np.zeros((N, N), dtype=np.complex128, order="F")), Fo, x, "batch")
        assert rel(Gb.download(np.zeros((N, N), dtype=np.complex128, order="F")), Go) <= TOL
        assert rel(Cb.download(np.zeros(N, dtype=np.complex128)), Co) <= TOL
    b.close()


@pytest.mark.parametrize("gorder,physics", [("linear", "disctriconstant"),
                                            ("quadratic", "disctriquadratic"),
                                            ("linear", "disctriquadratic")])
def test_sweep_matches_oracle(gorder, physics):
    mesh = disc_mesh(gorder, physics, f=2)
    N = mesh.n_sources
    ks = np.linspace(1.0, 9.0, 5)
    print(f"[case] disc sweep {gorder} {physics}", flush=True)
    dm = DeviceMesh(mesh)
    kb = DevBuf(8 * len(ks))
    kb.upload(ks)
    stride = N * N
    F, G, C = DevBuf(16 * stride * len(ks)), DevBuf(16 * stride * len(ks)), DevBuf(16 * N)
    dm.assemble_sweep(kb.ptr, len(ks), F.ptr, G.ptr, stride, C.ptr, n=3, m=2)
    sync()
    f = F.download(np.zeros(stride * len(ks), dtype=np.complex128))
    g = G.download(np.zeros(stride * len(ks), dtype=np.complex128))
    for q, k in enumerate(ks):
        Fo, Go, _ = oracle_np.assemble(mesh, k, n=3, m=2)
        check_F(f[q * stride:(q + 1) * stride].reshape(N, N, order="F"), Fo, mesh, f"k={k}")
        assert rel(g[q * stride:(q + 1) * stride].reshape(N, N, order="F"), Go) <= TOL
    dm.close()


def test_adjoint_rejects_discontinuous_physics():
    from bie_amd.adjoint import assembly_vjp
    mesh = disc_mesh("linear", "disctrilinear", f=2)
    N = mesh.n_sources
    rng = np.random.default_rng(1)
    a = rng.standard_normal(N) + 0j
    dm = DeviceMesh(mesh)
    with pytest.raises(Exception) as ei:
        assembly_vjp(dm, 2.0, a, a)
    assert "continuous physics" in str(ei.value)
    dm.close()


def test_taylor_rosebem_settings():
    """The reference's ROSEBEM example (examples/3d_rosebem.jl: quadratic geometry with
    disctriquadratic physics, (n, m) = (2, 2), gOn = false, M = 25): order 0 against the
    oracle's assembly, order 1 against its central difference in k (the rules do not depend
    on k), and the 26-term expansion at a nearby k against the oracle's assembly there."""
    from bie_amd.taylor import apply_taylor_expansion, taylor_assemble
    mesh = disc_mesh("quadratic", "disctriquadratic", f=2)
    k0, Mo = 2.0, 25
    print(f"[case] disc taylor N={mesh.n_sources} orders 0..{Mo}", flush=True)
    F, _, _ = taylor_assemble(mesh, k0, M=Mo, n=2, m=2, gOn=False)
    F0, _, _ = oracle_np.assemble(mesh, k0, n=2, m=2, gOn=False)
    check_F(F[:, :, 0], F0, mesh, "order 0")
    h = 1e-4
    Fp, _, _ = oracle_np.assemble(mesh, k0 + h, n=2, m=2, gOn=False)
    Fm, _, _ = oracle_np.assemble(mesh, k0 - h, n=2, m=2, gOn=False)
    fd = (Fp - Fm) / (2 * h)
    assert rel(F[:, :, 1], fd) <= 1e-6
    k1 = k0 + 0.25
    Fx = apply_taylor_expansion(F, k1, k0)
    F1, _, _ = oracle_np.assemble(mesh, k1, n=2, m=2, gOn=False)
    assert rel(Fx, F1) <= 1e-10


@pytest.mark.parametrize("go,po,tol", [
    ("linear", "disctriconstant", 0.0073), ("linear", "disctrilinear", 0.0044),
    ("linear", "disctriquadratic", 0.004), ("quadratic", "disctriconstant", 0.0008),
    ("quadratic", "disctrilinear", 4.2e-4), ("quadratic", "disctriquadratic", 4.9e-5)])
def test_rigid_sphere_reference_tolerance_table(go, po, tol):
    """test/3d/assembly_test.jl:12-44: plane-wave scattering by the rigid unit sphere on the
    reference's sphere_1m mesh with the default n = m = 4, the BEM solution against the
    analytic one, to the reference's own tolerance for each geometry and physics order."""
    import os

    import analytic
    gold = os.path.join(os.path.dirname(__file__), "golden")
    k = 2 * np.pi * 100.0 / 340.0
    mesh = M.load_comsol(os.path.join(gold, "sphere_1m"), go, po)
    print(f"[case] rigid sphere {go} {po} N={mesh.n_sources}", flush=True)
    F, _, C = assemble_parallel(mesh, k, gOn=False)
    pI = analytic.incoming_wave([np.pi / 2, 0.0], 1.0, mesh.sources, k)
    p = np.linalg.solve(F + np.diag(1.0 - C), pI)
    th = np.arccos(np.clip(mesh.sources[0], -1, 1))
    pa = analytic.plane_wave_scattering_sphere(k, 1.0, 1.0, th, 1e-6)
    err = rel(p, pa)
    print(f"  relative error {err:.3e} (reference tolerance {tol})", flush=True)
    assert err < tol
