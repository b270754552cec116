"""Batched frequency sweep (one launch for many k) against single-k assembly and the
oracle; the graph-captured form replayed several times with eager work in between."""
import numpy as np
import pytest

import oracle
from bie_amd import mesh as M
from bie_amd.assembly import DeviceMesh, assemble_parallel
from gpu_helpers import DevBuf, sync

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def run_sweep(mesh, ks, n, m, graph_replays=0, gOn=True):
    dm = DeviceMesh(mesh)
    N = mesh.n_sources
    nk = len(ks)
    stride = N * N + 7          # padded stride: F_q need not be packed
    kb = DevBuf(8 * nk)
    kb.upload(np.asarray(ks, dtype=np.float64))
    F, G, C = DevBuf(16 * stride * nk), DevBuf(16 * stride * nk), DevBuf(16 * N)
    outs = []
    if graph_replays:
        g = dm.sweep_graph(kb.ptr, nk, F.ptr, G.ptr, stride, C.ptr, n=n, m=m)
        for rep in range(graph_replays):
            g.launch()
            sync()
            outs.append(read(F, G, C, N, nk, stride))
            # eager work between replays: a plain assembly on the same mesh and buffers
            dm.assemble(1.0, F.ptr, G.ptr, C.ptr, n=n, m=m)
            sync()
        g.close()
    else:
        dm.assemble_sweep(kb.ptr, nk, F.ptr, G.ptr, stride, C.ptr, n=n, m=m, gOn=gOn)
        sync()
        outs.append(read(F, G, C, N, nk, stride))
    return outs


def read(F, G, C, N, nk, stride):
    f = F.download(np.zeros(stride * nk, dtype=np.complex128))
    g = G.download(np.zeros(stride * nk, dtype=np.complex128))
    c = C.download(np.zeros(N, dtype=np.complex128))
    Fs = [f[q * stride:q * stride + N * N].reshape(N, N, order="F") for q in range(nk)]
    Gs = [g[q * stride:q * stride + N * N].reshape(N, N, order="F") for q in range(nk)]
    return Fs, Gs, c


@pytest.mark.parametrize("name,ks,n,m", [
    ("lin", np.linspace(0.5, 30.0, 9), 4, 4),                  # uniform blocks + partial block
    ("lin", [1.0, 2.5, 2.6, 7.0, 11.0, 11.5], 3, 2),            # non-uniform: direct path
    ("quad", np.linspace(1.0, 12.0, 6), 3, 2),                  # quadratic: mid-edge pairs
])
def test_sweep_matches_single_k(name, ks, n, m):
    mesh = M.icosphere(4, "linear") if name == "lin" else M.icosphere(3, "quadratic")
    print(f"[case] sweep {name} nk={len(ks)} (n,m)=({n},{m})", flush=True)
    (Fs, Gs, c), = run_sweep(mesh, ks, n, m)
    for q, k in enumerate(ks):
        F1, G1, C1 = assemble_parallel(mesh, k, mesh.sources, n=n, m=m)
        assert rel(Fs[q], F1) < TOL and rel(Gs[q], G1) < TOL, (q, k)
    assert rel(c, C1) < TOL
    Fo, Go, Co = oracle.assemble(mesh, ks[-1], n=n, m=m)
    assert rel(Fs[-1], Fo) < TOL and rel(Gs[-1], Go) < TOL


@pytest.mark.parametrize("name,ks", [
    ("lin", np.linspace(0.5, 30.0, 19)),           # two uniform blocks of 8 + a partial one
    ("lin", [1.0, 2.5, 2.6, 7.0, 11.0, 11.5, 3.0, 4.0, 9.0]),  # non-uniform: direct path
    ("quad", np.linspace(1.0, 12.0, 11)),          # quadratic: uniform block of 4 + partial
])
def test_sweep_without_g_matches_single_k(name, ks):
    """gOn = false: two wavenumbers' F per slot group, twice the wavenumbers per block, the
    uniform-spacing recurrence over the longer blocks; every F_k against a single-k F
    assembly."""
    mesh = M.icosphere(4, "linear") if name == "lin" else M.icosphere(3, "quadratic")
    print(f"[case] sweep without G {name} nk={len(ks)}", flush=True)
    (Fs, Gs, c), = run_sweep(mesh, ks, 3, 2, gOn=False)
    for q, k in enumerate(ks):
        F1, _, C1 = assemble_parallel(mesh, k, mesh.sources, n=3, m=2, gOn=False)
        assert rel(Fs[q], F1) < TOL, (q, k)
        assert not Gs[q].any()
    assert rel(c, C1) < TOL


@pytest.mark.parametrize("gOn", [True, False])
def test_sweep_of_one_wavenumber(gOn):
    """nk = 1: a single partial block, with and without G."""
    mesh = M.icosphere(3, "quadratic")
    print(f"[case] sweep of one k gOn={gOn}", flush=True)
    (Fs, Gs, c), = run_sweep(mesh, [4.5], 3, 2, gOn=gOn)
    F1, G1, C1 = assemble_parallel(mesh, 4.5, mesh.sources, n=3, m=2, gOn=gOn)
    assert rel(Fs[0], F1) < TOL and rel(c, C1) < TOL
    assert (rel(Gs[0], G1) < TOL) if gOn else not Gs[0].any()


@pytest.mark.parametrize("quadratic", [False, True])
def test_sweep_tune_then_sweep_and_graph(quadratic):
    """The tile autotuner tries every (elements per tile, slots) candidate and keeps the
    fastest; later sweeps and graphs use it and agree with the untuned sweep."""
    mesh = M.icosphere(3, "quadratic") if quadratic else M.icosphere(4, "linear")
    ks = np.linspace(1.0, 9.0, 7)
    n, m = (3, 2) if quadratic else (3, 3)
    print(f"[case] sweep tune quadratic={quadratic}", flush=True)
    (F0, G0, C0), = run_sweep(mesh, ks, n, m)
    dm = DeviceMesh(mesh)
    N, nk = mesh.n_sources, len(ks)
    stride = N * N
    kb = DevBuf(8 * nk)
    kb.upload(np.asarray(ks, dtype=np.float64))
    F, G, C = DevBuf(16 * stride * nk), DevBuf(16 * stride * nk), DevBuf(16 * N)
    ept, slots, ms = dm.sweep_tune(kb.ptr, nk, F.ptr, G.ptr, stride, C.ptr, n=n, m=m)
    assert ept in (16, 32, 64, 128, 0) and slots in (0, 32) and ms > 0
    dm.assemble_sweep(kb.ptr, nk, F.ptr, G.ptr, stride, C.ptr, n=n, m=m)
    sync()
    Fs, Gs, c = read(F, G, C, N, nk, stride)
    for q in range(nk):
        assert rel(Fs[q], F0[q]) < 1e-14 and rel(Gs[q], G0[q]) < 1e-14
    assert rel(c, C0) < 1e-14
    g = dm.sweep_graph(kb.ptr, nk, F.ptr, G.ptr, stride, C.ptr, n=n, m=m)
    g.launch()
    sync()
    Fs, Gs, c = read(F, G, C, N, nk, stride)
    assert all(rel(Fs[q], F0[q]) < 1e-14 and rel(Gs[q], G0[q]) < 1e-14 for q in range(nk))
    g.close()


def test_sweep_graph_replays():
    mesh = M.icosphere(3, "linear")
    ks = np.linspace(2.0, 9.0, 8)
    outs = run_sweep(mesh, ks, 3, 3, graph_replays=3)
    Fo, Go, _ = oracle.assemble(mesh, ks[5], n=3, m=3)
    for Fs, Gs, c in outs:
        assert rel(Fs[5], Fo) < TOL and rel(Gs[5], Go) < TOL


def test_sweep_then_gmres_rigid_sphere():
    """Config-3 shape at reduced size: batched sweep, then per-k GMRES (tol 1e-10) on
    A_k = F_k + diag(1 - C) for a rigid sphere under p_inc = e^{ikz}; solutions against an
    oracle dense solve and the analytic series (theta measured from -z for +z incidence)."""
    import torch
    import analytic
    from bie_amd import solvers
    from torch_solvers import (collocation_operator, collocation_operator_batched, gmres,
                               gmres_batched)

    mesh = M.icosphere(8, "linear")
    N = mesh.n_sources
    ks = np.linspace(0.5, 6.0, 6)
    dev = torch.device("cuda:0")
    dm = DeviceMesh(mesh)
    kd = torch.as_tensor(ks, device=dev)
    F = torch.empty((len(ks), N, N), dtype=torch.complex128, device=dev)
    G = torch.empty_like(F)
    C = torch.empty(N, dtype=torch.complex128, device=dev)
    print("[case] sweep + gmres", flush=True)
    dm.assemble_sweep(kd.data_ptr(), len(ks), F.data_ptr(), G.data_ptr(), N * N, C.data_ptr(),
                      n=4, m=4, gOn=False)
    torch.cuda.synchronize()
    z = mesh.sources[2]
    th = np.arccos(np.clip(-z, -1, 1))
    sols = []
    for q, k in enumerate(ks):
        b = torch.as_tensor(np.exp(1j * k * z), device=dev)
        x, it, rr = gmres(torch, collocation_operator(torch, F[q], C), b, tol=1e-10)
        assert rr <= 1e-10 and it < 100
        p = x.cpu().numpy()
        Fo, _, Co = oracle.assemble(mesh, k, n=4, m=4, gOn=False)
        pd = np.linalg.solve(Fo + np.diag(1.0 - Co), np.exp(1j * k * z))
        assert rel(p, pd) < 1e-8
        pa = analytic.plane_wave_scattering_sphere(k, 1.0, 1.0, th, 1e-8)
        assert rel(p, pa) < 0.0045 * k + 0.001
        sols.append((p, it))
    # the same six systems in one lockstep batched solve
    B = torch.as_tensor(np.exp(1j * ks[:, None] * z[None, :]), device=dev)
    X, its, rrs = gmres_batched(torch, collocation_operator_batched(torch, F, C), B, tol=1e-10)
    for q in range(len(ks)):
        p, it = sols[q]
        assert rrs[q] <= 1e-10 and abs(int(its[q]) - it) <= 1
        assert rel(X[q].cpu().numpy(), p) < 1e-9
    # the library's native batched GMRES on the same systems: same iteration counts (the
    # same algorithm) and the same solutions
    Xn = torch.empty_like(B)
    itn, rrn = solvers.gmres_batched(N, len(ks), F.data_ptr(), N * N, N, C.data_ptr(),
                                     B.data_ptr(), N, Xn.data_ptr(), N, tol=1e-10)
    for q in range(len(ks)):
        p, it = sols[q]
        assert rrn[q] <= 1e-10 and abs(int(itn[q]) - it) <= 1, (q, itn[q], it)
        assert rel(Xn[q].cpu().numpy(), p) < 1e-9
        r = B[q] - (torch.mv(F[q].t(), Xn[q]) + (1.0 - C) * Xn[q])     # explicit residual
        assert float(torch.linalg.norm(r) / torch.linalg.norm(B[q])) <= 1.05e-10


@pytest.mark.parametrize("trans", ["N", "T", "C"])
def test_device_lu_solve_matches_oracle_dense_solve(trans):
    """(F + diag(1 - C)) X = B, and the transposed and adjoint systems, by LU on the device
    (bie_lu_factor_device / bie_lu_solve_device) on F where the assembly wrote it: against a
    host solve of the oracle's matrices (two right-hand sides, complex k)."""
    import torch
    from bie_amd import solvers

    mesh = M.icosphere(4, "quadratic")
    N = mesh.n_sources
    k = 3.0 + 0.1j
    dev = torch.device("cuda:0")
    dm = DeviceMesh(mesh)
    F = torch.empty((N, N), dtype=torch.complex128, device=dev)   # F^T view (column-major F)
    G = torch.empty_like(F)
    C = torch.empty(N, dtype=torch.complex128, device=dev)
    ipiv = torch.empty(N, dtype=torch.int32, device=dev)
    print(f"[case] device LU solve N={N} trans={trans}", flush=True)
    dm.assemble(k, F.data_ptr(), G.data_ptr(), C.data_ptr(), n=3, m=3, gOn=False)
    solvers.lu_factor(N, F.data_ptr(), N, ipiv.data_ptr(), C_ptr=C.data_ptr())
    z = mesh.sources[2]
    Bh = np.stack([np.exp(1j * k.real * z), np.cos(2.0 * z) + 0j], axis=1)
    Bd = torch.as_tensor(np.asfortranarray(Bh).T.copy(), device=dev)   # column-major N x 2
    solvers.lu_solve(N, F.data_ptr(), N, ipiv.data_ptr(), Bd.data_ptr(), N, 2, trans=trans)
    torch.cuda.synchronize()
    X = Bd.cpu().numpy().T
    Fo, _, Co = oracle.assemble(mesh, k, n=3, m=3, gOn=False)
    A = Fo + np.diag(1.0 - Co)
    A = {"N": A, "T": A.T, "C": A.conj().T}[trans]
    Xo = np.linalg.solve(A, Bh)
    assert rel(X, Xo) < 1e-10


def test_device_lu_rejects_a_singular_matrix():
    import torch
    from bie_amd import solvers
    dev = torch.device("cuda:0")
    A = torch.zeros((4, 4), dtype=torch.complex128, device=dev)
    A[0, 0] = 1.0
    ipiv = torch.empty(4, dtype=torch.int32, device=dev)
    with pytest.raises(ValueError, match="singular"):
        solvers.lu_factor(4, A.data_ptr(), 4, ipiv.data_ptr())


@pytest.mark.parametrize("trans", ["N", "T", "C"])
def test_device_gemv_against_numpy(trans):
    import torch
    from bie_amd import solvers
    rng = np.random.default_rng(4)
    m, n = 300, 170
    A = np.asfortranarray(rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n)))
    x = rng.standard_normal(n if trans == "N" else m) + 1j * rng.standard_normal(n if trans == "N" else m)
    dev = torch.device("cuda:0")
    Ad = torch.as_tensor(A.T.copy(), device=dev)
    xd = torch.as_tensor(x, device=dev)
    yd = torch.empty(m if trans == "N" else n, dtype=torch.complex128, device=dev)
    solvers.zgemv(m, n, Ad.data_ptr(), m, xd.data_ptr(), yd.data_ptr(), trans=trans)
    torch.cuda.synchronize()
    op = {"N": A, "T": A.T, "C": A.conj().T}[trans]
    ref = op @ x
    assert rel(yd.cpu().numpy(), ref) < 1e-14
