"""Sparse near-field assembly on the GPU against the oracle (sparse_assemble_parallel!,
src/3d/assembly_collocation_losses.jl:200-405).

Pattern (I, J) must match exactly; G to 1e-12 of its largest entry, F to 1e-12 of the
largest entry of F or G.  (On the elements that contain the source node the in-plane part of
n.d is rounding noise amplified by 1/r^3 at the smallest subdivision cells, about 1e-6 of
the element: F entries there are ~1e-9 of G's and made of that noise.  The kernel evaluates
the geometry of each rule point in the oracle's operation order without fused
multiply-adds, so it reproduces the noise too; a flat mesh's depth-1 F is all noise and is
compared against G's scale.)
"""
import ctypes

import numpy as np
import pytest

import oracle
from bie_amd import _lib
from bie_amd import mesh as M
from bie_amd.assembly import (DeviceMesh, SparsePattern, assemble_parallel,
                              sparse_assemble_parallel)

pytestmark = pytest.mark.gpu

CASES = {
    "lin_ellipsoid": lambda: M.ellipsoid(3, axes=(1.0, 0.8, 1.3)),
    "quad_sphere": lambda: M.icosphere(2, "quadratic"),
    "quad_ellipsoid_linphys": lambda: M.ellipsoid(2, axes=(1.2, 1.0, 0.7), order="quadratic",
                                                  physics="linear"),
}


def _check(I, J, F, G, ref, tag):
    Io, Jo, Fo, Go = ref
    assert np.array_equal(I, Io) and np.array_equal(J, Jo), tag
    gs = np.max(np.abs(Go))
    assert np.max(np.abs(G - Go)) <= 1e-12 * gs, tag
    fs = np.max(np.abs(Fo))
    assert np.max(np.abs(F - Fo)) <= 1e-12 * max(fs, gs), tag


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("depth", [1, 2])
@pytest.mark.parametrize("k", [0.0, 3.0, 2.0 + 0.4j])
def test_sparse_matches_oracle(name, depth, k):
    mesh = CASES[name]()
    print(f"sparse {name} depth={depth} k={k}", flush=True)
    Fs, Gs = sparse_assemble_parallel(mesh, k, mesh.sources, depth=depth)
    ref = oracle.sparse_assemble(mesh, k, depth=depth)
    Fc, Gc = Fs.tocoo(), Gs.tocoo()
    # scipy's COO of a CSR matrix lists entries row by row, columns ascending: the
    # reference's order
    _check(Fc.row, Fc.col, Fc.data, Gc.data, ref, name)
    assert Fs.shape == (mesh.n_sources, mesh.n_sources)


def test_sparse_flags_offset_and_explicit_sources():
    mesh = CASES["quad_sphere"]()
    k = 4.0
    pts = mesh.sources * 1.0  # a copy: explicit source points
    ref = oracle.sparse_assemble(mesh, k, pts, depth=2, offset=0.15)
    Fs, _ = sparse_assemble_parallel(mesh, k, pts, depth=2, offset=0.15, gOn=False)
    _, Gs = sparse_assemble_parallel(mesh, k, pts, depth=2, offset=0.15, fOn=False)
    Fc, Gc = Fs.tocoo(), Gs.tocoo()
    _check(Fc.row, Fc.col, Fc.data, Gc.data, ref, "offset")


def test_sparse_through_assemble_parallel_keyword():
    # assemble_parallel!(...; sparse=true) defaults to depth 2 (assembly_collocation.jl:259)
    mesh = CASES["lin_ellipsoid"]()
    F, G = assemble_parallel(mesh, 1.5, mesh.sources, sparse=True)
    ref = oracle.sparse_assemble(mesh, 1.5, depth=2)
    Fc, Gc = F.tocoo(), G.tocoo()
    _check(Fc.row, Fc.col, Fc.data, Gc.data, ref, "keyword")


def test_sparse_device_handle_reuse_and_hypot_flag():
    mesh = CASES["quad_sphere"]()
    dm = DeviceMesh(mesh, hypot="scaled")
    sp = SparsePattern(dm, depth=1)
    I, J = sp.pattern()
    for k in (1.0, 5.0):
        F, G = sp.assemble(k)
        ref = oracle.sparse_assemble(mesh, k, depth=1, hypot="scaled")
        _check(I, J, F, G, ref, f"device k={k}")
    sp.close()
    dm.close()


def test_sparse_rejects_bad_arguments():
    mesh = CASES["lin_ellipsoid"]()
    hm = _lib.HostMesh(mesh)
    L = _lib.lib()
    nnz = ctypes.c_int64()
    assert L.bie_sparse_count(ctypes.byref(hm.struct), 3, ctypes.byref(nnz)) == _lib.BIE_EINVAL
    assert L.bie_sparse_count(ctypes.byref(hm.struct), 1, ctypes.byref(nnz)) == 0
    n = nnz.value
    I = np.zeros(n + 1, dtype=np.int64)
    J = np.zeros(n + 1, dtype=np.int64)
    F = np.zeros(n + 1, dtype=np.complex128)
    G = np.zeros(n + 1, dtype=np.complex128)
    rc = L.bie_sparse_assemble(ctypes.byref(hm.struct), 1.0, 0.0, None, 0, 1, float("nan"),
                               3, n + 1, _lib._ptr(I), _lib._ptr(J), _lib._ptr(F),
                               _lib._ptr(G))
    assert rc == _lib.BIE_EINVAL and b"nnz" in L.bie_last_error()
    with pytest.raises(ValueError):
        sparse_assemble_parallel(mesh, 1.0, mesh.sources[:, :10], depth=1)


# ---- the fast operators' partial near field (partial_assemble_parallel!,
# src/3d/fast_multipole_method.jl:11-81) ----------------------------------------------------
def _partial_ref(mesh, k, depth, n, sources=None, fOn=True, gOn=True):
    import oracle_np
    els, cols = oracle.connected_sources(mesh, depth)
    src = mesh.sources if sources is None else sources
    F, G = oracle_np.partial_assemble(mesh, k, src, els, cols, n=n, fOn=fOn, gOn=gOn)
    I = np.concatenate([np.full(len(c), i) for i, c in enumerate(cols)])
    return I, np.concatenate(cols), F, G


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("depth", [1, 2])
@pytest.mark.parametrize("n", [1, 3, 7, 13])
def test_partial_matches_oracle(name, depth, n):
    from bie_amd.assembly import partial_assemble_parallel
    mesh = CASES[name]()
    k = 2.0 + (0.3j if n == 7 else 0.0)
    print(f"partial {name} depth={depth} n={n} k={k}", flush=True)
    Fs, Gs = partial_assemble_parallel(mesh, k, n=n, depth=depth)
    ref = _partial_ref(mesh, k, depth, n)
    Fc, Gc = Fs.tocoo(), Gs.tocoo()
    _check(Fc.row, Fc.col, Fc.data, Gc.data, ref, name)


def test_partial_device_handle_flags_and_sources():
    from bie_amd.assembly import partial_assemble_parallel
    mesh = CASES["quad_sphere"]()
    dm = DeviceMesh(mesh)
    sp = SparsePattern(dm, depth=2)
    I, J = sp.pattern()
    nnz = len(I)
    from gpu_helpers import DevBuf, sync
    F, G = DevBuf(16 * nnz), DevBuf(16 * nnz)
    for n in (4, 12, 4):   # the rule table is rebuilt when n changes
        sp.partial_device(3.0, n, F.ptr, G.ptr)
        sync()
        ref = _partial_ref(mesh, 3.0, 2, n)
        f = F.download(np.zeros(nnz, dtype=np.complex128))
        g = G.download(np.zeros(nnz, dtype=np.complex128))
        _check(I, J, f, g, ref, f"device n={n}")
    sp.close()
    dm.close()
    # explicit source points and the flags
    pts = mesh.sources * 1.0
    ref = _partial_ref(mesh, 1.5, 1, 3, sources=pts)
    Fs, _ = partial_assemble_parallel(mesh, 1.5, pts, gOn=False)
    _, Gs = partial_assemble_parallel(mesh, 1.5, pts, fOn=False)
    Fc, Gc = Fs.tocoo(), Gs.tocoo()
    _check(Fc.row, Fc.col, Fc.data, Gc.data, ref, "explicit sources")
    with pytest.raises(ValueError, match="n=5"):
        partial_assemble_parallel(mesh, 1.0, n=5)
