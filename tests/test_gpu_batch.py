"""Uneven batch (one persistent launch over many meshes, each with its own k) against
single-mesh device assembly and the oracle, for every schedule configuration."""
import numpy as np
import pytest

import oracle
from bie_amd import _lib
from bie_amd import mesh as M
from bie_amd.assembly import Batch, DeviceMesh
from gpu_helpers import DevBuf, sync

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def read(buf, shape):
    out = np.zeros(int(np.prod(shape)), dtype=np.complex128)
    buf.download(out)
    return out.reshape(shape, order="F")


def outputs(meshes):
    bufs = []
    for mesh in meshes:
        N = mesh.n_sources
        bufs.append((DevBuf(16 * N * N), DevBuf(16 * N * N), DevBuf(16 * N)))
    return bufs


def fill(bufs, value=np.nan):
    for F, G, C in bufs:
        for b in (F, G, C):
            b.upload(np.full(b.nbytes // 8, value))


def single(dm, k, n, m):
    N = dm.n_sources
    F, G, C = DevBuf(16 * N * N), DevBuf(16 * N * N), DevBuf(16 * N)
    dm.assemble(k, F.ptr, G.ptr, C.ptr, n=n, m=m)
    sync()
    return read(F, (N, N)), read(G, (N, N)), read(C, (N,))


def linear_batch():
    # uneven sizes, aspect ratios and wavenumbers (real and complex)
    return [M.ellipsoid(2, (1.0, 0.7, 0.5)), M.icosphere(4, "linear", 1.3),
            M.ellipsoid(3, (2.0, 0.5, 1.0)), M.icosphere(1, "linear"),
            M.ellipsoid(5, (0.6, 1.5, 0.9))], [1.0, 7.5 + 0.3j, 3.0, 0.5, 12.0]


def run_batch(meshes, ks, n, m, config=None, tune=False, flags=(True, True, True)):
    dms = [DeviceMesh(x) for x in meshes]
    b = Batch(dms, n=n, m=m)
    if config:
        b.configure(*config)
    bufs = outputs(meshes)
    fill(bufs)
    fOn, gOn, cOn = flags
    args = (ks, [F.ptr if fOn else 0 for F, _, _ in bufs], [G.ptr if gOn else 0 for _, G, _ in bufs],
            [C.ptr for _, _, C in bufs])
    if tune:
        best = b.tune(*args, fOn=fOn, gOn=gOn, cOn=cOn)
        assert best > 0
    else:
        b.assemble(*args, fOn=fOn, gOn=gOn, cOn=cOn)
    sync()
    res = [(read(F, (x.n_sources,) * 2), read(G, (x.n_sources,) * 2), read(C, (x.n_sources,)))
           for x, (F, G, C) in zip(meshes, bufs)]
    return res, dms, b


@pytest.mark.parametrize("config", [(64, 64, 0), (32, 32, 1), (128, 64, 1), (16, 32, 0),
                                    (1024, 64, 0)])
def test_linear_batch_matches_oracle(config):
    meshes, ks = linear_batch()
    print(f"[case] batch linear config={config}", flush=True)
    res, dms, b = run_batch(meshes, ks, 3, 3, config=config)
    assert b.info()["n_tiles"] > 0
    for mesh, k, (F, G, C) in zip(meshes, ks, res):
        Fo, Go, Co = oracle.assemble(mesh, k, n=3, m=3)
        assert rel(F, Fo) < TOL and rel(G, Go) < TOL and rel(C, Co) < TOL


def test_batch_equals_single_mesh_assembly():
    meshes, ks = linear_batch()
    res, dms, _ = run_batch(meshes, ks, 4, 4)
    for dm, k, (F, G, C) in zip(dms, ks, res):
        Fs, Gs, Cs = single(dm, k, 4, 4)
        # same pair quadrature; only the order of the shared-column additions differs
        assert rel(F, Fs) < 1e-14 and rel(G, Gs) < 1e-14 and rel(C, Cs) < 1e-14


def test_quadratic_batch_with_mid_edge_pairs():
    meshes = [M.icosphere(2, "quadratic"), M.ellipsoid(3, (1.0, 0.6, 1.4), "quadratic"),
              M.icosphere(1, "quadratic", 0.8)]
    ks = [2.0, 5.0 + 0.5j, 1.0]
    print("[case] batch quadratic", flush=True)
    res, _, _ = run_batch(meshes, ks, 3, 2, config=(32, 64, 0))
    for mesh, k, (F, G, C) in zip(meshes, ks, res):
        Fo, Go, Co = oracle.assemble(mesh, k, n=3, m=2)
        assert rel(F, Fo) < TOL and rel(G, Go) < TOL and rel(C, Co) < TOL


def test_batch_flags_zero_disabled_outputs():
    meshes, ks = linear_batch()
    meshes, ks = meshes[:3], ks[:3]
    res, _, _ = run_batch(meshes, ks, 2, 2, flags=(False, True, True))
    for mesh, k, (F, G, C) in zip(meshes, ks, res):
        Fo, Go, Co = oracle.assemble(mesh, k, n=2, m=2)
        assert np.isnan(F).all()    # disabled with a NULL pointer: never written
        assert rel(G, Go) < TOL and rel(C, Co) < TOL


def test_batch_disabled_output_given_is_zeroed():
    meshes, ks = linear_batch()
    meshes, ks = meshes[:2], ks[:2]
    dms = [DeviceMesh(x) for x in meshes]
    b = Batch(dms, n=2, m=2)
    bufs = outputs(meshes)
    fill(bufs)
    b.assemble(ks, [F.ptr for F, _, _ in bufs], [G.ptr for _, G, _ in bufs],
               [C.ptr for _, _, C in bufs], gOn=False)
    sync()
    for mesh, k, (Fb, Gb, Cb) in zip(meshes, ks, bufs):
        N = mesh.n_sources
        Fo, _, Co = oracle.assemble(mesh, k, n=2, m=2)
        assert rel(read(Fb, (N, N)), Fo) < TOL
        assert not read(Gb, (N, N)).any()


def test_batch_tune_then_repeat_calls():
    meshes, ks = linear_batch()
    res, dms, b = run_batch(meshes, ks, 2, 2, tune=True)
    cfg = b.info()
    assert cfg["elems_per_tile"] in (32, 64, 128) and cfg["slots"] in (0, 32)
    for mesh, k, (F, G, C) in zip(meshes, ks, res):
        Fo, Go, Co = oracle.assemble(mesh, k, n=2, m=2)
        assert rel(F, Fo) < TOL and rel(G, Go) < TOL and rel(C, Co) < TOL
    # repeated calls reuse the staging table: new k values each time
    bufs = outputs(meshes)
    for rep in range(3):
        ks2 = [k + 0.25 * (rep + 1) for k in ks]
        b.assemble(ks2, [F.ptr for F, _, _ in bufs], [G.ptr for _, G, _ in bufs],
                   [C.ptr for _, _, C in bufs])
    sync()
    for mesh, k, (Fb, Gb, Cb) in zip(meshes, ks2, bufs):
        N = mesh.n_sources
        Fo, Go, _ = oracle.assemble(mesh, k, n=2, m=2)
        assert rel(read(Fb, (N, N)), Fo) < TOL and rel(read(Gb, (N, N)), Go) < TOL


def test_batch_meshes_shared_between_batches_and_repeated():
    # one mesh twice (two wavenumbers) in a batch, and a second batch with another schedule
    # on the same meshes: each batch rebuilds what the other replaced
    meshes, ks = linear_batch()
    dms = [DeviceMesh(x) for x in meshes[:3]]
    inst = [dms[0], dms[1], dms[0], dms[2]]
    kk = [1.0, 2.0, 6.0, 3.0]
    b1, b2 = Batch(inst, n=2, m=2), Batch(dms, n=2, m=2)
    b1.configure(32, 32, 0)
    b2.configure(128, 64, 1)
    mesh_of = [meshes[0], meshes[1], meshes[0], meshes[2]]
    bufs1, bufs2 = outputs(mesh_of), outputs(meshes[:3])
    for rep in range(2):
        b1.assemble(kk, [F.ptr for F, _, _ in bufs1], [G.ptr for _, G, _ in bufs1],
                    [C.ptr for _, _, C in bufs1])
        b2.assemble(ks[:3], [F.ptr for F, _, _ in bufs2], [G.ptr for _, G, _ in bufs2],
                    [C.ptr for _, _, C in bufs2])
    sync()
    for mesh, k, (Fb, Gb, Cb) in list(zip(mesh_of, kk, bufs1)) + list(zip(meshes[:3], ks, bufs2)):
        N = mesh.n_sources
        Fo, Go, Co = oracle.assemble(mesh, k, n=2, m=2)
        assert rel(read(Fb, (N, N)), Fo) < TOL and rel(read(Gb, (N, N)), Go) < TOL
        assert rel(read(Cb, (N,)), Co) < TOL


def test_batch_rejects_mixed_kinds_and_bad_config():
    dms = [DeviceMesh(M.icosphere(1, "linear")), DeviceMesh(M.icosphere(1, "quadratic"))]
    with pytest.raises(_lib.BieError) as ei:
        Batch(dms, n=2, m=2)
    assert ei.value.code == _lib.BIE_EUNSUPPORTED
    with pytest.raises(_lib.BieError) as ei:
        Batch([dms[1], dms[1]], n=2, m=2)     # a quadratic mesh twice shares its queue
    assert ei.value.code == _lib.BIE_EINVAL
    b = Batch(dms[:1], n=2, m=2)
    for bad in ((4, 64, 0), (64, 48, 0), (64, 64, 2), (64, -1, 0)):
        with pytest.raises(_lib.BieError) as ei:
            b.configure(*bad)
        assert ei.value.code == _lib.BIE_EINVAL


@pytest.mark.parametrize("quadratic", [False, True])
def test_batch_graph_replays_match_eager_batch(quadratic):
    """The uneven batch captured into a HIP graph (kernels only), replayed three times with
    eager work on the same meshes in between, against the eager batched assembly."""
    if quadratic:
        meshes = [M.icosphere(2, "quadratic"), M.ellipsoid(3, (1.0, 0.6, 1.4), "quadratic")]
        ks, n, m = [2.0, 5.0], 3, 2
    else:
        meshes, ks = linear_batch()
        n, m = 3, 3
    print(f"[case] batch graph quadratic={quadratic}", flush=True)
    ref, dms, b = run_batch(meshes, ks, n, m)
    bufs = outputs(meshes)
    args = (ks, [F.ptr for F, _, _ in bufs], [G.ptr for _, G, _ in bufs],
            [C.ptr for _, _, C in bufs])
    g = b.graph(*args)
    other = outputs(meshes)
    for replay in range(3):
        fill(bufs)
        sync()
        g.launch()
        sync()
        # eager work on the same batch and meshes between the replays
        b.assemble([k + 1.0 for k in ks], [F.ptr for F, _, _ in other],
                   [G.ptr for _, G, _ in other], [C.ptr for _, _, C in other])
        dms[0].assemble(ks[0] + 2.0, other[0][0].ptr, other[0][1].ptr, other[0][2].ptr, n=n,
                        m=m)
        sync()
        for x, (F, G, C), (Fr, Gr, Cr) in zip(meshes, bufs, ref):
            N = x.n_sources
            assert rel(read(F, (N, N)), Fr) < 1e-14, replay
            assert rel(read(G, (N, N)), Gr) < 1e-14, replay
            assert rel(read(C, (N,)), Cr) < 1e-14, replay
    g.close()
