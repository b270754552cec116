"""The near field of discontinuous triangular physics on the GPU (the reference's
sparse_assemble_parallel! branch for it, src/3d/assembly_collocation_losses.jl:233-257:
connected_sources :41-96, polar rules create_polar_elements :409-447) against the NumPy
oracle (oracle_np.sparse_assemble_disc, pinned in test_sparse_disc_cpu.py), and the
partial near field on the same pattern.

Bar: the pattern exactly; G within 1e-12 of its largest entry; F within 1e-12 of the
largest entry of F or G, except on the row's own element, where the source lies on the
element and n.d carries its rounding noise divided by r^3 at the polar rule's innermost
points (the whole of F on a flat element: ~1e-12 - 5e-11 of G): there within 1e-9.
"""
import numpy as np
import pytest

import oracle_np
from bie_amd import mesh as M
from bie_amd.assembly import (DeviceMesh, SparsePattern, assemble_parallel,
                              partial_assemble_parallel, sparse_assemble_parallel)
from gpu_helpers import DevBuf, sync

pytestmark = pytest.mark.gpu

CASES = [("linear", "disctriconstant", 8), ("linear", "disctrilinear", 10),
         ("linear", "disctriquadratic", 40), ("quadratic", "disctriconstant", 8),
         ("quadratic", "disctrilinear", 12), ("quadratic", "disctriquadratic", 40)]


def disc_mesh(gorder, physics, f=2):
    return M.with_discontinuous_physics(M.perturb_radial(M.icosphere(f, gorder), 0.02, 7),
                                        physics)


def _check(I, J, F, G, ref, mesh, tag):
    Io, Jo, Fo, Go = ref
    assert np.array_equal(I, Io) and np.array_equal(J, Jo), tag
    gs = np.max(np.abs(Go))
    assert np.max(np.abs(G - Go)) <= 1e-12 * gs, tag
    scale = max(np.max(np.abs(Fo)), gs)
    npk = mesh.physics_topology.shape[0]
    own = (J // npk) == (I // npk)      # columns of the row's own element
    err = np.abs(F - Fo)
    assert np.max(err[~own], initial=0.0) <= 1e-12 * scale, tag
    assert np.max(err[own]) <= 1e-9 * scale, tag


@pytest.mark.parametrize("gorder,physics,ng", CASES)
@pytest.mark.parametrize("depth", [1, 2])
def test_near_field_matches_oracle(gorder, physics, ng, depth):
    mesh = disc_mesh(gorder, physics)
    k = 3.0 if depth == 1 else 2.0 + 0.4j
    print(f"[case] disc sparse {gorder} {physics} Ngauss={ng} depth={depth} k={k}", flush=True)
    Fs, Gs = sparse_assemble_parallel(mesh, k, depth=depth, Ngauss=ng)
    ref = oracle_np.sparse_assemble_disc(mesh, k, depth=depth, Ngauss=ng)
    Fc, Gc = Fs.tocoo(), Gs.tocoo()
    _check(Fc.row, Fc.col, Fc.data, Gc.data, ref, mesh, f"{gorder} {physics} d{depth}")
    assert Fs.shape == (mesh.n_sources, mesh.n_sources)


def test_k0_flags_and_explicit_sources():
    mesh = disc_mesh("quadratic", "disctrilinear")
    pts = mesh.sources * 1.0
    print("[case] disc sparse k=0 / flags / explicit sources", flush=True)
    ref = oracle_np.sparse_assemble_disc(mesh, 0.0, pts, depth=2, Ngauss=6)
    Fs, _ = sparse_assemble_parallel(mesh, 0.0, pts, depth=2, Ngauss=6, gOn=False)
    _, Gs = sparse_assemble_parallel(mesh, 0.0, pts, depth=2, Ngauss=6, fOn=False)
    Fc, Gc = Fs.tocoo(), Gs.tocoo()
    _check(Fc.row, Fc.col, Fc.data, Gc.data, ref, mesh, "k=0 flags")
    # the keyword route: assemble_parallel(...; sparse=true) takes depth 2 and Ngauss 40
    Fk, Gk = assemble_parallel(mesh, 1.5, sparse=True)
    ref = oracle_np.sparse_assemble_disc(mesh, 1.5, depth=2, Ngauss=40)
    Fc, Gc = Fk.tocoo(), Gk.tocoo()
    _check(Fc.row, Fc.col, Fc.data, Gc.data, ref, mesh, "sparse keyword")


def test_device_handle_and_forced_order():
    """The device handle with its own Ngauss; quadratic physics takes Ngauss = 40 whatever
    is asked (the reference overrides it)."""
    for physics, ng, ng_ref in (("disctrilinear", 6, 6), ("disctriquadratic", 6, 40)):
        mesh = disc_mesh("linear", physics)
        print(f"[case] disc sparse handle {physics} Ngauss={ng}", flush=True)
        dm = DeviceMesh(mesh)
        sp = SparsePattern(dm, depth=2, Ngauss=ng)
        I, J = sp.pattern()
        nnz = len(I)
        F, G = DevBuf(16 * nnz), DevBuf(16 * nnz)
        for k in (2.5, 4.0):
            sp.assemble_device(k, F.ptr, G.ptr)
            sync()
            ref = oracle_np.sparse_assemble_disc(mesh, k, depth=2, Ngauss=ng_ref)
            _check(I, J, F.download(np.zeros(nnz, dtype=np.complex128)),
                   G.download(np.zeros(nnz, dtype=np.complex128)), ref, mesh, f"k={k}")
        sp.close()
        dm.close()


@pytest.mark.parametrize("physics", ["disctriconstant", "disctrilinear", "disctriquadratic"])
@pytest.mark.parametrize("depth", [1, 2])
def test_partial_near_field(physics, depth):
    """partial_assemble_parallel! on discontinuous physics: the discontinuous pattern, every
    pair with gauss_points_triangle(n) and the discontinuous basis."""
    mesh = disc_mesh("quadratic", physics)
    npk = mesh.physics_topology.shape[0]
    print(f"[case] disc partial {physics} depth={depth}", flush=True)
    Fs, Gs = partial_assemble_parallel(mesh, 2.0 + 0.2j, n=3, depth=depth)
    els, cols = oracle_np.connected_sources_disc(mesh, depth)
    rows_el = [els[r // npk] for r in range(mesh.n_sources)]
    rows_col = [cols[r // npk] for r in range(mesh.n_sources)]
    Fo, Go = oracle_np.partial_assemble(mesh, 2.0 + 0.2j, mesh.sources, rows_el, rows_col, n=3)
    I = np.concatenate([np.full(len(c), r) for r, c in enumerate(rows_col)])
    ref = (I, np.concatenate(rows_col), Fo, Go)
    Fc, Gc = Fs.tocoo(), Gs.tocoo()
    _check(Fc.row, Fc.col, Fc.data, Gc.data, ref, mesh, f"partial {physics}")


def test_batched_pairs_with_moved_sources():
    """The batches are grouped by the rules the mesh sources select; explicit sources moved
    off the nodes select other rules for some rows, which the kernel must follow (it selects
    every row's rule itself).  Depth 2, continuous and discontinuous physics."""
    import oracle
    rng = np.random.default_rng(11)
    for mesh in (M.icosphere(2, "quadratic"), disc_mesh("quadratic", "disctrilinear")):
        N = mesh.n_sources
        pts = mesh.sources + 0.02 * rng.standard_normal((3, N))
        disc = mesh.physics_kind & 0x100
        print(f"[case] batched pairs, moved sources, disc={bool(disc)} N={N}", flush=True)
        Fs, Gs = sparse_assemble_parallel(mesh, 2.5, pts, depth=2, Ngauss=8)
        if disc:
            ref = oracle_np.sparse_assemble_disc(mesh, 2.5, pts, depth=2, Ngauss=8)
        else:
            ref = oracle.sparse_assemble(mesh, 2.5, pts, depth=2)
        Fc, Gc = Fs.tocoo(), Gs.tocoo()
        Io, Jo, Fo, Go = ref
        assert np.array_equal(Fc.row, Io) and np.array_equal(Fc.col, Jo)
        # off the surface nothing is rounding noise: every entry to 1e-12 of the largest
        scale = max(np.max(np.abs(Fo)), np.max(np.abs(Go)))
        assert np.max(np.abs(Gc.data - Go)) <= 1e-12 * np.max(np.abs(Go))
        assert np.max(np.abs(Fc.data - Fo)) <= 1e-12 * scale


@pytest.mark.parametrize("physics,depth,ng", [("disctrilinear", 1, 40), ("disctrilinear", 2, 6),
                                              ("disctriquadratic", 1, 40)])
def test_reference_mesh(physics, depth, ng):
    """The reference's own 840-element sphere_1m (quadratic geometry) with discontinuous
    physics from its loader, at the reference's default Ngauss and at a small one."""
    import os
    gold = os.path.join(os.path.dirname(__file__), "golden")
    mesh = M.load_comsol(os.path.join(gold, "sphere_1m"), "quadratic", physics)
    k = 2 * np.pi * 100.0 / 340.0
    print(f"[case] disc sparse sphere_1m {physics} depth={depth} Ngauss={ng} "
          f"N={mesh.n_sources}", flush=True)
    Fs, Gs = sparse_assemble_parallel(mesh, k, depth=depth, Ngauss=ng)
    ref = oracle_np.sparse_assemble_disc(mesh, k, depth=depth, Ngauss=ng)
    Fc, Gc = Fs.tocoo(), Gs.tocoo()
    _check(Fc.row, Fc.col, Fc.data, Gc.data, ref, mesh, f"sphere_1m {physics}")
