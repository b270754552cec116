"""Parity of the HIP assembly (through the C ABI) with the CPU oracle.

Bar: rule indices bit-exact; F, G, C within 1e-12 relative Frobenius error of the
oracle (the north star's tolerance; the reference itself runs @fastmath, so bitwise
agreement of floating-point sums is not expected).  Full-size checks use rows sampled
from the 10242-node quadratic sphere plus size-independent identities.
"""
import ctypes
import os

import numpy as np
import pytest

import analytic
import oracle
from bie_amd import _lib
from bie_amd import mesh as M
from bie_amd.assembly import DeviceMesh, assemble_parallel

GOLD = os.path.join(os.path.dirname(__file__), "golden")
TOL = 1e-12

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


class DevBuf:
    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        _lib.check(_lib.lib().bie_device_malloc(ctypes.byref(p), int(nbytes)))
        self.ptr = p.value or 0
        self.nbytes = nbytes

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        _lib.check(_lib.lib().bie_memcpy_h2d(ctypes.c_void_p(self.ptr), _lib._ptr(arr), arr.nbytes))

    def download(self, arr):
        _lib.check(_lib.lib().bie_memcpy_d2h(_lib._ptr(arr), ctypes.c_void_p(self.ptr), arr.nbytes))
        return arr

    def __del__(self):
        if self.ptr:
            _lib.lib().bie_device_free(ctypes.c_void_p(self.ptr))


def gpu_rules(mesh, targets=None, hypot="julia1.9"):
    dm = DeviceMesh(mesh, hypot=hypot)
    if targets is None:
        M_ = mesh.n_sources
        tbuf = None
    else:
        M_ = targets.shape[1]
        tbuf = DevBuf(24 * M_)
        tbuf.upload(np.asfortranarray(targets).ravel(order="F"))
    out = DevBuf(M_ * mesh.n_elements)
    dm.select_rules(out.ptr, targets_ptr=tbuf.ptr if tbuf else 0, n_targets=M_)
    _lib.check(_lib.lib().bie_device_synchronize())
    return out.download(np.zeros((M_, mesh.n_elements), dtype=np.int8))


MESHES = {
    "lin8": lambda: M.icosphere(8, "linear"),
    "quad4": lambda: M.icosphere(4, "quadratic"),
    "quad3_linphys": lambda: M.icosphere(3, "quadratic", physics="linear"),
    "sphere_1m": lambda: M.load_comsol(os.path.join(GOLD, "sphere_1m")),
    "sphere_1m_lin": lambda: M.load_comsol(os.path.join(GOLD, "sphere_1m"), "linear", "linear"),
    "ellipsoid": lambda: M.ellipsoid(5, (0.7, 1.3, 2.0), "linear"),
}


@pytest.mark.parametrize("name", ["lin8", "quad4", "sphere_1m", "ellipsoid"])
def test_rule_selection_bit_exact(name):
    mesh = MESHES[name]()
    print(f"[case] selection {name} N={mesh.n_sources}", flush=True)
    assert np.array_equal(gpu_rules(mesh), oracle.select(mesh))


def test_rule_selection_ties_and_field_points():
    mesh = M.icosphere(6, "quadratic")
    rng = np.random.default_rng(7)
    c = mesh.coordinates
    # exact midpoints between node pairs, element centroids, and scaled copies: tie-heavy
    i, j = rng.integers(0, c.shape[1], (2, 400))
    T = [0.5 * (c[:, i] + c[:, j]), c[:, mesh.topology[:3, :50]].mean(axis=1),
         1.7 * c[:, :200], 0.3 * c[:, :200], rng.standard_normal((3, 300))]
    T = np.hstack(T)
    assert np.array_equal(gpu_rules(mesh, T), oracle.select(mesh, T))


@pytest.mark.parametrize("name", ["lin8", "quad4", "sphere_1m"])
def test_rule_selection_with_the_scaled_hypot_bit_exact(name):
    """BIE_MESH_SCALED_HYPOT: the tied pairs take the later Julia releases' hypot rounding;
    rule indices bit-exact against the oracle's selection with that formula, which differs
    from the default one on these symmetric meshes."""
    mesh = MESHES[name]()
    print(f"[case] scaled-hypot selection {name}", flush=True)
    got = gpu_rules(mesh, hypot="scaled")
    assert np.array_equal(got, oracle.select(mesh, hypot="scaled"))
    if name != "sphere_1m":
        assert not np.array_equal(got, oracle.select(mesh))


def test_assembly_with_the_scaled_hypot():
    """F, G, C with BIE_MESH_SCALED_HYPOT against the oracle's frozen-rule assembly at the
    rules that formula selects, on a mesh where about 0.2 % of the pairs change rule."""
    mesh = M.icosphere(8, "linear")
    rules = oracle.select(mesh, hypot="scaled")
    assert (rules != oracle.select(mesh)).sum() > 0
    for k in (1.0, 4.0 + 0.5j):
        F, G, C = assemble_parallel(mesh, k, mesh.sources, n=2, m=2, hypot="scaled")
        Fo, Go, Co = oracle.assemble_frozen(mesh, k, rules, n=2, m=2)
        assert rel(F, Fo) < TOL and rel(G, Go) < TOL and rel(C, Co) < TOL, k


@pytest.mark.parametrize("rule", [1, 2, 3, 4, 5, 6])
def test_element_geometry_matches_oracle(rule):
    mesh = M.icosphere(3, "quadratic")
    dm = DeviceMesh(mesh)
    y, nr, jw = dm.element_geometry(rule, 4, 3)
    yo, no, jo = oracle.element_geometry(mesh, rule, 4, 3)
    assert np.max(np.abs(y - yo)) < 1e-14
    assert np.max(np.abs(nr - no)) < 1e-13
    assert np.max(np.abs(jw - jo)) < 1e-15


CASES = [
    ("lin8", 1.0, 2, 2),          # config 1: 642-node linear sphere, k=1, 4-point rule
    ("lin8", 7.5, 4, 4),
    ("quad4", 10.0, 3, 2),        # quadratic sphere at the full-size k and 6-point rule
    ("quad4", 10.0, 4, 4),
    ("quad4", 2.0 + 0.4j, 3, 3),  # complex k (lossy caller)
    ("quad3_linphys", 3.0, 4, 4), # quadratic geometry, linear physics
    ("sphere_1m", 2 * np.pi * 100 / 340, 4, 4),
    ("sphere_1m_lin", 4.0, 3, 2),
    ("ellipsoid", 5.0, 4, 4),
]


@pytest.mark.parametrize("name,k,n,m", CASES)
def test_assembly_parity(name, k, n, m):
    mesh = MESHES[name]()
    print(f"[case] assembly {name} N={mesh.n_sources} k={k} (n,m)=({n},{m})", flush=True)
    F, G, C = assemble_parallel(mesh, k, mesh.sources, n=n, m=m)
    Fo, Go, Co = oracle.assemble(mesh, k, n=n, m=m)
    assert rel(F, Fo) < TOL and rel(G, Go) < TOL and rel(C, Co) < TOL
    assert np.isfortran(F) and F.shape == (mesh.n_sources, mesh.n_sources)


@pytest.mark.parametrize("k,tol", [(60.0, TOL), (5.0e7, 1e-6)])
def test_assembly_parity_large_kr(k, tol):
    """e^{ikr} at kr up to 120 (table-driven sincos well away from small arguments) and at
    kr up to 1e8, where points on both sides of 2^26 take the table path and the library
    sincos.  There the phase can only agree to the rounding of r itself, |d(kr)| ~ kr 4e-16
    ~ 4e-8, hence the looser tolerance; C does not depend on k."""
    mesh = MESHES["lin8"]()
    print(f"[case] assembly lin8 N={mesh.n_sources} k={k} (n,m)=(2,2)", flush=True)
    F, G, C = assemble_parallel(mesh, k, mesh.sources, n=2, m=2)
    Fo, Go, Co = oracle.assemble(mesh, k, n=2, m=2)
    assert np.isfinite(F).all() and np.isfinite(G).all()
    assert rel(F, Fo) < tol and rel(G, Go) < tol and rel(C, Co) < TOL


def test_field_points_and_flags():
    mesh = MESHES["quad4"]()
    rng = np.random.default_rng(3)
    X = rng.standard_normal((3, 77))
    X *= (0.2 + 2.5 * rng.random(77)) / np.linalg.norm(X, axis=0)
    print("[case] field points quad4 M=77", flush=True)
    F, G, C = assemble_parallel(mesh, 4.0, X, n=3, m=2)
    Fo, Go, Co = oracle.assemble(mesh, 4.0, targets=X, n=3, m=2)
    assert rel(F, Fo) < TOL and rel(G, Go) < TOL and rel(C, Co) < TOL
    F, G, C = assemble_parallel(mesh, 4.0, X, n=3, m=2, fOn=False, cOn=False)
    assert not F.any() and not C.any() and rel(G, Go) < TOL
    F, G, C = assemble_parallel(mesh, 4.0, X, n=3, m=2, gOn=False)
    assert not G.any() and rel(F, Fo) < TOL and rel(C, Co) < TOL


def test_empty_targets():
    mesh = M.icosphere(2, "linear")
    F, G, C = assemble_parallel(mesh, 1.0, np.zeros((3, 0)), n=2, m=2)
    assert F.shape == (0, mesh.n_sources) and C.shape == (0,)


def _submesh(mesh, elems):
    """The elements `elems` of a mesh as a mesh of their own (nodes re-indexed)."""
    topo = mesh.topology[:, elems]
    ptopo = mesh.physics_topology[:, elems]
    g_nodes, g_inv = np.unique(topo, return_inverse=True)
    p_nodes, p_inv = np.unique(ptopo, return_inverse=True)
    return M.Mesh3d(mesh.coordinates[:, g_nodes], g_inv.reshape(topo.shape),
                    mesh.sources[:, p_nodes], p_inv.reshape(ptopo.shape),
                    mesh.geometry_kind, mesh.physics_kind)


@pytest.mark.parametrize("kind,n_el", [("linear", 1), ("linear", 3), ("quadratic", 2)])
def test_fewer_elements_than_waves_per_step(kind, n_el):
    """Meshes of 1-3 elements: most of the 8 waves of every slot-kernel step have no
    element; collocation and field targets (some on nodes, one far away) against the
    oracle."""
    full = M.icosphere(1, kind)
    # elements sharing nodes with element 0, so the patch is connected
    first = set(full.physics_topology[:, 0].tolist())
    near = [e for e in range(full.n_elements)
            if first & set(full.physics_topology[:, e].tolist())]
    mesh = _submesh(full, near[:n_el])
    assert mesh.n_elements == n_el
    rng = np.random.default_rng(n_el)
    X = np.concatenate([mesh.sources[:, :2], rng.standard_normal((3, 5)),
                        np.array([[30.0], [-4.0], [2.0]])], axis=1)
    for k in (0.0, 3.0):
        print(f"[case] {kind} patch of {n_el} element(s) k={k}", flush=True)
        for targets in (mesh.sources, X):
            F, G, C = assemble_parallel(mesh, k, targets, n=3, m=2)
            Fo, Go, Co = oracle.assemble(mesh, k, targets=None if targets is mesh.sources
                                         else targets, n=3, m=2)
            assert rel(G, Go) < TOL
            # on one flat triangle every collocation point lies in the element's plane, so
            # n.(x - y) = 0 and F, C vanish exactly: both sides are round-off there, and
            # their error is measured against the size of G (same integral scale on this
            # unit-sphere patch) instead of against themselves
            scale = max(np.linalg.norm(Fo), np.linalg.norm(Co), np.linalg.norm(Go))
            assert np.linalg.norm(F - Fo) < TOL * scale
            assert np.linalg.norm(C - Co) < TOL * scale


def _golden():
    import sys
    if GOLD not in sys.path:
        sys.path.insert(0, GOLD)
    import make_fixtures
    return make_fixtures, np.load(os.path.join(GOLD, "oracle_vectors.npz"))


@pytest.mark.parametrize("name", ["coarse_quad", "coarse_lin", "ico2_lin", "ico1_quad"])
def test_gpu_matches_golden_vectors(name):
    """The committed golden vectors (the reference's coarsest COMSOL sphere, quadratic and
    linear, and two small geodesic spheres at k = 0, 1, 5 + 0.3i): rule indices bit-exact,
    F, G, C within the parity tolerance."""
    mf, gv = _golden()
    mesh = mf.meshes()[name]
    print(f"[case] golden vectors {name}", flush=True)
    assert np.array_equal(gpu_rules(mesh), gv[f"{name}_rules"])
    for j, k in enumerate(mf.KS):
        if f"{name}_k{j}_F" not in gv:
            continue
        F, G, C = assemble_parallel(mesh, k, n=mf.NM[0], m=mf.NM[1])
        assert rel(F, gv[f"{name}_k{j}_F"]) < TOL
        assert rel(G, gv[f"{name}_k{j}_G"]) < TOL
        assert rel(C, gv[f"{name}_k{j}_C"]) < TOL


@pytest.mark.parametrize("i", range(3))
def test_gpu_far_entries_match_mpmath(i):
    """Single flat elements, far targets, (n,m) = (12,12): the HIP path against mpmath's
    adaptive quadrature directly (no oracle in between)."""
    mf, gv = _golden()
    one = mf.single_element(mf.meshes()["coarse_lin"], int(gv[f"far_{i}_e"]))
    x = gv[f"far_{i}_x"].reshape(3, 1)
    k = complex(gv[f"far_{i}_k"])
    print(f"[case] far entry {i} vs mpmath (12,12)", flush=True)
    F, G, _ = assemble_parallel(one, k, x, n=12, m=12)
    assert rel(F[0], gv[f"far_{i}_F"]) < 1e-13
    assert rel(G[0], gv[f"far_{i}_G"]) < 1e-13


def test_reference_tolerance_table_on_gpu_output():
    # test/3d/assembly_test.jl:12-44 with the GPU's F, C
    k = 2 * np.pi * 100.0 / 340.0
    for go, po, tol in [("quadratic", "geometry", 2.8e-5), ("linear", "linear", 0.037),
                        ("quadratic", "linear", 0.048)]:
        mesh = M.load_comsol(os.path.join(GOLD, "sphere_1m"), go, po)
        F, _, C = assemble_parallel(mesh, k, mesh.sources, gOn=False)
        pI = analytic.incoming_wave([np.pi / 2, 0.0], 1.0, mesh.sources, k)
        p = np.linalg.solve(F + np.diag(1.0 - C), pI)
        th = np.arccos(np.clip(mesh.sources[0], -1, 1))
        pa = analytic.plane_wave_scattering_sphere(k, 1.0, 1.0, th, 1e-6)
        assert rel(p, pa) < tol, (go, po)


def test_config1_pulsating_sphere_lu():
    mesh = M.icosphere(8, "linear")
    Z0 = 1.21 * 343.0
    F, G, C = assemble_parallel(mesh, 1.0, mesh.sources, n=2, m=2)
    p = np.linalg.solve(F + np.diag(1 - C), -1j * Z0 * (G @ np.ones(mesh.n_sources)))
    pa = analytic.pulsating_sphere_exterior(1.0, 1.0, 1.0, Z0, 1.0)
    assert np.max(np.abs(p - pa)) / abs(pa) < 5e-3


@pytest.mark.parametrize("nm", [(3, 2), (4, 4)])
def test_config2_full_size_rows_and_identities(nm):
    n, m = nm
    mesh = M.icosphere(16, "quadratic")   # 10242 nodes, 5120 quadratic triangles
    k = 10.0
    print(f"[case] full-size quadratic sphere N={mesh.n_sources} k={k} (n,m)={nm}", flush=True)
    F, G, C = assemble_parallel(mesh, k, mesh.sources, n=n, m=m)
    assert np.isfinite(F).all() and np.isfinite(G).all() and np.isfinite(C).all()
    rows = np.sort(np.random.default_rng(11).choice(mesh.n_sources, 40, replace=False))
    Fo, Go, Co = oracle.assemble(mesh, k, targets=mesh.sources[:, rows], n=n, m=m)
    assert rel(F[rows], Fo) < TOL and rel(G[rows], Go) < TOL and rel(C[rows], Co) < TOL
    # size-independent identities over all rows
    assert np.max(np.abs(C - 0.5)) < 1e-3
    rg, rf = analytic.sphere_row_sum_G(k, 1.0), analytic.sphere_row_sum_F(k, 1.0)
    assert np.max(np.abs(G.sum(axis=1) - rg)) / abs(rg) < 5e-3
    assert np.max(np.abs(F.sum(axis=1) - rf)) / abs(rf) < 2e-2
    del F, G
    F0, _, C0 = assemble_parallel(mesh, 0.0, mesh.sources, n=n, m=m, gOn=False)
    assert np.max(np.abs(F0.sum(axis=1) - C0)) < 1e-12


@pytest.mark.parametrize("name,k,n,m", [("lin8", 1.0, 2, 2), ("quad4", 10.0, 3, 2),
                                        ("quad4", 2.0 + 0.4j, 4, 4)])
def test_slot_and_atomic_kernels_agree(name, k, n, m, monkeypatch):
    """Both scatter strategies (LDS slots with owned/shared flush, and all-atomic) give
    the oracle's result."""
    mesh = MESHES[name]()
    Fo, Go, Co = oracle.assemble(mesh, k, n=n, m=m)
    for kern in ("atomic", "slots"):
        monkeypatch.setenv("BIE_AMD_KERNEL", kern)
        print(f"[case] kernel={kern} {name} k={k} (n,m)=({n},{m})", flush=True)
        F, G, C = assemble_parallel(mesh, k, mesh.sources, n=n, m=m)
        assert rel(F, Fo) < TOL and rel(G, Go) < TOL and rel(C, Co) < TOL, kern


def test_device_path_repeated_calls_reuse_caches():
    """Second and later calls on one device mesh (cached rules, schedule, recorded
    mid-edge pairs) agree with the oracle, also after switching rule orders and after a
    call with explicit targets in between (which needs the pair queue again)."""
    mesh = MESHES["quad4"]()
    dm = DeviceMesh(mesh)
    N = mesh.n_sources
    bufs = [DevBuf(16 * N * N), DevBuf(16 * N * N), DevBuf(16 * N)]
    pts = M.fibonacci_sphere(37, 1.02)
    P = pts.shape[1]
    tb = DevBuf(24 * P)
    tb.upload(np.asfortranarray(pts).ravel(order="F"))
    fb = [DevBuf(16 * P * N), DevBuf(16 * P * N), DevBuf(16 * P)]
    outs = []
    for step, nm in enumerate([(3, 2), (4, 4), (3, 2), (3, 2)]):
        if step == 3:
            dm.assemble(6.0, fb[0].ptr, fb[1].ptr, fb[2].ptr, targets_ptr=tb.ptr, n_targets=P,
                        n=3, m=2)
        dm.assemble(6.0, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, n=nm[0], m=nm[1])
        _lib.check(_lib.lib().bie_device_synchronize())
        F = bufs[0].download(np.zeros(N * N, dtype=np.complex128)).reshape(N, N, order="F")
        G = bufs[1].download(np.zeros(N * N, dtype=np.complex128)).reshape(N, N, order="F")
        C = bufs[2].download(np.zeros(N, dtype=np.complex128))
        outs.append((F, G, C))
    ref = {nm: oracle.assemble(mesh, 6.0, n=nm[0], m=nm[1]) for nm in [(3, 2), (4, 4)]}
    for (F, G, C), nm in zip(outs, [(3, 2), (4, 4), (3, 2), (3, 2)]):
        Fo, Go, Co = ref[nm]
        assert rel(F, Fo) < TOL and rel(G, Go) < TOL and rel(C, Co) < TOL
    Ff = fb[0].download(np.zeros(P * N, dtype=np.complex128)).reshape(P, N, order="F")
    Ffo, _, _ = oracle.assemble(mesh, 6.0, targets=pts, n=3, m=2)
    assert rel(Ff, Ffo) < TOL


def test_sources_prefix_as_targets_does_not_record_collocation_pairs():
    """The device sources pointer with fewer targets than sources is not collocation: the
    recorded mid-edge pairs of a later full collocation call must still be complete."""
    mesh = MESHES["quad4"]()
    dm = DeviceMesh(mesh)
    N = mesh.n_sources
    P = N // 3
    fb = [DevBuf(16 * P * N), DevBuf(16 * P * N), DevBuf(16 * P)]
    dm.assemble(4.0, fb[0].ptr, fb[1].ptr, fb[2].ptr, targets_ptr=dm.sources_ptr, n_targets=P,
                n=3, m=2)
    bufs = [DevBuf(16 * N * N), DevBuf(16 * N * N), DevBuf(16 * N)]
    for _ in range(2):   # first call records, second uses the record
        dm.assemble(4.0, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, n=3, m=2)
    _lib.check(_lib.lib().bie_device_synchronize())
    F = bufs[0].download(np.zeros(N * N, dtype=np.complex128)).reshape(N, N, order="F")
    Fp = fb[0].download(np.zeros(P * N, dtype=np.complex128)).reshape(P, N, order="F")
    Fo, _, _ = oracle.assemble(mesh, 4.0, n=3, m=2)
    assert rel(F, Fo) < TOL and rel(Fp, Fo[:P]) < TOL


def test_host_calls_reuse_the_mesh_across_wavenumbers_rules_and_targets():
    """The host-array entry point keeps the last mesh's device form: repeated calls on one
    mesh at other k and (n,m), with field targets in between, then another mesh, then the
    first again (its arrays copied, so a new but identical mesh object), all against the
    oracle."""
    a = M.icosphere(3, "quadratic")
    b = M.icosphere(4, "linear")
    X = 1.8 * np.random.default_rng(2).standard_normal((3, 25))
    seq = [(a, 2.0, None, 3, 2), (a, 5.0 + 0.2j, None, 3, 2), (a, 3.0, None, 4, 4),
           (a, 3.0, X, 3, 2), (a, 1.0, None, 3, 2), (b, 2.0, None, 3, 3),
           (M.Mesh3d(a.coordinates.copy(), a.topology.copy(), a.sources.copy(),
                     a.physics_topology.copy(), a.geometry_kind, a.physics_kind),
            4.0, None, 3, 2)]
    for mesh, k, tg, n, m in seq:
        print(f"[case] host call N={mesh.n_sources} k={k} targets={tg is not None} ({n},{m})",
              flush=True)
        F, G, C = assemble_parallel(mesh, k, mesh.sources if tg is None else tg, n=n, m=m)
        Fo, Go, Co = oracle.assemble(mesh, k, targets=tg, n=n, m=m)
        assert rel(F, Fo) < TOL and rel(G, Go) < TOL and rel(C, Co) < TOL


def test_host_copy_of_large_outputs_is_exact():
    """Outputs of 256 MiB or more come back through pinned double buffers and host threads:
    a 4412-node sphere (311 MB per matrix: two full 128 MiB chunks and a partial one)
    against the same assembly copied back plainly from device buffers.  Two assemblies
    agree to round-off only (shared columns add FP64 atomics from several workgroups in
    scheduling order); a misplaced or missing chunk would be an O(1) error."""
    mesh = M.icosphere(21, "linear")
    N = mesh.n_sources
    print(f"[case] staged host copy N={N}", flush=True)
    F, G, C = assemble_parallel(mesh, 3.0, mesh.sources, n=2, m=2)
    dm = DeviceMesh(mesh)
    dF, dG, dC = DevBuf(16 * N * N), DevBuf(16 * N * N), DevBuf(16 * N)
    dm.assemble(3.0, dF.ptr, dG.ptr, dC.ptr, n=2, m=2)
    _lib.check(_lib.lib().bie_device_synchronize())
    Fd = dF.download(np.empty((N, N), dtype=np.complex128, order="F"))
    Gd = dG.download(np.empty((N, N), dtype=np.complex128, order="F"))
    assert rel(F, Fd) < 1e-14 and rel(G, Gd) < 1e-14
    assert np.max(np.abs(F - Fd)) <= 1e-13 * np.max(np.abs(Fd))
    assert rel(C, dC.download(np.empty(N, dtype=np.complex128))) < 1e-14
