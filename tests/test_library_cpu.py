"""C-ABI library checks that need no GPU: it loads, exports every declared symbol, and
rejects malformed input on the host before any HIP call (the reference's
ArgumentError / missing-method paths)."""
import ctypes
import os
import re

import numpy as np
import pytest

from bie_amd import _lib
from bie_amd import mesh as M
from bie_amd.assembly import assemble_parallel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(bie_[a-z0-9_]+)\s*\(",
                                text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    # and the Python binding declares a signature for each of them
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_version():
    assert _lib.lib().bie_version() >= 1


def _struct(mesh, **over):
    hm = _lib.HostMesh(mesh)
    for k, v in over.items():
        setattr(hm.struct, k, v)
    return hm


def _call(hm, M_=None):
    N = hm.struct.n_sources
    M_ = N if M_ is None else M_
    F = np.zeros((M_, N), dtype=np.complex128)
    G = np.zeros_like(F)
    C = np.zeros(M_, dtype=np.complex128)
    return _lib.lib().bie_assemble(ctypes.byref(hm.struct), 1.0, 0.0, None, M_, 2, 2, 7,
                                   _lib._ptr(F), _lib._ptr(G), _lib._ptr(C))


def test_rejects_unsupported_element_kinds():
    mesh = M.icosphere(1, "linear")
    assert _call(_struct(mesh, geometry_kind=4)) == _lib.BIE_EUNSUPPORTED
    assert _call(_struct(mesh, physics_kind=6)) == _lib.BIE_EUNSUPPORTED  # quad physics, linear geometry
    assert b"TRI" in _lib.lib().bie_last_error() or b"reference" in _lib.lib().bie_last_error()


def test_rejects_out_of_range_topology_and_null():
    mesh = M.icosphere(1, "linear")
    bad = M.icosphere(1, "linear")
    bad.topology[0, 3] = bad.n_coordinates + 5
    assert _call(_lib.HostMesh(bad)) == _lib.BIE_EINVAL
    assert b"out of range" in _lib.lib().bie_last_error()
    assert _call(_struct(mesh, index_base=2)) == _lib.BIE_EINVAL
    assert _call(_struct(mesh, coordinates=None)) == _lib.BIE_EINVAL
    rc = _lib.lib().bie_assemble(None, 1.0, 0.0, None, 0, 2, 2, 7, None, None, None)
    assert rc == _lib.BIE_EINVAL


def test_one_based_indices_are_validated_too():
    mesh = M.icosphere(1, "linear")
    hm = _lib.HostMesh(mesh)
    hm.struct.index_base = 1   # 0-based data read as 1-based -> index 0 is out of range
    assert _call(hm) == _lib.BIE_EINVAL


def test_python_mirror_argument_errors():
    mesh = M.icosphere(1, "linear")
    # checked before any HIP call
    with pytest.raises(ValueError):
        assemble_parallel(mesh, 1.0, sparse=True, depth=3)
    with pytest.raises(ValueError):
        assemble_parallel(mesh, 1.0, mesh.sources[:, :5], sparse=True)
    with pytest.raises(ValueError):
        assemble_parallel(mesh, 1.0, np.zeros((2, 5)))


@pytest.mark.parametrize("depth", [1, 2])
def test_sparse_pattern_count_matches_oracle(depth):
    # bie_sparse_count is host code (connected_sources), no GPU needed
    import oracle
    for mesh in (M.icosphere(2, "linear"), M.icosphere(1, "quadratic")):
        hm = _lib.HostMesh(mesh)
        nnz = ctypes.c_int64()
        _lib.check(_lib.lib().bie_sparse_count(ctypes.byref(hm.struct), depth,
                                               ctypes.byref(nnz)))
        _, cols = oracle.connected_sources(mesh, depth)
        assert nnz.value == sum(len(c) for c in cols)
    assert _lib.lib().bie_sparse_count(ctypes.byref(hm.struct), 3,
                                       ctypes.byref(nnz)) == _lib.BIE_EINVAL


def test_apply_rejects_leading_dimension_beyond_buffer_offsets():
    # checked on the host before any HIP call: the apply kernel addresses a K split with
    # 32-bit buffer offsets, so lda is limited to ~1.6e6 rows (the pointers are never used)
    p = ctypes.c_void_p(0x1000)
    lib = _lib.lib()
    rc = lib.bie_zgemm_device(10, 1, 4, p, 2_000_000, p, 4, p, 2_000_000, None)
    assert rc == _lib.BIE_EINVAL
    assert b"leading dimension too large" in lib.bie_last_error()


def test_taylor_entry_points_validate_arguments_on_the_host():
    # every check below runs before any HIP call (the pointers are never used)
    p = ctypes.c_void_p(0x1000)
    lib = _lib.lib()
    coef = (ctypes.c_double * 4)(1.0, 0.0, 0.5, 0.0)
    assert lib.bie_taylor_expand_device(0, 10, p, 10, coef, p, None) == _lib.BIE_EINVAL
    assert lib.bie_taylor_expand_device(65, 10, p, 10, coef, p, None) == _lib.BIE_EINVAL
    assert lib.bie_taylor_expand_device(2, 10, p, 9, coef, p, None) == _lib.BIE_EINVAL
    assert b"stride" in lib.bie_last_error()
    assert lib.bie_taylor_expand_device(2, 10, p, 10, None, p, None) == _lib.BIE_EINVAL
    # the output may be the first term but no other (fake addresses: rejected before use)
    a0 = 0x100000
    assert lib.bie_taylor_expand_device(2, 10, ctypes.c_void_p(a0), 10, coef,
                                        ctypes.c_void_p(a0 + 160), None) == _lib.BIE_EINVAL
    assert b"overlaps" in lib.bie_last_error()
    assert lib.bie_taylor_expand_device(2, 10, ctypes.c_void_p(a0), 10, coef,
                                        ctypes.c_void_p(a0 + 80), None) == _lib.BIE_EINVAL
    bad = (ctypes.c_double * 4)(1.0, 0.0, float("nan"), 0.0)
    assert lib.bie_taylor_expand_device(2, 10, p, 10, bad, p, None) == _lib.BIE_EINVAL
    assert lib.bie_assemble_taylor_device(None, 1.0, 0.0, 3, None, 4, 3, 3, 7, p, p, 16, p,
                                          None) == _lib.BIE_EINVAL
    from bie_amd.taylor import taylor_assemble, taylor_coefficients
    with pytest.raises(ValueError):
        taylor_assemble(M.icosphere(1, "linear"), 1.0, M=64)
    c = taylor_coefficients(1.5 + 0.5j, 1.0, 25)
    dk = 0.5 + 0.5j
    assert c[0] == 1.0 and abs(c[1] - dk) < 1e-16
    assert abs(c[24] - dk ** 24 / 620448401733239439360000) < 1e-16 * abs(c[24])


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(ImportError):
        _lib.lib()


@pytest.mark.parametrize("name", ["c1", "c2", "c3", "c4", "sphere_1m", "quadlin", "ellipsoid"])
@pytest.mark.parametrize("min_segments", [1, 13, 64, 300])
def test_slot_schedule_invariants(name, min_segments):
    """The LDS-slot kernel's schedule: every element swept once, slots never double-held,
    every touched node flushed exactly once per segment, shared/owned flags consistent."""
    mesh = {"c1": lambda: M.icosphere(8, "linear"),
            "c2": lambda: M.icosphere(16, "quadratic"),
            "c3": lambda: M.icosphere(20, "linear"),
            "c4": lambda: M.icosphere(16, "linear"),
            "sphere_1m": lambda: M.load_comsol(os.path.join(ROOT, "tests", "golden", "sphere_1m")),
            "quadlin": lambda: M.icosphere(6, "quadratic", physics="linear"),
            "ellipsoid": lambda: M.ellipsoid(9, (0.5, 1.0, 2.0), "linear")}[name]()
    hm = _lib.HostMesh(mesh)
    st = (ctypes.c_int64 * 6)()
    rc = _lib.lib().bie_schedule_check(ctypes.byref(hm.struct), min_segments, st)
    assert rc == 0, _lib.lib().bie_last_error()
    n_seg, n_steps, max_live, n_shared, n_zero, n_flush = list(st)
    max_size = -(-mesh.n_elements // min_segments)   # segments hold at most this many elements
    assert n_seg >= -(-mesh.n_elements // max_size)
    assert max_live <= 64
    assert n_zero == n_shared            # closed surfaces: every node is touched
    assert n_flush >= mesh.n_sources     # each node flushed at least once
