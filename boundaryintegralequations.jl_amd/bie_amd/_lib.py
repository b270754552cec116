"""ctypes binding of libbie_amd.so (the C ABI in include/bie_amd.h).

Loading fails loudly when the HIP library has not been built: there is no CPU
fallback on the product path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import mesh as _mesh

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbie_amd.so")

BIE_OK = 0
BIE_EINVAL = 1
BIE_EUNSUPPORTED = 2
BIE_ENOMEM = 3
BIE_EHIP = 4
BIE_F, BIE_G, BIE_C = 1, 2, 4

_dp = ctypes.c_void_p


class BieMesh(ctypes.Structure):
    _fields_ = [("coordinates", _dp), ("topology", _dp), ("sources", _dp),
                ("physics_topology", _dp), ("n_coordinates", ctypes.c_int64),
                ("n_elements", ctypes.c_int64), ("n_sources", ctypes.c_int64),
                ("geometry_kind", ctypes.c_int32), ("physics_kind", ctypes.c_int32),
                ("index_base", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("physics_beta", ctypes.c_double)]

# bie_mesh.flags: which n-ary hypot rule selection compares distances with
BIE_MESH_SCALED_HYPOT = 1
HYPOT_FLAGS = {"julia1.9": 0, "scaled": BIE_MESH_SCALED_HYPOT}


def hypot_flag(hypot: str) -> int:
    """bie_mesh.flags for the hypot formula of rule selection: "julia1.9" (the reference's
    CI Julia: maxabs * sqrt(sum((x/maxabs)^2))) or "scaled" (later Julia releases:
    power-of-two scaled sqrt(x^2 + y^2 + z^2))."""
    if hypot not in HYPOT_FLAGS:
        raise ValueError(f"hypot must be one of {sorted(HYPOT_FLAGS)}")
    return HYPOT_FLAGS[hypot]


class BieError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bie_amd error {code}: {msg}")
        self.code = code


# name -> (restype, argtypes)
SIGNATURES = {
    "bie_last_error": (ctypes.c_char_p, []),
    "bie_version": (ctypes.c_int, []),
    "bie_assemble": (ctypes.c_int, [ctypes.POINTER(BieMesh), ctypes.c_double, ctypes.c_double,
                                    _dp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                    ctypes.c_uint32, _dp, _dp, _dp]),
    "bie_dmesh_create": (ctypes.c_int, [ctypes.POINTER(BieMesh), ctypes.c_int,
                                        ctypes.POINTER(_dp)]),
    "bie_dmesh_destroy": (ctypes.c_int, [_dp]),
    "bie_dmesh_n_sources": (ctypes.c_int64, [_dp]),
    "bie_dmesh_n_elements": (ctypes.c_int64, [_dp]),
    "bie_dmesh_sources": (_dp, [_dp]),
    "bie_assemble_device": (ctypes.c_int, [_dp, ctypes.c_double, ctypes.c_double, _dp,
                                           ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_uint32, _dp, _dp, _dp, _dp]),
    "bie_select_rules_device": (ctypes.c_int, [_dp, _dp, ctypes.c_int64, _dp, _dp]),
    "bie_element_geometry": (ctypes.c_int, [_dp, ctypes.c_int32, ctypes.c_int32,
                                            ctypes.c_int32, _dp, _dp, _dp]),
    "bie_schedule_check": (ctypes.c_int, [ctypes.POINTER(BieMesh), ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_int64)]),
    "bie_assembly_vjp_device": (ctypes.c_int, [_dp, ctypes.c_double, _dp, ctypes.c_int64, _dp,
                                                _dp, _dp, _dp, ctypes.c_int32, ctypes.c_int32,
                                                _dp, _dp, _dp]),
    "bie_assembly_vjp_complex_device": (ctypes.c_int, [_dp, ctypes.c_double, ctypes.c_double,
                                                        _dp, ctypes.c_int64, _dp, _dp, _dp, _dp,
                                                        ctypes.c_int32, ctypes.c_int32, _dp, _dp,
                                                        _dp]),
    "bie_zgemv_batched_device": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, _dp,
                                                 ctypes.c_int64, ctypes.c_int64, _dp,
                                                 ctypes.c_int64, _dp, ctypes.c_int64, _dp,
                                                 ctypes.c_int32, _dp]),
    "bie_zgemm_device": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _dp,
                                         ctypes.c_int64, _dp, ctypes.c_int64, _dp,
                                         ctypes.c_int64, _dp]),
    "bie_zgemm_device_noalloc": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                                 _dp, ctypes.c_int64, _dp, ctypes.c_int64, _dp,
                                                 ctypes.c_int64, _dp]),
    "bie_host_cache_info": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int64)] * 4),
    "bie_lu_factor_device": (ctypes.c_int, [ctypes.c_int64, _dp, ctypes.c_int64, _dp, _dp, _dp]),
    "bie_lu_solve_device": (ctypes.c_int, [ctypes.c_int64, _dp, ctypes.c_int64, _dp,
                                            ctypes.c_int32, _dp, ctypes.c_int64, ctypes.c_int64,
                                            _dp]),
    "bie_zgemv_device": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, _dp,
                                         ctypes.c_int64, _dp, _dp, _dp]),
    "bie_gmres_batched_device": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, _dp,
                                                 ctypes.c_int64, ctypes.c_int64, _dp, _dp,
                                                 ctypes.c_int64, _dp, ctypes.c_int64,
                                                 ctypes.c_double, ctypes.c_int32,
                                                 ctypes.c_int32, _dp, _dp, _dp]),
    "bie_host_cache_release": (ctypes.c_int, []),
    "bie_assemble_sweep_device": (ctypes.c_int, [_dp, _dp, ctypes.c_int32, _dp, ctypes.c_int64,
                                                  ctypes.c_int32, ctypes.c_int32,
                                                  ctypes.c_uint32, _dp, _dp, ctypes.c_int64,
                                                  _dp, _dp]),
    "bie_sweep_tune": (ctypes.c_int, [_dp, _dp, ctypes.c_int32, _dp, ctypes.c_int64,
                                      ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32, _dp, _dp,
                                      ctypes.c_int64, _dp, _dp,
                                      ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_double)]),
    "bie_sweep_graph_create": (ctypes.c_int, [_dp, _dp, ctypes.c_int32, _dp, ctypes.c_int64,
                                              ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32,
                                              _dp, _dp, ctypes.c_int64, _dp,
                                              ctypes.POINTER(_dp)]),
    "bie_assemble_taylor_device": (ctypes.c_int, [_dp, ctypes.c_double, ctypes.c_double,
                                                   ctypes.c_int32, _dp, ctypes.c_int64,
                                                   ctypes.c_int32, ctypes.c_int32,
                                                   ctypes.c_uint32, _dp, _dp, ctypes.c_int64,
                                                   _dp, _dp]),
    "bie_taylor_expand_device": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, _dp,
                                                 ctypes.c_int64, _dp, _dp, _dp]),
    "bie_graph_launch": (ctypes.c_int, [_dp, _dp]),
    "bie_graph_destroy": (ctypes.c_int, [_dp]),
    "bie_batch_create": (ctypes.c_int, [ctypes.POINTER(_dp), ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_int32, ctypes.POINTER(_dp)]),
    "bie_batch_destroy": (ctypes.c_int, [_dp]),
    "bie_batch_configure": (ctypes.c_int, [_dp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "bie_batch_info": (ctypes.c_int, [_dp, ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_int64)]),
    "bie_batch_assemble": (ctypes.c_int, [_dp, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double), ctypes.c_uint32,
                                          ctypes.POINTER(_dp), ctypes.POINTER(_dp),
                                          ctypes.POINTER(_dp), _dp]),
    "bie_batch_tune": (ctypes.c_int, [_dp, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double), ctypes.c_uint32,
                                      ctypes.POINTER(_dp), ctypes.POINTER(_dp),
                                      ctypes.POINTER(_dp), _dp,
                                      ctypes.POINTER(ctypes.c_double)]),
    "bie_batch_graph_create": (ctypes.c_int, [_dp, ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_double), ctypes.c_uint32,
                                              ctypes.POINTER(_dp), ctypes.POINTER(_dp),
                                              ctypes.POINTER(_dp), ctypes.POINTER(_dp)]),
    "bie_batch_dfdk_device": (ctypes.c_int, [_dp, ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(_dp), ctypes.POINTER(_dp), _dp,
                                             _dp]),
    "bie_sparse_count": (ctypes.c_int, [ctypes.POINTER(BieMesh), ctypes.c_int32,
                                        ctypes.POINTER(ctypes.c_int64)]),
    "bie_sparse_assemble": (ctypes.c_int, [ctypes.POINTER(BieMesh), ctypes.c_double,
                                           ctypes.c_double, _dp, ctypes.c_int64,
                                           ctypes.c_int32, ctypes.c_double, ctypes.c_uint32,
                                           ctypes.c_int64, _dp, _dp, _dp, _dp]),
    "bie_sparse_assemble_ex": (ctypes.c_int, [ctypes.POINTER(BieMesh), ctypes.c_double,
                                              ctypes.c_double, _dp, ctypes.c_int64,
                                              ctypes.c_int32, ctypes.c_double, ctypes.c_int32,
                                              ctypes.c_uint32, ctypes.c_int64, _dp, _dp, _dp,
                                              _dp]),
    "bie_partial_assemble": (ctypes.c_int, [ctypes.POINTER(BieMesh), ctypes.c_double,
                                            ctypes.c_double, _dp, ctypes.c_int64,
                                            ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32,
                                            ctypes.c_int64, _dp, _dp, _dp, _dp]),
    "bie_sparse_create": (ctypes.c_int, [_dp, ctypes.c_int32, ctypes.c_double,
                                         ctypes.POINTER(_dp)]),
    "bie_sparse_create_ex": (ctypes.c_int, [_dp, ctypes.c_int32, ctypes.c_double,
                                            ctypes.c_int32, ctypes.POINTER(_dp)]),
    "bie_sparse_destroy": (ctypes.c_int, [_dp]),
    "bie_sparse_nnz": (ctypes.c_int64, [_dp]),
    "bie_sparse_pattern": (ctypes.c_int, [_dp, ctypes.c_int32, _dp, _dp]),
    "bie_sparse_assemble_device": (ctypes.c_int, [_dp, ctypes.c_double, ctypes.c_double, _dp,
                                                   ctypes.c_uint32, _dp, _dp, _dp]),
    "bie_sparse_partial_device": (ctypes.c_int, [_dp, ctypes.c_double, ctypes.c_double,
                                                  ctypes.c_int32, _dp, ctypes.c_uint32, _dp,
                                                  _dp, _dp]),
    "bie_timing_enable": (ctypes.c_int, [ctypes.c_int]),
    "bie_timing_collect": (ctypes.c_int, [ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_int64)]),
    "bie_device_malloc": (ctypes.c_int, [ctypes.POINTER(_dp), ctypes.c_int64]),
    "bie_device_free": (ctypes.c_int, [_dp]),
    "bie_memcpy_h2d": (ctypes.c_int, [_dp, _dp, ctypes.c_int64]),
    "bie_memcpy_d2h": (ctypes.c_int, [_dp, _dp, ctypes.c_int64]),
    "bie_device_synchronize": (ctypes.c_int, []),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        # PyTorch (when installed) is imported before the library is loaded: with the
        # library loaded first, PyTorch's later device initialisation in the same process
        # fails ("No HIP GPUs are available"); in this order both work.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != BIE_OK:
        msg = lib().bie_last_error()
        raise BieError(rc, msg.decode() if msg else "")


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class DeviceBuffer:
    """Device buffer from the library's allocator (bie_device_malloc), optionally filled
    from a host array; freed with the object."""

    def __init__(self, arr: np.ndarray | None = None, nbytes: int = 0):
        if arr is not None:
            arr = np.ascontiguousarray(arr)
            nbytes = arr.nbytes
        p = ctypes.c_void_p()
        check(lib().bie_device_malloc(ctypes.byref(p), int(max(nbytes, 8))))
        self.ptr = p.value
        self.nbytes = nbytes
        if arr is not None and arr.nbytes:
            check(lib().bie_memcpy_h2d(ctypes.c_void_p(self.ptr), _ptr(arr), arr.nbytes))

    @classmethod
    def empty(cls, nbytes: int) -> "DeviceBuffer":
        return cls(None, nbytes)

    def get(self, out: np.ndarray) -> np.ndarray:
        check(lib().bie_memcpy_d2h(_ptr(out), ctypes.c_void_p(self.ptr), out.nbytes))
        return out

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().bie_device_free(ctypes.c_void_p(self.ptr))


class HostMesh:
    """Keeps Fortran-ordered copies alive while a bie_mesh struct points at them."""

    def __init__(self, m: "_mesh.Mesh3d", flags: int = 0):
        self.coords = np.asfortranarray(m.coordinates, dtype=np.float64)
        self.topo = np.asfortranarray(m.topology, dtype=np.int64)
        self.sources = np.asfortranarray(m.sources, dtype=np.float64)
        self.ptopo = np.asfortranarray(m.physics_topology, dtype=np.int64)
        self.struct = BieMesh(_ptr(self.coords), _ptr(self.topo), _ptr(self.sources),
                              _ptr(self.ptopo), m.n_coordinates, m.n_elements, m.n_sources,
                              m.geometry_kind, m.physics_kind, 0, flags,
                              float(getattr(m, "physics_beta", 0.0)))
