"""Dense operator apply on the GPU (bie_zgemm_device): Y = A X for complex128
column-major matrices, the step after assembly in the reference's workflows
(examples/3d_pulsating_sphere.jl:42,62)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def zgemm_device(M, nrhs, K, A_ptr, lda, X_ptr, ldx, Y_ptr, ldy, stream=0):
    """Asynchronous Y = A X on device pointers (complex128, column-major)."""
    _lib.check(_lib.lib().bie_zgemm_device(
        int(M), int(nrhs), int(K), ctypes.c_void_p(A_ptr), int(lda), ctypes.c_void_p(X_ptr),
        int(ldx), ctypes.c_void_p(Y_ptr), int(ldy), ctypes.c_void_p(stream or None)))


def zgemv_batched_device(M, K, A_ptr, strideA, lda, X_ptr, ldx, Y_ptr, ldy, idx_ptr, n_idx,
                         stream=0):
    """Asynchronous y_q += A_q x_q for q in the device int32 array idx (complex128; strides
    and leading dimensions in complex entries)."""
    _lib.check(_lib.lib().bie_zgemv_batched_device(
        int(M), int(K), ctypes.c_void_p(A_ptr), int(strideA), int(lda), ctypes.c_void_p(X_ptr),
        int(ldx), ctypes.c_void_p(Y_ptr), int(ldy), ctypes.c_void_p(idx_ptr), int(n_idx),
        ctypes.c_void_p(stream or None)))


class _Dev:
    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        _lib.check(_lib.lib().bie_device_malloc(ctypes.byref(p), int(max(nbytes, 1))))
        self.ptr = p.value

    def __del__(self):
        if getattr(self, "ptr", None):
            _lib.lib().bie_device_free(ctypes.c_void_p(self.ptr))


def apply(A: np.ndarray, X: np.ndarray) -> np.ndarray:
    """Host convenience: copies A (M x K) and X (K x nrhs) to the GPU and returns A @ X."""
    A = np.asfortranarray(A, dtype=np.complex128)
    X = np.asfortranarray(X, dtype=np.complex128)
    if X.ndim == 1:
        X = X.reshape(-1, 1, order="F")
    M, K = A.shape
    if X.shape[0] != K:
        raise ValueError("inner dimensions differ")
    nrhs = X.shape[1]
    dA, dX, dY = _Dev(A.nbytes), _Dev(X.nbytes), _Dev(16 * M * nrhs)
    L = _lib.lib()
    _lib.check(L.bie_memcpy_h2d(ctypes.c_void_p(dA.ptr), _lib._ptr(A), A.nbytes))
    _lib.check(L.bie_memcpy_h2d(ctypes.c_void_p(dX.ptr), _lib._ptr(X), X.nbytes))
    zgemm_device(M, nrhs, K, dA.ptr, max(M, 1), dX.ptr, max(K, 1), dY.ptr, max(M, 1))
    _lib.check(L.bie_device_synchronize())
    Y = np.empty((M, nrhs), dtype=np.complex128, order="F")
    _lib.check(L.bie_memcpy_d2h(_lib._ptr(Y), ctypes.c_void_p(dY.ptr), Y.nbytes))
    return Y
