"""Allocation-free apply (bie_zgemm_device_noalloc) through the C ABI against the apply
that allocates its workspace (bie_zgemm_device): two calls in a row, the second one
smaller (the shared scratch array keeps the first call's packed X, which must not leak
into the second), and the size limit of the scratch array."""
import ctypes

import numpy as np
import pytest

from bie_amd import _lib
from gpu_helpers import DevBuf, sync

pytestmark = pytest.mark.gpu


def _apply(name, A, X, M, nrhs, K):
    L = _lib.lib()
    Y = DevBuf(16 * M * nrhs)
    rc = getattr(L, name)(M, nrhs, K, ctypes.c_void_p(A.ptr), M, ctypes.c_void_p(X.ptr), K,
                          ctypes.c_void_p(Y.ptr), M, None)
    _lib.check(rc)
    sync()
    return Y.download(np.zeros((M, nrhs), dtype=np.complex128, order="F"))


def test_noalloc_apply_matches_the_allocating_apply_on_two_calls():
    rng = np.random.default_rng(11)
    for M, K, nrhs in [(3000, 2600, 40), (700, 900, 7)]:     # the second call is smaller
        print(f"[case] noalloc apply M={M} K={K} nrhs={nrhs}", flush=True)
        A = np.asfortranarray(rng.standard_normal((M, K)) + 1j * rng.standard_normal((M, K)))
        X = np.asfortranarray(rng.standard_normal((K, nrhs)) + 1j * rng.standard_normal((K, nrhs)))
        dA, dX = DevBuf(A.nbytes), DevBuf(X.nbytes)
        dA.upload(A.ravel(order="F"))
        dX.upload(X.ravel(order="F"))
        Y0 = _apply("bie_zgemm_device_noalloc", dA, dX, M, nrhs, K)
        Y1 = _apply("bie_zgemm_device", dA, dX, M, nrhs, K)
        ref = A @ X
        assert np.linalg.norm(Y0 - Y1) <= 1e-14 * np.linalg.norm(Y1)
        assert np.linalg.norm(Y0 - ref) <= 1e-13 * np.linalg.norm(ref)


def test_noalloc_apply_rejects_sizes_beyond_the_scratch_array():
    L = _lib.lib()
    A, X, Y = DevBuf(16), DevBuf(16), DevBuf(16)
    # K * ceil(nrhs/32)*32 = 40000 * 32 > 2^20 complex entries
    rc = L.bie_zgemm_device_noalloc(1, 32, 40000, ctypes.c_void_p(A.ptr), 1,
                                    ctypes.c_void_p(X.ptr), 40000, ctypes.c_void_p(Y.ptr), 1,
                                    None)
    assert rc == _lib.BIE_EINVAL
    assert b"scratch" in L.bie_last_error()
