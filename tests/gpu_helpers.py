"""Device buffers over the C ABI's allocation helpers, for GPU tests."""
import ctypes

import numpy as np

from bie_amd import _lib


class DevBuf:
    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        _lib.check(_lib.lib().bie_device_malloc(ctypes.byref(p), int(max(nbytes, 1))))
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


def sync():
    _lib.check(_lib.lib().bie_device_synchronize())
