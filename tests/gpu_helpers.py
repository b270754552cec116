"""Device buffers over the C ABI's allocation helpers, and comparison helpers, for GPU
tests."""
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


TOL = 1e-12


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def own_blocks(mesh):
    """Mask of the (row, column) entries of an element's own nodes."""
    own = np.zeros((mesh.n_sources, mesh.n_sources), dtype=bool)
    for e in range(mesh.n_elements):
        idx = mesh.physics_topology[:, e]
        own[np.ix_(idx, idx)] = True
    return own


def check_F(F, Fo, mesh, tag=""):
    """Collocation F where a node may lie inside or on an element without the Duffy collapse
    at it (discontinuous physics, whose nodes sit inside the element away from the corner its
    rule is clustered at; quadrilaterals, whose one Gauss rule has no clustering).  On the
    element's own nodes the integrand then carries the rounding noise of n.d amplified by
    1/r^3 at the rule points nearest the node: on flat elements n.d is 0 exactly there and F
    is pure noise (~1e-12 of max |F|); on curved ones the noise sits on top of the true
    value.  Those entries are held to 1e-9 of max |F|, every other entry to the bar."""
    own = own_blocks(mesh)
    fmax = np.max(np.abs(Fo))
    err_own = np.max(np.abs(F[own] - Fo[own])) / fmax
    print(f"  {tag} F off own blocks rel {rel(F[~own], Fo[~own]):.2e}, own blocks max "
          f"{err_own:.2e} of max|F|", flush=True)
    assert rel(F[~own], Fo[~own]) <= TOL
    assert err_own <= 1e-9


