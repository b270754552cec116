"""Solves of the assembled collocation systems on the device, through the C ABI.

The reference solves its rigid-sphere system A p = p_inc, A = F + diag(1 - C)
(examples/3d_rigid_sphere.jl:47-52), by Julia's `\\` (LAPACK getrf/getrs) or with
IterativeSolvers.gmres, and forms field pressures p_inc - F_field p
(examples/3d_rosebem.jl:81).  Here the same steps run where the assembly wrote F:

* ``lu_factor`` / ``lu_solve``: A := F + diag(1 - C) and its LU factorisation in place,
  then solves with A, A^T or A^H (bie_lu_factor_device, bie_lu_solve_device);
* ``zgemv``: y = op(A) x (bie_zgemv_device);
* ``gmres_batched``: restarted GMRES for the systems (F_q + diag(1 - C)) x_q = b_q of a
  frequency sweep at once (bie_gmres_batched_device).

Arguments are raw device pointers (any allocator: the library's bie_device_malloc, or a
framework's tensors) to column-major complex128 arrays, and an optional HIP stream handle.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

TRANS = {"N": 0, "T": 1, "C": 2}


def _p(x):
    return ctypes.c_void_p(int(x) if x else None)


def lu_factor(n, A_ptr, lda, ipiv_ptr, *, C_ptr=0, stream=0):
    """A := F + diag(1 - C) (no free term when C_ptr is 0) and its LU factors in place; the
    n int32 pivots go to ipiv_ptr.  Waits for the factorisation; a singular matrix raises
    ValueError."""
    rc = _lib.lib().bie_lu_factor_device(int(n), _p(A_ptr), int(lda), _p(C_ptr), _p(ipiv_ptr),
                                         _p(stream))
    if rc == _lib.BIE_EINVAL:
        raise ValueError(_lib.lib().bie_last_error().decode())
    _lib.check(rc)


def lu_solve(n, A_ptr, lda, ipiv_ptr, B_ptr, ldb, nrhs=1, *, trans="N", stream=0):
    """B := op(A)^-1 B with the factors of lu_factor (op: "N", "T" or "C").  Asynchronous."""
    _lib.check(_lib.lib().bie_lu_solve_device(int(n), _p(A_ptr), int(lda), _p(ipiv_ptr),
                                              TRANS[trans], _p(B_ptr), int(ldb), int(nrhs),
                                              _p(stream)))


def zgemv(m, n, A_ptr, lda, x_ptr, y_ptr, *, trans="N", stream=0):
    """y := op(A) x for A m x n (op: "N", "T" or "C").  Asynchronous."""
    _lib.check(_lib.lib().bie_zgemv_device(TRANS[trans], int(m), int(n), _p(A_ptr), int(lda),
                                           _p(x_ptr), _p(y_ptr), _p(stream)))


def gmres_batched(n, n_systems, F_ptr, strideF, ldf, C_ptr, B_ptr, ldb, X_ptr, ldx, *,
                  tol=1e-10, restart=60, maxiter=2000, stream=0):
    """(F_q + diag(1 - C)) x_q = b_q for q < n_systems: F_q at F_ptr + 16 q strideF bytes,
    b_q and x_q at + 16 q ldb / ldx bytes (x0 = 0).  Restarted GMRES (CGS2 Arnoldi, Givens
    rotations) with all systems in lockstep.  Returns (iterations, relative residuals) as
    numpy arrays; waits for the solves."""
    its = np.zeros(int(n_systems), dtype=np.int32)
    res = np.zeros(int(n_systems), dtype=np.float64)
    _lib.check(_lib.lib().bie_gmres_batched_device(
        int(n), int(n_systems), _p(F_ptr), int(strideF), int(ldf), _p(C_ptr), _p(B_ptr),
        int(ldb), _p(X_ptr), int(ldx), float(tol), int(restart), int(maxiter), _lib._ptr(its),
        _lib._ptr(res), _p(stream)))
    return its, res
