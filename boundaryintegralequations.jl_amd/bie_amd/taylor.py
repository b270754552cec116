"""Wavenumber derivatives of the dense operators, and their Taylor expansion in k, on the GPU.

``taylor_assemble(mesh, k, in_sources; M=0, fOn, gOn, cOn, m=3, n=3, progress, U)`` mirrors
``taylor_assemble!`` (src/3d/rosebem.jl:56-157): same keywords and defaults, returning
``(F, G, C)`` with ``F[:, :, q]`` = d^q F / dk^q at k for q = 0..M (complex128, column-major,
the derivative order last, as the reference lays them out), all orders from one device call
(``bie_assemble_taylor_device``).  With a basis U (N x r) the derivatives are reduced to
U^H F_q U as in the reference (rosebem.jl:141-151), with two device ZGEMMs per order.

``apply_taylor_expansion(Abasis, k, k0[, bbasis])`` mirrors rosebem.jl:164-199: the sum
A(k) = sum_i (k - k0)^i / i! A^{(i)}, one streaming pass on the device
(``bie_taylor_expand_device``).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from .assembly import DeviceMesh, _flags
from .mesh import Mesh3d
from .operators import _Dev, zgemm_device

MAX_ORDERS = 64


def taylor_coefficients(k, k0, n_terms: int) -> np.ndarray:
    """(k - k0)^i / i! for i = 0 .. n_terms-1 (the weights of rosebem.jl:170-177)."""
    dk = complex(k) - complex(k0)
    return np.array([dk ** i / math.factorial(i) for i in range(int(n_terms))],
                    dtype=np.complex128)


def taylor_expand_device(n_terms, n, A_ptr, strideA, coef, Y_ptr, stream=0):
    """Asynchronous Y = sum_t coef[t] A_t over device arrays of n complex entries, A_t at
    A_ptr + 16 * t * strideA bytes; coef is a host sequence of n_terms complex values."""
    c = np.ascontiguousarray(np.asarray(coef, dtype=np.complex128).reshape(-1))
    if c.size != int(n_terms):
        raise ValueError("need one coefficient per term")
    _lib.check(_lib.lib().bie_taylor_expand_device(
        int(n_terms), int(n), ctypes.c_void_p(A_ptr), int(strideA), _lib._ptr(c),
        ctypes.c_void_p(Y_ptr), ctypes.c_void_p(stream or None)))


def assemble_taylor_device(dm: DeviceMesh, k, n_orders, F_ptr, G_ptr, stride, C_ptr, *,
                           targets_ptr=0, n_targets=None, n=3, m=3, fOn=True, gOn=True,
                           cOn=True, stream=0):
    """Asynchronous d^q F/dk^q, d^q G/dk^q (q < n_orders) into caller-owned device buffers:
    order q at F_ptr + 16 * q * stride bytes."""
    k = complex(k)
    Mt = dm.n_sources if n_targets is None else int(n_targets)
    _lib.check(_lib.lib().bie_assemble_taylor_device(
        dm.handle, k.real, k.imag, int(n_orders), ctypes.c_void_p(targets_ptr or None), Mt,
        int(n), int(m), _flags(fOn, gOn, cOn), ctypes.c_void_p(F_ptr), ctypes.c_void_p(G_ptr),
        int(stride), ctypes.c_void_p(C_ptr), ctypes.c_void_p(stream or None)))


def _reduce(L, dX, nt, N, n_orders, U, dU, dUH, dT):
    """U^H X_q U for every order q of the device stack dX (nt x N each), back on the host."""
    r = U.shape[1]
    out = np.empty((r, r, n_orders), dtype=np.complex128, order="F")
    dO = _Dev(16 * r * r * n_orders)
    for q in range(n_orders):
        zgemm_device(nt, r, N, dX.ptr + 16 * q * nt * N, max(nt, 1), dU.ptr, max(N, 1),
                     dT.ptr, max(nt, 1))
        zgemm_device(r, r, nt, dUH.ptr, max(r, 1), dT.ptr, max(nt, 1), dO.ptr + 16 * q * r * r,
                     max(r, 1))
    _lib.check(L.bie_device_synchronize())
    _lib.check(L.bie_memcpy_d2h(_lib._ptr(out), ctypes.c_void_p(dO.ptr), out.nbytes))
    return out


def taylor_assemble(mesh: Mesh3d, k, in_sources=None, *, M=0, fOn=True, gOn=True, cOn=True,
                    m=3, n=3, progress=True, U=None, device=0):
    """(F, G, C): F[:, :, q] and G[:, :, q] the q-th k-derivatives (q = 0..M) of the dense
    operators at k, C the integral free term (drop-in for taylor_assemble!)."""
    del progress
    n_orders = int(M) + 1
    if not 1 <= n_orders <= MAX_ORDERS:
        raise ValueError(f"M must be in 0..{MAX_ORDERS - 1}")
    if in_sources is None or in_sources is mesh.sources:
        targets = None
        nt = mesh.n_sources
    else:
        targets = np.asfortranarray(in_sources, dtype=np.float64)
        if targets.ndim != 2 or targets.shape[0] != 3:
            raise ValueError("in_sources must be a 3 x M array")
        nt = targets.shape[1]
    N = mesh.n_sources
    if U is not None:
        U = np.asfortranarray(U, dtype=np.complex128)
        if U.ndim != 2 or U.shape[0] != N or nt != N:
            raise ValueError("U must be n_sources x r, with the collocation targets")
    L = _lib.lib()
    dm = DeviceMesh(mesh, device)
    # disabled outputs get no device buffer; they are returned as zeros (rosebem.jl:98-100)
    dF = _Dev(16 * nt * N * n_orders) if fOn else None
    dG = _Dev(16 * nt * N * n_orders) if gOn else None
    dC = _Dev(16 * nt)
    dX = None
    if targets is not None:
        dX = _Dev(targets.nbytes)
        _lib.check(L.bie_memcpy_h2d(ctypes.c_void_p(dX.ptr), _lib._ptr(targets), targets.nbytes))
    assemble_taylor_device(dm, k, n_orders, dF.ptr if dF else 0, dG.ptr if dG else 0,
                           nt * N, dC.ptr,
                           targets_ptr=dX.ptr if dX else 0, n_targets=nt, n=n, m=m, fOn=fOn,
                           gOn=gOn, cOn=cOn)
    _lib.check(L.bie_device_synchronize())
    C = np.empty(nt, dtype=np.complex128)
    _lib.check(L.bie_memcpy_d2h(_lib._ptr(C), ctypes.c_void_p(dC.ptr), C.nbytes))
    if U is not None:
        r = U.shape[1]
        UH = np.asfortranarray(U.conj().T)
        dU, dUH, dT = _Dev(U.nbytes), _Dev(UH.nbytes), _Dev(16 * nt * r)
        _lib.check(L.bie_memcpy_h2d(ctypes.c_void_p(dU.ptr), _lib._ptr(U), U.nbytes))
        _lib.check(L.bie_memcpy_h2d(ctypes.c_void_p(dUH.ptr), _lib._ptr(UH), UH.nbytes))
        zero = np.zeros((r, r, n_orders), dtype=np.complex128, order="F")
        return (_reduce(L, dF, nt, N, n_orders, U, dU, dUH, dT) if dF else zero,
                _reduce(L, dG, nt, N, n_orders, U, dU, dUH, dT) if dG else zero.copy(), C)
    out = []
    for d in (dF, dG):
        X = np.zeros((nt, N, n_orders), dtype=np.complex128, order="F")
        if d:
            _lib.check(L.bie_memcpy_d2h(_lib._ptr(X), ctypes.c_void_p(d.ptr), X.nbytes))
        out.append(X)
    return out[0], out[1], C


def _expand_host(basis, k, k0):
    basis = np.asfortranarray(basis, dtype=np.complex128)
    n_terms = basis.shape[-1]
    if not 1 <= n_terms <= MAX_ORDERS:
        raise ValueError(f"between 1 and {MAX_ORDERS} terms")
    n = basis.size // n_terms
    L = _lib.lib()
    dA, dY = _Dev(basis.nbytes), _Dev(16 * n)
    _lib.check(L.bie_memcpy_h2d(ctypes.c_void_p(dA.ptr), _lib._ptr(basis), basis.nbytes))
    taylor_expand_device(n_terms, n, dA.ptr, n, taylor_coefficients(k, k0, n_terms), dY.ptr)
    _lib.check(L.bie_device_synchronize())
    out = np.empty(basis.shape[:-1], dtype=np.complex128, order="F")
    _lib.check(L.bie_memcpy_d2h(_lib._ptr(out), ctypes.c_void_p(dY.ptr), out.nbytes))
    return out


def apply_taylor_expansion(Abasis, k, k0, bbasis=None):
    """A(k) = sum_i (k - k0)^i / i! Abasis[..., i] (and the same for bbasis when given:
    returns (A, b)), as rosebem.jl:164-199."""
    A = _expand_host(Abasis, k, k0)
    if bbasis is None:
        return A
    return A, _expand_host(bbasis, k, k0)
