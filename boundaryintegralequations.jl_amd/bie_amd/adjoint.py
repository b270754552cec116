"""Adjoint gradient of a rigid-sphere scattering objective through the dense operators.

Objective (config 4 of the build's benchmarks): incident plane wave p_inc = e^{i k d.x}
on a rigid (sound-hard) body, surface pressure from the reference's exterior formulation
A p = p_inc with A = F + diag(1 - C) (test/3d/assembly_test.jl:37,
examples/3d_rigid_sphere.jl:50), field pressure u = p_inc - F_field p at exterior points
(the field convention of examples/3d_rosebem.jl:81), and

    J = sum_m |u_m|^2.

With lambda solving A^H lambda = F_field^H u,

    dJ = 2 Re[ u^H dp_inc,field - u^H (dF_field) p - lambda^H db + lambda^H (dA) p ],

where the operator terms come from two calls of the assembly's vector-Jacobian product
(bie_assembly_vjp_device: surface rows with moving targets, field rows with fixed
targets) and the plane-wave terms are explicit.  Rule selection is held fixed, as the
derivative of a piecewise-constant choice is zero almost everywhere.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .assembly import DeviceMesh, assemble_parallel
from .mesh import Mesh3d


class _Dev:
    def __init__(self, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        p = ctypes.c_void_p()
        _lib.check(_lib.lib().bie_device_malloc(ctypes.byref(p), int(max(arr.nbytes, 8))))
        self.ptr = p.value
        self.nbytes = arr.nbytes
        if arr.nbytes:
            _lib.check(_lib.lib().bie_memcpy_h2d(ctypes.c_void_p(self.ptr), _lib._ptr(arr),
                                                 arr.nbytes))

    def get(self, out: np.ndarray) -> np.ndarray:
        _lib.check(_lib.lib().bie_memcpy_d2h(_lib._ptr(out), ctypes.c_void_p(self.ptr), out.nbytes))
        return out

    def __del__(self):
        if getattr(self, "ptr", None):
            _lib.lib().bie_device_free(ctypes.c_void_p(self.ptr))


def source_nodes(mesh: Mesh3d) -> np.ndarray:
    """Index into mesh.coordinates of every collocation node (they move with the mesh)."""
    lut = {tuple(c): j for j, c in enumerate(mesh.coordinates.T)}
    out = np.array([lut.get(tuple(s), -1) for s in mesh.sources.T], dtype=np.int32)
    if (out < 0).any():
        raise ValueError("collocation nodes are not mesh coordinates")
    return out


def assembly_vjp(dmesh: DeviceMesh, k: float, alpha, p, *, beta=None, targets=None,
                 target_node=None, n=4, m=4):
    """Gradient of Re[alpha^T F p + beta^T C] w.r.t. mesh coordinates (3 x Nc) and k."""
    mesh = dmesh.mesh
    M = mesh.n_sources if targets is None else targets.shape[1]
    dA = _Dev(np.asarray(alpha, dtype=np.complex128))
    dP = _Dev(np.asarray(p, dtype=np.complex128))
    dB = _Dev(np.asarray(beta, dtype=np.complex128)) if beta is not None else None
    dT = _Dev(np.asfortranarray(targets, dtype=np.float64).ravel(order="F")) if targets is not None else None
    dN = _Dev(np.asarray(target_node, dtype=np.int32)) if target_node is not None else None
    gX = _Dev(np.zeros(3 * mesh.n_coordinates))
    gk = _Dev(np.zeros(1))
    _lib.check(_lib.lib().bie_assembly_vjp_device(
        dmesh.handle, float(k), ctypes.c_void_p(dT.ptr if dT else None), int(M),
        ctypes.c_void_p(dN.ptr if dN else None), ctypes.c_void_p(dA.ptr),
        ctypes.c_void_p(dB.ptr if dB else None), ctypes.c_void_p(dP.ptr), int(n), int(m),
        ctypes.c_void_p(gX.ptr), ctypes.c_void_p(gk.ptr), None))
    _lib.check(_lib.lib().bie_device_synchronize())
    g = gX.get(np.zeros(3 * mesh.n_coordinates)).reshape(mesh.n_coordinates, 3).T.copy()
    return g, float(gk.get(np.zeros(1))[0])


def plane_wave(x, k, direction=(0.0, 0.0, 1.0)):
    d = np.asarray(direction, dtype=np.float64)
    return np.exp(1j * k * (d @ x))


def scattering_objective(mesh: Mesh3d, k: float, field_points, *, n=4, m=4,
                         direction=(0.0, 0.0, 1.0)):
    """J and the forward quantities (p, u, A, F_field) on the GPU-assembled operators."""
    F, _, C = assemble_parallel(mesh, k, mesh.sources, n=n, m=m, gOn=False)
    Ff, _, _ = assemble_parallel(mesh, k, field_points, n=n, m=m, gOn=False, cOn=False)
    A = F + np.diag(1.0 - C)
    b = plane_wave(mesh.sources, k, direction)
    p = np.linalg.solve(A, b)
    u = plane_wave(field_points, k, direction) - Ff @ p
    return float(np.real(np.vdot(u, u))), dict(A=A, Ff=Ff, p=p, u=u, b=b)


def scattering_gradient(mesh: Mesh3d, k: float, field_points, *, n=4, m=4,
                        direction=(0.0, 0.0, 1.0), dmesh: DeviceMesh | None = None):
    """J, dJ/d(mesh coordinates) (3 x Nc) and dJ/dk by the adjoint method."""
    J, fw = scattering_objective(mesh, k, field_points, n=n, m=m, direction=direction)
    A, Ff, p, u, b = fw["A"], fw["Ff"], fw["p"], fw["u"], fw["b"]
    lam = np.linalg.solve(A.conj().T, Ff.conj().T @ u)
    dm = dmesh if dmesh is not None else DeviceMesh(mesh)
    gs, gks = assembly_vjp(dm, k, np.conj(lam), p, beta=-np.conj(lam) * p,
                           target_node=source_nodes(mesh), n=n, m=m)
    gf, gkf = assembly_vjp(dm, k, -np.conj(u), p, targets=np.asarray(field_points), n=n, m=m)
    grad = 2.0 * (gs + gf)
    dk = 2.0 * (gks + gkf)
    d = np.asarray(direction, dtype=np.float64)
    # -lambda^H db with b_i = e^{ik d.x_i}: d b_i/d x_i = i k d b_i, d b_i/dk = i (d.x_i) b_i
    nodes = source_nodes(mesh)
    coef = np.real(-np.conj(lam) * 1j * k * b)
    for c in range(3):
        np.add.at(grad[c], nodes, 2.0 * d[c] * coef)
    dk += 2.0 * np.real(np.sum(-np.conj(lam) * 1j * (d @ mesh.sources) * b))
    # u^H dp_inc,field/dk
    pf = plane_wave(field_points, k, direction)
    dk += 2.0 * np.real(np.sum(np.conj(u) * 1j * (d @ field_points) * pf))
    return J, grad, dk
