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
from .assembly import DeviceMesh
from .mesh import Mesh3d


_Dev = _lib.DeviceBuffer


def source_nodes(mesh: Mesh3d) -> np.ndarray:
    """Index into mesh.coordinates of every collocation node (they move with the mesh)."""
    lut = {tuple(c): j for j, c in enumerate(mesh.coordinates.T)}
    out = np.array([lut.get(tuple(s), -1) for s in mesh.sources.T], dtype=np.int32)
    if (out < 0).any():
        raise ValueError("collocation nodes are not mesh coordinates")
    return out


def assembly_vjp(dmesh: DeviceMesh, k, alpha, p, *, beta=None, targets=None,
                 target_node=None, n=4, m=4):
    """Gradient of Re[alpha^T F p + beta^T C] w.r.t. mesh coordinates (3 x Nc) and k.

    Real k: returns (grad_X, d/dk).  Complex k (the lossy systems' wavenumbers): returns
    (grad_X, (d/dRe k, d/dIm k))."""
    mesh = dmesh.mesh
    kc = complex(k)
    cplx = kc.imag != 0.0
    M = mesh.n_sources if targets is None else targets.shape[1]
    dA = _Dev(np.asarray(alpha, dtype=np.complex128))
    dP = _Dev(np.asarray(p, dtype=np.complex128))
    dB = _Dev(np.asarray(beta, dtype=np.complex128)) if beta is not None else None
    dT = _Dev(np.asfortranarray(targets, dtype=np.float64).ravel(order="F")) if targets is not None else None
    dN = _Dev(np.asarray(target_node, dtype=np.int32)) if target_node is not None else None
    gX = _Dev(np.zeros(3 * mesh.n_coordinates))
    gk = _Dev(np.zeros(2 if cplx else 1))
    common = (ctypes.c_void_p(dT.ptr if dT else None), int(M),
              ctypes.c_void_p(dN.ptr if dN else None), ctypes.c_void_p(dA.ptr),
              ctypes.c_void_p(dB.ptr if dB else None), ctypes.c_void_p(dP.ptr), int(n), int(m),
              ctypes.c_void_p(gX.ptr), ctypes.c_void_p(gk.ptr), None)
    if cplx:
        _lib.check(_lib.lib().bie_assembly_vjp_complex_device(dmesh.handle, kc.real, kc.imag,
                                                              *common))
    else:
        _lib.check(_lib.lib().bie_assembly_vjp_device(dmesh.handle, kc.real, *common))
    _lib.check(_lib.lib().bie_device_synchronize())
    g = gX.get(np.zeros(3 * mesh.n_coordinates)).reshape(mesh.n_coordinates, 3).T.copy()
    if cplx:
        v = gk.get(np.zeros(2))
        return g, (float(v[0]), float(v[1]))
    return g, float(gk.get(np.zeros(1))[0])


def plane_wave(x, k, direction=(0.0, 0.0, 1.0)):
    d = np.asarray(direction, dtype=np.float64)
    return np.exp(1j * k * (d @ x))


class _Forward:
    """Device state of one objective evaluation: A = F + diag(1 - C) factored in place, the
    field operator F_field, p and the small host vectors."""


def _forward(dm: DeviceMesh, k: float, field_points, n, m, direction, stream=0):
    """J and the forward quantities: the surface and field operators assembled into device
    buffers, A p = b by LU on the device (A overwrites F), u = p_inc - F_field p; only
    vectors of N and P entries come to the host."""
    from . import solvers
    mesh = dm.mesh
    pts = np.asfortranarray(field_points, dtype=np.float64)
    N, P = mesh.n_sources, pts.shape[1]
    fw = _Forward()
    fw.dF = _Dev.empty(16 * N * N)               # F, then A and its LU factors
    fw.dFf = _Dev.empty(16 * P * N)              # F_field, P x N column-major
    dG = _Dev.empty(16 * N * max(N, P))          # G is not needed (zero-filled)
    dC = _Dev.empty(16 * N)
    dCf = _Dev.empty(16 * P)
    fw.dpts = _Dev(pts.ravel(order="F"))
    dm.assemble(k, fw.dF.ptr, dG.ptr, dC.ptr, n=n, m=m, gOn=False, stream=stream)
    dm.assemble(k, fw.dFf.ptr, dG.ptr, dCf.ptr, targets_ptr=fw.dpts.ptr, n_targets=P, n=n,
                m=m, gOn=False, cOn=False, stream=stream)
    del dG
    fw.ipiv = _Dev.empty(4 * N)
    solvers.lu_factor(N, fw.dF.ptr, N, fw.ipiv.ptr, C_ptr=dC.ptr, stream=stream)
    fw.b = plane_wave(mesh.sources, k, direction)
    fw.pf = plane_wave(pts, k, direction)
    dp = _Dev(fw.b)
    solvers.lu_solve(N, fw.dF.ptr, N, fw.ipiv.ptr, dp.ptr, N, 1, stream=stream)
    dfp = _Dev.empty(16 * P)
    solvers.zgemv(P, N, fw.dFf.ptr, P, dp.ptr, dfp.ptr, stream=stream)
    _lib.check(_lib.lib().bie_device_synchronize())
    fw.p = dp.get(np.empty(N, dtype=np.complex128))
    fw.u = fw.pf - dfp.get(np.empty(P, dtype=np.complex128))
    fw.dp = dp
    fw.pts = pts
    J = float(np.sum(fw.u.real ** 2 + fw.u.imag ** 2))
    return J, fw


def scattering_objective(mesh: Mesh3d, k: float, field_points, *, n=4, m=4,
                         direction=(0.0, 0.0, 1.0), dmesh: DeviceMesh | None = None,
                         forward=True):
    """J and the forward quantities (p, u, b as numpy arrays), computed on the GPU; with
    forward=False only J."""
    dm = dmesh if dmesh is not None else DeviceMesh(mesh)
    J, fw = _forward(dm, float(k), field_points, n, m, direction)
    if not forward:
        return J, None
    return J, {"p": fw.p, "u": fw.u, "b": fw.b}


def scattering_gradient(mesh: Mesh3d, k: float, field_points, *, n=4, m=4,
                        direction=(0.0, 0.0, 1.0), dmesh: DeviceMesh | None = None):
    """J, dJ/d(mesh coordinates) (3 x Nc numpy array) and dJ/dk by the adjoint method.  The
    operators, their LU factors, both solves and both vector-Jacobian products stay on the
    GPU; vectors of N and P entries, J, the gradient and dJ/dk come back."""
    from . import solvers
    dm = dmesh if dmesh is not None else DeviceMesh(mesh)
    k = float(k)
    J, fw = _forward(dm, k, field_points, n, m, direction)
    N, P = mesh.n_sources, fw.pts.shape[1]
    # lambda = A^-H F_field^H u on the device
    du = _Dev(fw.u)
    dlam = _Dev.empty(16 * N)
    solvers.zgemv(P, N, fw.dFf.ptr, P, du.ptr, dlam.ptr, trans="C")
    solvers.lu_solve(N, fw.dF.ptr, N, fw.ipiv.ptr, dlam.ptr, N, 1, trans="C")
    _lib.check(_lib.lib().bie_device_synchronize())
    lam = dlam.get(np.empty(N, dtype=np.complex128))
    lam_c = np.conj(lam)
    p, u, b, pf = fw.p, fw.u, fw.b, fw.pf
    nodes = source_nodes(mesh)
    gs, gks = assembly_vjp(dm, k, lam_c, p, beta=-lam_c * p, target_node=nodes, n=n, m=m)
    gf, gkf = assembly_vjp(dm, k, -np.conj(u), p, targets=fw.pts, n=n, m=m)
    grad = 2.0 * (gs + gf)
    dk = 2.0 * (gks + gkf)
    d = np.asarray(direction, dtype=np.float64)
    # -lambda^H db with b_i = e^{ik d.x_i}: d b_i/d x_i = i k d b_i, d b_i/dk = i (d.x_i) b_i
    coef = (-lam_c * 1j * k * b).real
    np.add.at(grad, (slice(None), nodes), 2.0 * d[:, None] * coef[None, :])
    dk += 2.0 * (-lam_c * 1j * (d @ mesh.sources) * b).sum().real
    # u^H dp_inc,field/dk
    dk += 2.0 * (np.conj(u) * 1j * (d @ fw.pts) * pf).sum().real
    return J, grad, float(dk)
