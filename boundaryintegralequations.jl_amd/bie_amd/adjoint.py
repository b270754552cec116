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


def _forward(torch, dm: DeviceMesh, k: float, field_points, n, m, direction):
    """J and the forward quantities, on the device of dm: the surface and field operators
    assembled into device buffers, A p = b and the field pressure by device LU solves."""
    mesh = dm.mesh
    dev = torch.device("cuda", dm.device)
    stream = torch.cuda.current_stream(dev).cuda_stream
    pts = np.asarray(field_points, dtype=np.float64)
    N, P = mesh.n_sources, pts.shape[1]
    cplx = torch.complex128
    Fm = torch.empty((N, N), dtype=cplx, device=dev)        # column-major F: Fm = F^T
    Ffm = torch.empty((N, P), dtype=cplx, device=dev)       # column-major F_field
    Gs = torch.empty(N * max(N, P), dtype=cplx, device=dev)  # G is not needed (zero-filled)
    C = torch.empty(N, dtype=cplx, device=dev)
    Cf = torch.empty(P, dtype=cplx, device=dev)
    tp = torch.as_tensor(np.asfortranarray(pts).ravel(order="F"), device=dev)
    dm.assemble(k, Fm.data_ptr(), Gs.data_ptr(), C.data_ptr(), n=n, m=m, gOn=False,
                stream=stream)
    dm.assemble(k, Ffm.data_ptr(), Gs.data_ptr(), Cf.data_ptr(), targets_ptr=tp.data_ptr(),
                n_targets=P, n=n, m=m, gOn=False, cOn=False, stream=stream)
    del Gs
    d = torch.as_tensor(np.asarray(direction, dtype=np.float64), device=dev)
    xs = torch.as_tensor(np.ascontiguousarray(mesh.sources), device=dev)
    xf = torch.as_tensor(np.ascontiguousarray(pts), device=dev)
    b = torch.exp(1j * k * (d @ xs))
    pf = torch.exp(1j * k * (d @ xf))
    A = Fm.T + torch.diag(1.0 - C)
    p = torch.linalg.solve(A, b)
    Ff = Ffm.T
    u = pf - Ff @ p
    J = float(torch.sum(u.real ** 2 + u.imag ** 2))
    return J, dict(A=A, Ff=Ff, p=p, u=u, b=b, pf=pf, d=d, xs=xs, xf=xf, tp=tp,
                   stream=stream, dev=dev)


def scattering_objective(mesh: Mesh3d, k: float, field_points, *, n=4, m=4,
                         direction=(0.0, 0.0, 1.0), dmesh: DeviceMesh | None = None,
                         forward=True):
    """J and the forward quantities (p, u, A, F_field as numpy arrays), computed on the GPU;
    with forward=False only J (nothing else leaves the device)."""
    import torch
    dm = dmesh if dmesh is not None else DeviceMesh(mesh)
    J, fw = _forward(torch, dm, float(k), field_points, n, m, direction)
    if not forward:
        return J, None
    return J, {key: fw[key].cpu().numpy() for key in ("A", "Ff", "p", "u", "b")}


def _vjp_device(torch, dm: DeviceMesh, k, alpha, p, beta, targets_ptr, n_targets, target_node,
                n, m, stream, dev):
    """assembly_vjp on device tensors (complex128 alpha, p, beta; int32 target_node)."""
    gX = torch.zeros(3 * dm.mesh.n_coordinates, dtype=torch.float64, device=dev)
    gk = torch.zeros(1, dtype=torch.float64, device=dev)
    # the library reads raw memory: materialise lazy conjugate / negative views first
    def raw(t):
        return t.resolve_conj().resolve_neg().contiguous()

    alpha, p = raw(alpha), raw(p)
    beta = raw(beta) if beta is not None else None
    _lib.check(_lib.lib().bie_assembly_vjp_device(
        dm.handle, float(k), ctypes.c_void_p(targets_ptr or None), int(n_targets),
        ctypes.c_void_p(target_node.data_ptr() if target_node is not None else None),
        ctypes.c_void_p(alpha.data_ptr()),
        ctypes.c_void_p(beta.data_ptr() if beta is not None else None),
        ctypes.c_void_p(p.data_ptr()), int(n), int(m), ctypes.c_void_p(gX.data_ptr()),
        ctypes.c_void_p(gk.data_ptr()), ctypes.c_void_p(stream or None)))
    return gX.view(dm.mesh.n_coordinates, 3).T, gk[0]


def scattering_gradient(mesh: Mesh3d, k: float, field_points, *, n=4, m=4,
                        direction=(0.0, 0.0, 1.0), dmesh: DeviceMesh | None = None):
    """J, dJ/d(mesh coordinates) (3 x Nc numpy array) and dJ/dk by the adjoint method.  The
    operators, both solves and both vector-Jacobian products stay on the GPU; only J, the
    gradient and dJ/dk come back."""
    import torch
    dm = dmesh if dmesh is not None else DeviceMesh(mesh)
    k = float(k)
    J, fw = _forward(torch, dm, k, field_points, n, m, direction)
    A, Ff, p, u, b = fw["A"], fw["Ff"], fw["p"], fw["u"], fw["b"]
    stream, dev = fw["stream"], fw["dev"]
    lam = torch.linalg.solve(A.conj().T, Ff.conj().T @ u)
    nodes = torch.as_tensor(source_nodes(mesh), device=dev)
    P = u.shape[0]
    lam_c = torch.conj_physical(lam)
    gs, gks = _vjp_device(torch, dm, k, lam_c, p, -lam_c * p, 0, mesh.n_sources, nodes, n, m,
                          stream, dev)
    gf, gkf = _vjp_device(torch, dm, k, -torch.conj_physical(u), p, None, fw["tp"].data_ptr(),
                          P, None, n, m, stream, dev)
    grad = 2.0 * (gs + gf)
    dk = 2.0 * (gks + gkf)
    d, xs, xf, pf = fw["d"], fw["xs"], fw["xf"], fw["pf"]
    # -lambda^H db with b_i = e^{ik d.x_i}: d b_i/d x_i = i k d b_i, d b_i/dk = i (d.x_i) b_i
    coef = (-lam_c * 1j * k * b).real
    grad = grad.index_add(1, nodes.long(), 2.0 * d[:, None] * coef[None, :])
    dk = dk + 2.0 * (-lam_c * 1j * (d @ xs) * b).sum().real
    # u^H dp_inc,field/dk
    dk = dk + 2.0 * (u.conj() * 1j * (d @ xf) * pf).sum().real
    return J, grad.cpu().numpy(), float(dk)
