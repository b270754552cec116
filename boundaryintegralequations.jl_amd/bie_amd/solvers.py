"""Restarted GMRES on the GPU for the dense collocation systems the assembly produces.

The reference solves its rigid-sphere system A p = p_inc, A = F + diag(1 - C)
(examples/3d_rigid_sphere.jl:47-52), with a direct or Krylov solver from the Julia
ecosystem (IterativeSolvers.gmres in its examples and tests).  This is the same
iteration: Arnoldi with classical Gram-Schmidt applied twice (CGS2), Givens
rotations for the small least-squares problem, restart after `restart` steps, stopping at
||b - A x|| <= tol ||b||.  Device tensors are torch complex128; the operator is any
callable x -> A x on them, so the assembled matrix stays where the assembly wrote it.
"""
from __future__ import annotations

import math


def collocation_operator(torch, F_storage, C):
    """x -> (F + diag(1 - C)) x for F stored column-major N x N (the library's layout):
    F_storage is the (N, N) torch view of that memory, i.e. F transposed."""
    d = 1.0 - C

    def apply(x):
        return torch.mv(F_storage.t(), x) + d * x

    return apply


def gmres(torch, A, b, *, tol=1e-10, restart=60, maxiter=2000, x0=None):
    """Returns (x, iterations, relative residual ||b - A x|| / ||b||)."""
    n = b.shape[0]
    x = torch.zeros_like(b) if x0 is None else x0.clone()
    bnorm = float(torch.linalg.vector_norm(b))
    if bnorm == 0.0:
        return x, 0, 0.0
    it = 0
    r = b - A(x) if x0 is not None else b.clone()
    beta = float(torch.linalg.vector_norm(r))
    while it < maxiter and beta > tol * bnorm:
        m = min(restart, maxiter - it, n)
        V = torch.empty((m + 1, n), dtype=b.dtype, device=b.device)
        V[0] = r / beta
        H = [[0j] * m for _ in range(m + 1)]
        cs, sn = [0.0] * m, [0j] * m
        g = [0j] * (m + 1)
        g[0] = complex(beta)
        k_used = 0
        for j in range(m):
            w = A(V[j])
            # classical Gram-Schmidt with one reorthogonalisation pass, both on the device;
            # the coefficients and the norm come to the host in one transfer per step
            h = torch.mv(V[: j + 1].conj(), w)
            w = w - torch.mv(V[: j + 1].t(), h)
            h2 = torch.mv(V[: j + 1].conj(), w)
            w = w - torch.mv(V[: j + 1].t(), h2)
            hv = torch.cat([h + h2, torch.linalg.vector_norm(w).reshape(1).to(h.dtype)])
            hv = hv.cpu().tolist()
            for i in range(j + 1):
                H[i][j] += hv[i]
            hn = float(hv[j + 1].real)
            H[j + 1][j] = complex(hn)
            if hn > 0.0:
                V[j + 1] = w / hn
            # apply previous rotations to column j, then form the new one
            for i in range(j):
                t = cs[i] * H[i][j] + sn[i] * H[i + 1][j]
                H[i + 1][j] = -sn[i].conjugate() * H[i][j] + cs[i] * H[i + 1][j]
                H[i][j] = t
            a, bb = H[j][j], H[j + 1][j]
            den = math.sqrt(abs(a) ** 2 + abs(bb) ** 2)
            if den == 0.0:
                cs[j], sn[j] = 1.0, 0j
            else:
                cs[j] = abs(a) / den
                phase = a / abs(a) if abs(a) > 0 else 1.0
                sn[j] = phase * bb.conjugate() / den
            H[j][j] = cs[j] * a + sn[j] * bb
            H[j + 1][j] = 0j
            g[j + 1] = -sn[j].conjugate() * g[j]
            g[j] = cs[j] * g[j]
            it += 1
            k_used = j + 1
            if abs(g[j + 1]) <= tol * bnorm or hn == 0.0:
                break
        # back substitution for y, then x += V^T y
        y = [0j] * k_used
        for i in range(k_used - 1, -1, -1):
            s = g[i] - sum(H[i][l] * y[l] for l in range(i + 1, k_used))
            y[i] = s / H[i][i]
        yt = torch.tensor(y, dtype=b.dtype, device=b.device)
        x = x + torch.mv(V[:k_used].t(), yt)
        r = b - A(x)
        beta = float(torch.linalg.vector_norm(r))
    return x, it, beta / bnorm
