"""Torch restatement of the library's restarted GMRES and dense solve (test reference).

The library solves the rigid-sphere systems A p = p_inc, A = F + diag(1 - C)
(examples/3d_rigid_sphere.jl:47-52), natively behind its C ABI (bie_gmres_batched_device,
bie_lu_factor_device / bie_lu_solve_device).  This module states the same iteration
independently with torch tensors - Arnoldi with classical Gram-Schmidt applied twice
(CGS2), Givens rotations for the small least-squares problem, restart after `restart`
steps, stop at ||b - A x|| <= tol ||b|| - so the tests can check the native solver's
iteration counts and solutions against it (on CPU tensors without a GPU, against direct
solves).
"""
from __future__ import annotations

import math


def _upper_solve(R, rhs):
    """y with R y = rhs for the upper-triangular Hessenberg factor R (k x k)."""
    import numpy as np
    from scipy.linalg import solve_triangular
    return solve_triangular(np.asarray(R, dtype=np.complex128),
                            np.asarray(rhs, dtype=np.complex128), lower=False)


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
        # back substitution for y (upper triangular k_used x k_used), then x += V^T y
        y = _upper_solve([row[:k_used] for row in H[:k_used]], g[:k_used])
        yt = torch.as_tensor(y, dtype=b.dtype, device=b.device)
        x = x + torch.mv(V[:k_used].t(), yt)
        r = b - A(x)
        beta = float(torch.linalg.vector_norm(r))
    return x, it, beta / bnorm


def lu_solve(torch, F_storage, C, B, *, overwrite=False):
    """X with (F + diag(1 - C)) X = B by LU on the device (torch.linalg.solve, i.e.
    rocSOLVER getrf/getrs): the direct solve of the rigid-sphere system
    (examples/3d_rigid_sphere.jl:50-52, test/3d/assembly_test.jl:37) with F left where the
    assembly wrote it.  F_storage is the (N, N) torch view of the column-major F (F
    transposed), C the length-N free term, B (N,) or (N, nrhs).  With overwrite=True the
    diagonal term is added into F_storage in place (no second N x N buffer; F is then A)."""
    d = 1.0 - C
    if overwrite:
        F_storage.diagonal().add_(d)
        A = F_storage.t()
    else:
        A = F_storage.t() + torch.diag(d)
    return torch.linalg.solve(A, B)


def collocation_operator_batched(torch, F_storage, C):
    """(X, active) -> Y with Y[q] = (F_q + diag(1 - C)) X[q] for the systems q marked in the
    boolean numpy array `active` (the others get only the diagonal term, which
    gmres_batched does not use): F_storage[q] is the (N, N) torch view of the column-major
    F_q, as the batched sweep writes them."""
    import numpy as np
    from bie_amd.operators import zgemv_batched_device
    d = 1.0 - C
    nb, n, _ = F_storage.shape
    assert F_storage.is_contiguous() and F_storage.dtype == torch.complex128

    def apply(X, active):
        assert X.stride(-1) == 1 and X.dtype == torch.complex128
        Y = d * X                                   # contiguous (nb, N)
        idx = np.flatnonzero(active).astype(np.int32)
        if idx.size:
            if F_storage.is_cuda:                   # one batched GEMV launch on the device
                di = torch.as_tensor(idx, device=X.device)
                zgemv_batched_device(n, n, F_storage.data_ptr(), n * n, n, X.data_ptr(),
                                     X.stride(0), Y.data_ptr(), n, di.data_ptr(), idx.size,
                                     torch.cuda.current_stream(X.device).cuda_stream)
            else:                                   # host tensors (tests): per-system mv
                for q in idx.tolist():
                    Y[q] += torch.mv(F_storage[q].t(), X[q])
        return Y

    return apply


def gmres_batched(torch, A, B, *, tol=1e-10, restart=60, maxiter=2000):
    """Restarted GMRES for nb independent systems A_q x_q = b_q advanced in lockstep: the
    same iteration as `gmres` per system (CGS2 Arnoldi, Givens rotations, restart, stop at
    ||b_q - A_q x_q|| <= tol ||b_q||), with the Gram-Schmidt products batched over the
    systems and one host transfer per Arnoldi step for all of them; a system that has
    converged takes no further operator applications.  A(X, active) applies the operators to
    the rows of X (nb, N) marked in the numpy mask `active`.  Returns (X, iterations per
    system, relative residual per system)."""
    import numpy as np
    nb, n = B.shape
    dev, dt = B.device, B.dtype
    X = torch.zeros_like(B)
    bnorm = torch.linalg.vector_norm(B, dim=1).cpu().numpy()
    its = np.zeros(nb, dtype=np.int64)
    R = B.clone()
    beta = bnorm.copy()
    done = beta <= tol * bnorm
    total = 0
    while not done.all() and total < maxiter:
        m = min(restart, maxiter - total, n)
        V = torch.zeros((nb, m + 1, n), dtype=dt, device=dev)
        inv0 = np.where(done, 0.0, 1.0 / np.where(beta > 0, beta, 1.0))
        V[:, 0] = R * torch.as_tensor(inv0, dtype=R.real.dtype, device=dev)[:, None]
        H = np.zeros((nb, m + 1, m), dtype=np.complex128)
        cs = np.zeros((nb, m))
        sn = np.zeros((nb, m), dtype=np.complex128)
        g = np.zeros((nb, m + 1), dtype=np.complex128)
        g[:, 0] = np.where(done, 0.0, beta)
        active = ~done
        k_used = np.zeros(nb, dtype=np.int64)
        steps = 0
        for j in range(m):
            W = A(V[:, j], active)
            Vj = V[:, : j + 1]
            # h = Vj^H w = conj(sum_n Vj[:, :, n] conj(w[n])): an elementwise product and a
            # reduction stream Vj at full bandwidth (a batched GEMM with one column does not)
            h1 = (Vj * W.conj().unsqueeze(1)).sum(-1).conj()
            W = W - torch.bmm(h1.unsqueeze(1), Vj).squeeze(1)
            h2 = (Vj * W.conj().unsqueeze(1)).sum(-1).conj()
            W = W - torch.bmm(h2.unsqueeze(1), Vj).squeeze(1)
            hn = torch.linalg.vector_norm(W, dim=1)
            hv = torch.cat([h1 + h2, hn.unsqueeze(1).to(dt)], dim=1).cpu().numpy()
            hnr = hv[:, j + 1].real
            live = active & (hnr > 0)
            inv = np.where(live, 1.0 / np.where(hnr > 0, hnr, 1.0), 0.0)
            V[:, j + 1] = W * torch.as_tensor(inv, dtype=W.real.dtype, device=dev)[:, None]
            H[:, : j + 1, j] = hv[:, : j + 1]
            H[:, j + 1, j] = hnr
            for i in range(j):   # previous rotations on column j, all systems at once
                t = cs[:, i] * H[:, i, j] + sn[:, i] * H[:, i + 1, j]
                H[:, i + 1, j] = -np.conj(sn[:, i]) * H[:, i, j] + cs[:, i] * H[:, i + 1, j]
                H[:, i, j] = t
            a, bb = H[:, j, j].copy(), H[:, j + 1, j].copy()
            den = np.sqrt(np.abs(a) ** 2 + np.abs(bb) ** 2)
            nz = den > 0
            dsafe = np.where(nz, den, 1.0)
            aabs = np.abs(a)
            phase = np.where(aabs > 0, a / np.where(aabs > 0, aabs, 1.0), 1.0)
            cs[:, j] = np.where(nz, aabs / dsafe, 1.0)
            sn[:, j] = np.where(nz, phase * np.conj(bb) / dsafe, 0.0)
            H[:, j, j] = cs[:, j] * a + sn[:, j] * bb
            H[:, j + 1, j] = 0.0
            g[:, j + 1] = -np.conj(sn[:, j]) * g[:, j]
            g[:, j] = cs[:, j] * g[:, j]
            its[active] += 1
            k_used[active] = j + 1
            steps = j + 1
            active = active & ~((np.abs(g[:, j + 1]) <= tol * bnorm) | (hnr == 0.0))
            if not active.any():
                break
        total += steps
        # back substitution per system, then X += V^T y for all at once
        Y = np.zeros((nb, m + 1), dtype=np.complex128)
        for q in range(nb):
            kq = int(k_used[q])
            if kq:
                Y[q, :kq] = _upper_solve(H[q, :kq, :kq], g[q, :kq])
        X = X + torch.bmm(torch.as_tensor(Y, device=dev).unsqueeze(1), V).squeeze(1)
        R = B - A(X, np.ones(nb, dtype=bool))
        beta = torch.linalg.vector_norm(R, dim=1).cpu().numpy()
        done = beta <= tol * bnorm
    return X, its, beta / np.where(bnorm > 0, bnorm, 1.0)
