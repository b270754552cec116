"""The torch restatement of the library's restarted GMRES (tests/torch_solvers.py, the
reference the GPU tests hold the native bie_gmres_batched_device to) against direct solves,
on CPU tensors."""
import numpy as np
import torch

from torch_solvers import (collocation_operator, collocation_operator_batched, gmres,
                           gmres_batched)


def test_gmres_matches_direct_solve_with_restarts():
    rng = np.random.default_rng(0)
    n = 120
    A = np.eye(n) * 3 + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    At = torch.as_tensor(A)
    bt = torch.as_tensor(b)
    for restart in (5, 30, 200):
        x, it, rr = gmres(torch, lambda v: At @ v, bt, tol=1e-12, restart=restart, maxiter=5000)
        assert rr <= 1e-12
        xd = np.linalg.solve(A, b)
        assert np.linalg.norm(x.numpy() - xd) <= 1e-10 * np.linalg.norm(xd)
        assert it >= 1


def test_collocation_operator_reads_column_major_storage():
    rng = np.random.default_rng(1)
    n = 17
    F = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    C = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    storage = torch.as_tensor(np.asfortranarray(F).ravel(order="F").reshape(n, n))  # F^T view
    op = collocation_operator(torch, storage, torch.as_tensor(C))
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = (F + np.diag(1.0 - C)) @ x
    assert np.allclose(op(torch.as_tensor(x)).numpy(), ref, rtol=1e-13, atol=1e-13)


def test_gmres_zero_rhs_and_exact_start():
    A = torch.eye(4, dtype=torch.complex128) * 2
    x, it, rr = gmres(torch, lambda v: A @ v, torch.zeros(4, dtype=torch.complex128))
    assert it == 0 and rr == 0.0 and not x.any()
    b = torch.ones(4, dtype=torch.complex128)
    x, it, rr = gmres(torch, lambda v: A @ v, b, x0=b / 2)
    assert it == 0 and rr == 0.0


def _systems(nb, n, seed):
    rng = np.random.default_rng(seed)
    As = [np.eye(n) * (1.5 + q) + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
          / np.sqrt(n) for q in range(nb)]
    bs = [rng.standard_normal(n) + 1j * rng.standard_normal(n) for _ in range(nb)]
    return As, bs


def test_gmres_batched_matches_direct_and_single_system_counts():
    nb, n = 5, 90
    As, bs = _systems(nb, n, 2)
    bs[3] = np.zeros(n, dtype=np.complex128)          # a zero right-hand side in the batch
    At = torch.as_tensor(np.stack(As))
    B = torch.as_tensor(np.stack(bs))

    def op(X, active):
        Y = torch.zeros_like(X)
        for q in active.nonzero()[0].tolist():
            Y[q] = At[q] @ X[q]
        return Y

    for restart in (7, 40):
        X, its, rr = gmres_batched(torch, op, B, tol=1e-12, restart=restart, maxiter=4000)
        for q in range(nb):
            if q == 3:
                assert its[q] == 0 and rr[q] == 0.0 and not X[q].any()
                continue
            xd = np.linalg.solve(As[q], bs[q])
            assert rr[q] <= 1e-12
            assert np.linalg.norm(X[q].numpy() - xd) <= 1e-10 * np.linalg.norm(xd)
            _, it1, _ = gmres(torch, lambda v: At[q] @ v, B[q], tol=1e-12, restart=restart,
                              maxiter=4000)
            assert abs(int(its[q]) - it1) <= 1, (q, its[q], it1)


def test_collocation_operator_batched_reads_column_major_storage():
    rng = np.random.default_rng(4)
    nb, n = 3, 11
    Fs = [rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) for _ in range(nb)]
    C = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    storage = torch.as_tensor(np.stack([np.asfortranarray(F).ravel(order="F").reshape(n, n)
                                        for F in Fs]))
    op = collocation_operator_batched(torch, storage, torch.as_tensor(C))
    X = rng.standard_normal((nb, n)) + 1j * rng.standard_normal((nb, n))
    active = np.array([True, False, True])
    Y = op(torch.as_tensor(X), active).numpy()
    for q in range(nb):
        ref = (Fs[q] + np.diag(1.0 - C)) @ X[q] if active[q] else np.diag(1.0 - C) @ X[q]
        assert np.allclose(Y[q], ref, rtol=1e-13, atol=1e-13)
