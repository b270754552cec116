"""Restarted GMRES (the caller of the assembled operators) on the CPU with torch tensors."""
import numpy as np
import torch

from bie_amd.solvers import collocation_operator, gmres


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
