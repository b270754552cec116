"""Dense apply Y = A X on FP64 MFMA vs NumPy (complex128)."""
import numpy as np
import pytest

from bie_amd import mesh as M
from bie_amd.assembly import assemble_parallel
from bie_amd.operators import apply

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def crandn(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)


@pytest.mark.parametrize("m,k,nrhs", [(1, 1, 1), (17, 5, 3), (100, 100, 32), (129, 257, 33),
                                      (1000, 777, 32), (64, 4000, 7), (513, 96, 70)])
def test_zgemm_matches_numpy(m, k, nrhs):
    rng = np.random.default_rng(m * 7 + k + nrhs)
    A = crandn(rng, m, k)
    X = crandn(rng, k, nrhs)
    print(f"[case] zgemm {m}x{k} nrhs={nrhs}", flush=True)
    Y = apply(A, X)
    assert rel(Y, A @ X) < 1e-14


def test_zgemm_structured_values_pin_operand_layout():
    # integer-valued entries make every product exact: any lane/operand mix-up is visible
    m, k, nrhs = 48, 20, 32
    i, j = np.meshgrid(np.arange(m), np.arange(k), indexing="ij")
    A = (i + 3 * j) + 1j * (2 * i - j)
    kk, c = np.meshgrid(np.arange(k), np.arange(nrhs), indexing="ij")
    X = (kk - c) + 1j * (kk + 2 * c)
    Y = apply(A.astype(np.complex128), X.astype(np.complex128))
    assert np.array_equal(Y, A @ X)


def test_apply_assembled_operator():
    mesh = M.icosphere(6, "quadratic")
    F, G, C = assemble_parallel(mesh, 5.0, mesh.sources, n=3, m=2)
    rng = np.random.default_rng(0)
    X = crandn(rng, mesh.n_sources, 32)
    assert rel(apply(F, X), F @ X) < 1e-13
    assert rel(apply(G, X[:, :1]), G @ X[:, :1]) < 1e-13


@pytest.mark.parametrize("nb,m,k", [(5, 300, 257), (3, 4002, 4002), (2, 1, 1), (4, 64, 9)])
def test_zgemv_batched_adds_into_listed_systems_only(nb, m, k):
    """y_q += A_q x_q for the listed q (column-major A_q in one strided block, x_q rows of a
    strided view, y pre-filled), other systems untouched."""
    import torch
    from bie_amd.operators import zgemv_batched_device
    rng = np.random.default_rng(nb * 1000 + m + k)
    dev = torch.device("cuda:0")
    As = [crandn(rng, m, k) for _ in range(nb)]
    # storage[q] is the row-major (k, m) view of column-major A_q, as the sweep writes F_q
    S = torch.as_tensor(np.stack([np.asfortranarray(A).ravel(order="F").reshape(k, m)
                                  for A in As]), device=dev)
    Xfull = torch.as_tensor(crandn(rng, nb, 3, k), device=dev)
    X = Xfull[:, 1]                                   # row stride 3 k
    Y0 = crandn(rng, nb, m)
    Y = torch.as_tensor(Y0, device=dev)
    idx = np.array([q for q in range(nb) if q != 1 or nb == 1], dtype=np.int32)
    di = torch.as_tensor(idx, device=dev)
    print(f"[case] zgemv batched nb={nb} {m}x{k} systems={idx.tolist()}", flush=True)
    zgemv_batched_device(m, k, S.data_ptr(), m * k, m, X.data_ptr(), X.stride(0), Y.data_ptr(),
                         m, di.data_ptr(), idx.size, torch.cuda.current_stream().cuda_stream)
    Yh = Y.cpu().numpy()
    Xh = X.cpu().numpy()
    for q in range(nb):
        ref = Y0[q] + As[q] @ Xh[q] if q in idx else Y0[q]
        assert rel(Yh[q], ref) < 1e-14, q
