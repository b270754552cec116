"""k-derivatives of the dense operators (taylor_assemble!, src/3d/rosebem.jl:56-157) and
their Taylor expansion (apply_taylor_expansion, rosebem.jl:164-199) on the GPU, against the
oracle's Taylor operator (itself pinned by central differences in k) and, at a larger size,
against the assembly at a nearby wavenumber."""
import numpy as np
import pytest

import oracle
from bie_amd import mesh as M
from bie_amd.assembly import assemble_parallel
from bie_amd.taylor import apply_taylor_expansion, taylor_assemble, taylor_coefficients

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.mark.parametrize("kind,f", [("linear", 2), ("quadratic", 1)])
@pytest.mark.parametrize("k", [3.0, 2.0 + 0.3j])
@pytest.mark.parametrize("field", [False, True])
def test_taylor_orders_match_oracle(kind, f, k, field):
    """Ten orders: partial last block of orders on both element kinds, complex k,
    collocation and field targets."""
    mesh = M.icosphere(f, kind)
    targets = None
    if field:
        rng = np.random.default_rng(5)
        targets = np.concatenate([1.5 * rng.standard_normal((3, 40)), mesh.sources[:, :3]],
                                 axis=1)
    Mo = 9
    print(f"[case] taylor {kind} f={f} k={k} field={field} orders 0..{Mo}", flush=True)
    F, G, C = taylor_assemble(mesh, k, targets, M=Mo, n=3, m=3)
    nt = mesh.n_sources if targets is None else targets.shape[1]
    assert F.shape == (nt, mesh.n_sources, Mo + 1) and G.shape == F.shape
    for q in range(Mo + 1):
        Fo, Go = oracle.assemble_taylor(mesh, k, q, targets=targets, n=3, m=3)
        assert rel(F[:, :, q], Fo) < TOL, q
        assert rel(G[:, :, q], Go) < TOL, q
    _, _, Co = oracle.assemble(mesh, k, targets=targets, n=3, m=3)
    assert rel(C, Co) < TOL


def test_taylor_flags_zero_disabled_outputs():
    mesh = M.icosphere(2, "linear")
    print("[case] taylor flags", flush=True)
    F, G, C = taylor_assemble(mesh, 2.0, M=2, fOn=False, cOn=False)
    assert not F.any() and not C.any()
    Go = oracle.assemble_taylor(mesh, 2.0, 2, n=3, m=3)[1]
    assert rel(G[:, :, 2], Go) < TOL


def test_taylor_reduced_basis():
    """U^H F_q U, U^H G_q U as the reference forms them for a basis U (rosebem.jl:141-151)."""
    mesh = M.icosphere(2, "linear")
    rng = np.random.default_rng(11)
    U = rng.standard_normal((mesh.n_sources, 7)) + 1j * rng.standard_normal((mesh.n_sources, 7))
    print("[case] taylor reduced basis r=7", flush=True)
    Fr, Gr, C = taylor_assemble(mesh, 4.0, M=3, U=U)
    assert Fr.shape == (7, 7, 4)
    for q in range(4):
        Fo, Go = oracle.assemble_taylor(mesh, 4.0, q, n=3, m=3)
        assert rel(Fr[:, :, q], U.conj().T @ Fo @ U) < 1e-12
        assert rel(Gr[:, :, q], U.conj().T @ Go @ U) < 1e-12


@pytest.mark.parametrize("shape,k,k0", [((33, 17, 1), 2.0, 1.0), ((64, 64, 12), 5.3, 5.0),
                                        ((1000, 3, 25), 1.0 + 0.2j, 1.5), ((7, 1, 64), 1.1, 1.0)])
def test_apply_taylor_expansion_matches_host_sum(shape, k, k0):
    rng = np.random.default_rng(shape[0])
    A = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    b = rng.standard_normal((shape[0], shape[2])) + 1j * rng.standard_normal((shape[0], shape[2]))
    print(f"[case] taylor expansion {shape}", flush=True)
    Ak, bk = apply_taylor_expansion(A, k, k0, b)
    c = taylor_coefficients(k, k0, shape[2])
    assert rel(Ak, A @ c) < 1e-14 and rel(bk, b @ c) < 1e-14


def test_expansion_reproduces_the_assembly_at_a_nearby_wavenumber():
    """642-node sphere, orders 0..9 about k0 = 5: the expansion at k0 + 0.1 agrees with
    the assembly at that k to within the truncation term (0.1 * 2)^10 / 10! ~ 3e-14
    (r <= 2 on the unit sphere)."""
    mesh = M.icosphere(8, "linear")
    k0, dk = 5.0, 0.1
    print("[case] taylor expansion vs assembly at k0 + dk, N=642", flush=True)
    F, G, C = taylor_assemble(mesh, k0, M=9, n=3, m=3)
    Fk = apply_taylor_expansion(F, k0 + dk, k0)
    Gk = apply_taylor_expansion(G, k0 + dk, k0)
    Fa, Ga, Ca = assemble_parallel(mesh, k0 + dk, n=3, m=3)
    assert rel(Fk, Fa) < 1e-11 and rel(Gk, Ga) < 1e-11
    assert rel(C, Ca) < TOL


def test_expansion_in_place_on_the_first_term():
    """dY = dA: the order-0 array is overwritten with the sum (each entry is read before it
    is written, by the same thread)."""
    from bie_amd.taylor import taylor_expand_device
    from gpu_helpers import DevBuf, sync
    rng = np.random.default_rng(4)
    n, nt = 100003, 7
    A = rng.standard_normal((nt, n)) + 1j * rng.standard_normal((nt, n))
    c = taylor_coefficients(2.2, 2.0, nt)
    buf = DevBuf(A.nbytes)
    buf.upload(np.ascontiguousarray(A))
    print("[case] taylor expansion in place", flush=True)
    taylor_expand_device(nt, n, buf.ptr, n, c, buf.ptr)
    sync()
    out = buf.download(np.zeros(nt * n, dtype=np.complex128))
    assert rel(out[:n], c @ A) < 1e-14
    assert np.array_equal(out[n:].reshape(nt - 1, n), A[1:])


@pytest.mark.parametrize("kind,f", [("linear", 2), ("quadratic", 1)])
@pytest.mark.parametrize("Mo", [9, 12])
def test_taylor_without_g_matches_oracle(kind, f, Mo):
    """gOn = false: the kernel packs two orders' F into each slot group and carries twice
    the orders per block; 10 and 13 orders leave partial blocks (13: an odd count, so the
    last group's second half is past the end).  Collocation and field targets, complex k."""
    mesh = M.icosphere(f, kind)
    rng = np.random.default_rng(Mo)
    targets = np.concatenate([1.7 * rng.standard_normal((3, 30)), mesh.sources[:, :4]], axis=1)
    k = 2.5 + 0.2j
    for tg in (None, targets):
        print(f"[case] taylor without G {kind} orders 0..{Mo} field={tg is not None}", flush=True)
        F, G, C = taylor_assemble(mesh, k, tg, M=Mo, gOn=False, n=3, m=3)
        assert not G.any()
        for q in range(Mo + 1):
            Fo, _ = oracle.assemble_taylor(mesh, k, q, targets=tg, n=3, m=3)
            assert rel(F[:, :, q], Fo) < TOL, q
        _, _, Co = oracle.assemble(mesh, k, targets=tg, n=3, m=3)
        assert rel(C, Co) < TOL


def test_taylor_edge_cases():
    """One order (M = 0) equals the plain assembly; no targets gives empty outputs; an empty
    expansion is a no-op."""
    from bie_amd.taylor import taylor_expand_device
    mesh = M.icosphere(2, "quadratic")
    print("[case] taylor edge cases", flush=True)
    F, G, C = taylor_assemble(mesh, 3.0 + 0.1j, M=0, n=3, m=3)
    Fa, Ga, Ca = assemble_parallel(mesh, 3.0 + 0.1j, n=3, m=3)
    assert F.shape[2] == 1
    assert rel(F[:, :, 0], Fa) < TOL and rel(G[:, :, 0], Ga) < TOL and rel(C, Ca) < TOL
    F, G, C = taylor_assemble(mesh, 2.0, np.zeros((3, 0)), M=3)
    assert F.shape == (0, mesh.n_sources, 4) and C.shape == (0,)
    taylor_expand_device(3, 0, 0, 0, taylor_coefficients(1.0, 0.5, 3), 0)
