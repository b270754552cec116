"""Adjoint gradients against central finite differences of the oracle at frozen rule
selection (h = 1e-6), and dJ/dk against the oracle's Taylor-in-k operator."""
import numpy as np
import pytest

import oracle
from bie_amd import mesh as M
from bie_amd.adjoint import assembly_vjp, scattering_gradient, source_nodes, plane_wave
from bie_amd.assembly import DeviceMesh

pytestmark = pytest.mark.gpu


def crandn(rng, n):
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def functional(mesh, X, k, rules, alpha, beta, p, targets, n, m):
    mh = M.with_coordinates(mesh, X)
    F, _, C = oracle.assemble_frozen(mh, k, rules, targets=targets, n=n, m=m, gOn=False)
    v = alpha @ (F @ p)
    if beta is not None:
        v += beta @ C
    return float(np.real(v))


@pytest.mark.parametrize("kind,field", [("linear", False), ("linear", True),
                                         ("quadratic", False), ("quadratic", True)])
def test_vjp_matches_frozen_rule_finite_differences(kind, field):
    mesh = M.perturb_radial(M.icosphere(3, kind), 0.01, 1)
    rng = np.random.default_rng(5)
    n, m, k = 3, 3, 4.0
    if field:
        targets = M.fibonacci_sphere(25, 2.5)
        Mt = 25
        beta = None
        tnode = None
    else:
        targets = None
        Mt = mesh.n_sources
        beta = crandn(rng, Mt)
        tnode = source_nodes(mesh)
    alpha = crandn(rng, Mt)
    p = crandn(rng, mesh.n_sources)
    print(f"[case] vjp {kind} field={field}", flush=True)
    g, gk = assembly_vjp(DeviceMesh(mesh), k, alpha, p, beta=beta, targets=targets,
                         target_node=tnode, n=n, m=m)
    rules = oracle.select(mesh, targets)
    X = mesh.coordinates
    for trial in range(3):
        V = rng.standard_normal(X.shape)
        h = 1e-6
        fp = functional(mesh, X + h * V, k, rules, alpha, beta, p, targets, n, m)
        fm = functional(mesh, X - h * V, k, rules, alpha, beta, p, targets, n, m)
        fd = (fp - fm) / (2 * h)
        ad = float(np.sum(g * V))
        assert abs(ad - fd) <= 1e-6 * abs(fd), (trial, ad, fd)
    # dk against the oracle's d/dk operator (taylor_assemble!, order 1)
    F1, _ = oracle.assemble_taylor(mesh, k, 1, targets=targets, n=n, m=m)
    ref = float(np.real(alpha @ (F1 @ p)))
    assert abs(gk - ref) <= 1e-9 * abs(ref)


@pytest.mark.parametrize("kind,field", [("linear", False), ("quadratic", False),
                                         ("quadratic", True)])
def test_vjp_matches_dual_number_derivatives(kind, field):
    """The VJP dotted with random directions against forward-mode dual numbers through
    the assembly (tests/dual_oracle.py): exact derivatives, so a tight tolerance."""
    from dual_oracle import assemble_F_C
    mesh = M.perturb_radial(M.icosphere(3, kind), 0.01, 4)
    rng = np.random.default_rng(8)
    n, m, k = 3, 2, 6.0
    targets = M.fibonacci_sphere(20, 2.0) if field else None
    Mt = 20 if field else mesh.n_sources
    beta = None if field else crandn(rng, Mt)
    alpha, p = crandn(rng, Mt), crandn(rng, mesh.n_sources)
    print(f"[case] vjp dual {kind} field={field}", flush=True)
    g, gk = assembly_vjp(DeviceMesh(mesh), k, alpha, p, beta=beta, targets=targets,
                         target_node=None if field else source_nodes(mesh), n=n, m=m)
    rules = oracle.select(mesh, targets)
    X = mesh.coordinates
    for trial in range(2):
        V = rng.standard_normal(X.shape)
        F, C = assemble_F_C(mesh, X, V, k, 0.0, rules, targets=targets, n=n, m=m)
        dd = np.real(alpha @ (F.d @ p) + (beta @ C.d if beta is not None else 0.0))
        assert abs(float(np.sum(g * V)) - dd) <= 1e-10 * abs(dd), (trial, dd)
    F, C = assemble_F_C(mesh, X, np.zeros_like(X), k, 1.0, rules, targets=targets, n=n, m=m)
    ddk = np.real(alpha @ (F.d @ p))
    assert abs(gk - ddk) <= 1e-10 * abs(ddk)


@pytest.mark.parametrize("kind,field", [("linear", False), ("quadratic", True)])
def test_vjp_complex_k_matches_dual_number_derivatives(kind, field):
    """Complex wavenumber (the lossy systems' case): the VJP in X and the two k-derivatives
    (d/dRe k, d/dIm k) against forward-mode dual numbers and against central differences
    of the oracle's frozen-rule assembly."""
    from dual_oracle import assemble_F_C
    mesh = M.perturb_radial(M.icosphere(3, kind), 0.01, 6)
    rng = np.random.default_rng(9)
    n, m, k = 3, 3, 4.0 + 0.7j
    targets = M.fibonacci_sphere(16, 2.2) if field else None
    Mt = 16 if field else mesh.n_sources
    beta = None if field else crandn(rng, Mt)
    alpha, p = crandn(rng, Mt), crandn(rng, mesh.n_sources)
    print(f"[case] vjp complex k {kind} field={field}", flush=True)
    g, (gkr, gki) = assembly_vjp(DeviceMesh(mesh), k, alpha, p, beta=beta, targets=targets,
                                 target_node=None if field else source_nodes(mesh), n=n, m=m)
    rules = oracle.select(mesh, targets)
    X = mesh.coordinates
    V = rng.standard_normal(X.shape)
    F, C = assemble_F_C(mesh, X, V, k, 0.0, rules, targets=targets, n=n, m=m)
    dd = np.real(alpha @ (F.d @ p) + (beta @ C.d if beta is not None else 0.0))
    assert abs(float(np.sum(g * V)) - dd) <= 1e-10 * abs(dd)
    h = 1e-6
    fp = functional(mesh, X + h * V, k, rules, alpha, beta, p, targets, n, m)
    fm = functional(mesh, X - h * V, k, rules, alpha, beta, p, targets, n, m)
    assert abs(float(np.sum(g * V)) - (fp - fm) / (2 * h)) <= 1e-6 * abs(dd)
    for dk, got in ((1.0, gkr), (1j, gki)):
        F, _ = assemble_F_C(mesh, X, np.zeros_like(X), k, dk, rules, targets=targets, n=n, m=m)
        ref = np.real(alpha @ (F.d @ p))
        assert abs(got - ref) <= 1e-10 * abs(ref), dk


def test_scattering_gradient_matches_finite_differences():
    # config-4 shape at reduced size: perturbed linear sphere, k = 5, field shell r = 3
    mesh = M.perturb_radial(M.icosphere(4, "linear"), 0.01, 1)
    pts = M.fibonacci_sphere(40, 3.0)
    k, n, m = 5.0, 3, 3
    J, grad, dk = scattering_gradient(mesh, k, pts, n=n, m=m)
    rules_s = oracle.select(mesh)
    rules_f = oracle.select(mesh, pts)

    def objective(X, kk):
        mh = M.with_coordinates(mesh, X)
        F, _, C = oracle.assemble_frozen(mh, kk, rules_s, n=n, m=m, gOn=False)
        Ff, _, _ = oracle.assemble_frozen(mh, kk, rules_f, targets=pts, n=n, m=m, gOn=False,
                                          cOn=False)
        p = np.linalg.solve(F + np.diag(1.0 - C), plane_wave(mh.sources, kk))
        u = plane_wave(pts, kk) - Ff @ p
        return float(np.real(np.vdot(u, u)))

    X = mesh.coordinates
    assert abs(objective(X, k) - J) <= 1e-10 * J
    rng = np.random.default_rng(2)
    h = 1e-6
    for flat in rng.choice(X.size, 10, replace=False):
        d, j = np.unravel_index(flat, X.shape)
        Xp = X.copy(); Xp[d, j] += h
        Xm = X.copy(); Xm[d, j] -= h
        fd = (objective(Xp, k) - objective(Xm, k)) / (2 * h)
        assert abs(grad[d, j] - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(grad).max()), (d, j)
    fdk = (objective(X, k + h) - objective(X, k - h)) / (2 * h)
    assert abs(dk - fdk) <= 1e-6 * abs(fdk)


def test_config4_adjoint_full_size_against_finite_differences():
    """Full config 4: 2562-node linear sphere with radial N(0, 0.01^2) perturbation (seed 1),
    k = 5, 100 Fibonacci field points on r = 3, (n, m) = (4, 4); central differences with
    h = 1e-6 on 50 coordinates drawn with seed 2 (draws whose +-h evaluations change a rule
    choice are replaced), and on k.  The differences use the GPU assembly and a GPU LU
    solve."""
    import torch

    mesh = M.perturb_radial(M.icosphere(16, "linear"), 0.01, 1)
    assert mesh.n_sources == 2562 and mesh.n_elements == 5120
    pts = M.fibonacci_sphere(100, 3.0)
    k, n, m = 5.0, 4, 4
    dev = torch.device("cuda:0")
    print("[case] config4 adjoint full size", flush=True)
    J, grad, dk = scattering_gradient(mesh, k, pts, n=n, m=m)
    N, P = mesh.n_sources, pts.shape[1]
    tp = torch.as_tensor(np.asfortranarray(pts).ravel(order="F"), device=dev)

    def rules(dm):
        rs = torch.empty(N * dm.n_elements, dtype=torch.int8, device=dev)
        rf = torch.empty(P * dm.n_elements, dtype=torch.int8, device=dev)
        dm.select_rules(rs.data_ptr())
        dm.select_rules(rf.data_ptr(), targets_ptr=tp.data_ptr(), n_targets=P)
        torch.cuda.synchronize()
        return rs, rf

    def objective(X, kk):
        mh = M.with_coordinates(mesh, X)
        dm = DeviceMesh(mh)
        F = torch.empty((N, N), dtype=torch.complex128, device=dev)   # F.T in memory
        G = torch.empty_like(F)
        C = torch.empty(N, dtype=torch.complex128, device=dev)
        Ff = torch.empty((N, P), dtype=torch.complex128, device=dev)  # Ff.T in memory
        Gf = torch.empty_like(Ff)
        Cf = torch.empty(P, dtype=torch.complex128, device=dev)
        dm.assemble(kk, F.data_ptr(), G.data_ptr(), C.data_ptr(), n=n, m=m, gOn=False)
        dm.assemble(kk, Ff.data_ptr(), Gf.data_ptr(), Cf.data_ptr(), targets_ptr=tp.data_ptr(),
                    n_targets=P, n=n, m=m, gOn=False, cOn=False)
        torch.cuda.synchronize()
        A = F.T + torch.diag(1.0 - C)
        b = torch.as_tensor(plane_wave(mh.sources, kk), device=dev)
        p = torch.linalg.solve(A, b)
        u = torch.as_tensor(plane_wave(pts, kk), device=dev) - Ff.T @ p
        return float(torch.sum(torch.abs(u) ** 2)), rules(dm)

    X = mesh.coordinates
    J0, (rs0, rf0) = objective(X, k)
    assert abs(J0 - J) <= 1e-10 * J
    rng = np.random.default_rng(2)
    h = 1e-6
    ad, fd, skipped = [], [], 0
    # 50 coordinates at which J is differentiable: a draw whose +-h evaluations change a
    # rule choice (the selection is piecewise constant in X) is replaced by the next one
    for flat in rng.permutation(X.size):
        if len(fd) == 50:
            break
        d, j = np.unravel_index(flat, X.shape)
        Xp = X.copy(); Xp[d, j] += h
        Xm = X.copy(); Xm[d, j] -= h
        Jp, (rsp, rfp) = objective(Xp, k)
        Jm, (rsm, rfm) = objective(Xm, k)
        if not (torch.equal(rsp, rs0) and torch.equal(rsm, rs0) and torch.equal(rfp, rf0)
                and torch.equal(rfm, rf0)):
            skipped += 1            # a rule choice flips inside [-h, h]: not differentiable
            continue
        fd.append((Jp - Jm) / (2 * h))
        ad.append(grad[d, j])
    ad, fd = np.array(ad), np.array(fd)
    err = np.abs(ad - fd)
    print(f"[config4] J={J:.6e} skipped={skipped} max|fd|={np.abs(fd).max():.3e} "
          f"max abs err={err.max():.3e} vec rel err={np.linalg.norm(ad - fd) / np.linalg.norm(fd):.3e}",
          flush=True)
    assert len(fd) == 50 and skipped <= 50
    assert np.linalg.norm(ad - fd) <= 1e-6 * np.linalg.norm(fd)
    Jkp, _ = objective(X, k + h)
    Jkm, _ = objective(X, k - h)
    fdk = (Jkp - Jkm) / (2 * h)
    print(f"[config4] dJ/dk adjoint={dk:.12e} fd={fdk:.12e}", flush=True)
    assert abs(dk - fdk) <= 1e-6 * abs(fdk)
