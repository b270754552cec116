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
