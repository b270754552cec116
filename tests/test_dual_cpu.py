"""Forward-mode dual-number derivatives of the assembly (tests/dual_oracle.py) against the
C++ oracle: values equal the oracle's frozen-rule assembly, directional derivatives in the
coordinates match central differences of the oracle, and d/dk matches the oracle's
derivative operator (taylor_assemble! order 1, src/3d/rosebem.jl:56-157).  These are the
derivatives the adjoint kernel (bie_assembly_vjp_device) is checked against on the GPU."""
import numpy as np
import pytest

import oracle
from bie_amd import mesh as M
from dual_oracle import assemble_F_C


def _functional(F, C, alpha, beta, p):
    v = alpha @ (F @ p)
    if beta is not None:
        v = v + beta @ C
    return np.real(v)


@pytest.mark.parametrize("kind", ["linear", "quadratic"])
@pytest.mark.parametrize("field", [False, True])
def test_dual_directional_derivative_matches_oracle(kind, field):
    mesh = M.perturb_radial(M.icosphere(2, kind), 0.02, 3)
    rng = np.random.default_rng(11)
    n, m, k = 3, 3, 3.0
    targets = M.fibonacci_sphere(12, 2.2) if field else None
    Mt = 12 if field else mesh.n_sources
    rules = oracle.select(mesh, targets)
    alpha = rng.standard_normal(Mt) + 1j * rng.standard_normal(Mt)
    beta = None if field else rng.standard_normal(Mt) + 1j * rng.standard_normal(Mt)
    p = rng.standard_normal(mesh.n_sources) + 1j * rng.standard_normal(mesh.n_sources)
    X = mesh.coordinates
    V = rng.standard_normal(X.shape)
    F, C = assemble_F_C(mesh, X, V, k, 0.0, rules, targets=targets, n=n, m=m)
    # values: the oracle's frozen-rule assembly
    Fo, _, Co = oracle.assemble_frozen(mesh, k, rules, targets=targets, n=n, m=m, gOn=False)
    assert np.max(np.abs(F.v - Fo)) <= 1e-12 * np.max(np.abs(Fo))
    # (for field points off the surface C is a cancelling sum near zero: scale by F)
    assert np.max(np.abs(C.v - Co)) <= 1e-12 * max(np.max(np.abs(Co)), np.max(np.abs(Fo)))
    ad = _functional(F.d, C.d, alpha, beta, p)

    def J(Xs):
        mh = M.with_coordinates(mesh, Xs)
        Fh, _, Ch = oracle.assemble_frozen(mh, k, rules, targets=targets, n=n, m=m, gOn=False)
        return _functional(Fh, Ch, alpha, beta, p)

    h = 1e-6
    fd = (J(X + h * V) - J(X - h * V)) / (2 * h)
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)
    # d/dk through the dual against the oracle's derivative operator
    Fk, _ = assemble_F_C(mesh, X, np.zeros_like(X), k, 1.0, rules, targets=targets, n=n, m=m)
    F1, _ = oracle.assemble_taylor(mesh, k, 1, targets=targets, n=n, m=m)
    assert np.max(np.abs(Fk.d - F1)) <= 1e-10 * np.max(np.abs(F1))
