"""The fast operators' partial near field (partial_assemble_parallel!,
src/3d/fast_multipole_method.jl:11-81) in the NumPy oracle, pinned:

* gauss_points_triangle (adaptive_integration.jl:13-64): weights sum to the triangle's area
  1/2 and each rule integrates the monomials u^a v^b exactly up to its degree (1, 2, 3, 5,
  6, 7 for n = 1, 3, 4, 7, 12, 13) - a mistyped point or weight of the table fails it;
* on the pattern's elements that do not touch the source node the integrand is smooth: the
  13-point values agree with the dense oracle's (8, 8) values of that element to the
  13-point rule's accuracy (basis, kernels, normal and the column of every entry);
* on the elements around the node (1/r singular at a corner) the single-layer row sums at
  k = 0 converge to the closed-form integral as n grows.
"""
from math import factorial

import numpy as np
import pytest

import oracle
import oracle_np
from bie_amd import mesh as M

DEGREE = {1: 1, 3: 2, 4: 3, 7: 5, 12: 6, 13: 7}


@pytest.mark.parametrize("n", sorted(DEGREE))
def test_gauss_points_triangle_exactness(n):
    u, v, w = oracle_np.gauss_points_triangle(n)
    assert len(w) == n and abs(w.sum() - 0.5) < 1e-14
    for a in range(DEGREE[n] + 1):
        for b in range(DEGREE[n] + 1 - a):
            exact = factorial(a) * factorial(b) / factorial(a + b + 2)
            assert abs(np.sum(w * u ** a * v ** b) - exact) < 2e-14, (a, b)


def _pattern(mesh, depth):
    els, cols = oracle.connected_sources(mesh, depth)
    return els, cols


@pytest.mark.parametrize("kind", ["linear", "quadratic"])
def test_smooth_entries_match_dense(kind):
    mesh = M.perturb_radial(M.icosphere(2, kind), 0.02, 5)
    els, cols = _pattern(mesh, 2)
    k = 2.0
    checked = 0
    for i in range(0, mesh.n_sources, 11):
        for e in els[i]:
            if i in mesh.physics_topology[:, e]:
                continue          # the element touches the source node: singular
            Fp, Gp = oracle_np.partial_assemble(mesh, k, mesh.sources[:, [i]], [[e]], [cols[i]],
                                                n=13)
            # the dense oracle on a mesh of this one element, target x_i, (8, 8)
            one = M.Mesh3d(mesh.coordinates, mesh.topology[:, [e]], mesh.sources,
                           mesh.physics_topology[:, [e]], mesh.geometry_kind,
                           mesh.physics_kind)
            Fd, Gd, _ = oracle_np.assemble(one, k, mesh.sources[:, [i]], n=8, m=8)
            for c in mesh.physics_topology[:, e]:
                t = list(cols[i]).index(c)
                # the 13-point rule's own error at one element's distance is 1e-5 - 2e-4
                # (quadratic bases); a wrong column, basis or normal is an O(1) error
                assert abs(Gp[t] - Gd[0, c]) <= 1e-3 * np.max(np.abs(Gd))
                assert abs(Fp[t] - Fd[0, c]) <= 5e-3 * np.max(np.abs(Fd))
                checked += 1
    assert checked > 50


def test_singular_rows_converge():
    # k = 0, flat elements, source at a vertex: sum over the row's depth-1 entries of G is the
    # integral of 1/(4 pi r) over the elements around the node
    mesh = M.icosphere(2, "linear")
    els, cols = _pattern(mesh, 1)
    i = 5
    exact = 0.0
    for e in els[i]:
        P = mesh.coordinates[:, mesh.topology[:, e]]
        j = list(mesh.topology[:, e]).index(i)
        A, B, Cc = P[:, j], P[:, (j + 1) % 3], P[:, (j + 2) % 3]
        exact += _vertex_potential(A, B, Cc)
    errs = []
    for n in (1, 3, 7, 13):
        _, Gp = oracle_np.partial_assemble(mesh, 0.0, mesh.sources[:, [i]], [els[i]], [cols[i]],
                                           n=n)
        errs.append(abs(Gp.sum().real - exact) / exact)
    assert errs[-1] < errs[0] and errs[-1] < 0.2


def _vertex_potential(A, B, C):
    """Closed form of the integral of 1/(4 pi |y - A|) over the flat triangle ABC (A a vertex):
    h / (4 pi) * [asinh(tan t) ] between the angles of AB and AC seen from the foot of the
    altitude from A onto BC."""
    BC = C - B
    L = np.linalg.norm(BC)
    t = np.dot(A - B, BC) / L ** 2
    foot = B + t * BC
    h = np.linalg.norm(A - foot)
    s0 = np.dot(B - foot, BC) / L
    s1 = np.dot(C - foot, BC) / L
    return h / (4 * np.pi) * (np.arcsinh(s1 / h) - np.arcsinh(s0 / h))
