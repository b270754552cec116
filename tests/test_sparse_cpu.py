"""Pins of the oracle's sparse near-field assembly (sparse_assemble_parallel!,
src/3d/assembly_collocation_losses.jl:200-405) against mathematics.

* subdivision rule (adaptive_integration.jl:172-246): weights sum to the reference
  triangle's area, linear functions integrate exactly, and the integral of 1/rho from a
  vertex and from an edge midpoint matches its closed form;
* connected_sources (:98-132) against the incidence-matrix formulation
  (depth 1: nonzeros of B B^T, depth 2: of (B B^T)^2; elements B, (B B^T) B);
* assembled values on flat (linear) triangles at k = 0: each row's single-layer sum equals
  the closed-form integral of 1/(4 pi r) over the elements around the node, and each
  depth-2 row's double-layer sum over an element off the node equals the element's solid
  angle / (4 pi) (van Oosterom-Strackee);
* on curved quadratic elements, depth-2 entries whose every element is in the near field
  agree with the dense oracle (a different quadrature of the same integrals).
None of these restates the oracle's formula.
"""
import numpy as np
import pytest
import scipy.sparse as sps

import oracle
from bie_amd import mesh as M

SOURCES = [[1, 0, 0], [0, 1, 0], [0, 0, 1], [0.5, 0.5, 0], [0, 0.5, 0.5], [0.5, 0, 0.5]]


def _lin(u, v):
    return np.stack([1 - u - v, u, v])


def _apex_integral(P, A, B):
    """Integral of 1/|x - P| over the flat triangle (P, A, B)."""
    t = (B - A) / np.linalg.norm(B - A)
    foot = A + np.dot(P - A, t) * t
    h = np.linalg.norm(P - foot)
    return h * (np.arcsinh(np.dot(B - foot, t) / h) - np.arcsinh(np.dot(A - foot, t) / h))


def _solid_angle(x, P1, P2, P3):
    """Signed solid angle of triangle (P1, P2, P3) seen from x, positive when
    ((P2 - P1) x (P3 - P1)) . (P1 - x) > 0 (van Oosterom and Strackee, 1983)."""
    a, b, c = P1 - x, P2 - x, P3 - x
    la, lb, lc = np.linalg.norm(a), np.linalg.norm(b), np.linalg.norm(c)
    num = np.dot(a, np.cross(b, c))
    den = la * lb * lc + np.dot(a, b) * lc + np.dot(a, c) * lb + np.dot(b, c) * la
    return 2.0 * np.arctan2(num, den)


# --------------------------------------------------------------- subdivision rule ---
@pytest.mark.parametrize("offr", [0.0, 0.1, 0.15])
@pytest.mark.parametrize("c", range(6))
def test_subdivision_rule_moments(c, offr):
    src = np.array([offr if x == 0 else x for x in SOURCES[c]], dtype=float)
    u, v, w = oracle.subdivision_rule(src)
    assert len(u) % 3 == 0 and np.all(w > 0)
    assert abs(w.sum() - 0.5) < 1e-14
    assert abs((w * u).sum() - 1 / 6) < 1e-14
    assert abs((w * v).sum() - 1 / 6) < 1e-14


def test_subdivision_rule_refines_toward_the_source():
    u, v, w = oracle.subdivision_rule(np.array([1.0, 0.0, 0.0]))
    # the source is the vertex (u, v) = (0, 0): the smallest cells sit there
    d = np.hypot(u, v)
    assert w[np.argmin(d)] < 1e-10 * w.max()
    far = oracle.subdivision_rule(np.array([1.0, 0.5, 0.5]))[2]  # off the plane
    assert len(far) < len(w)


@pytest.mark.parametrize("c", range(6))
def test_subdivision_rule_singular_integral(c):
    S = np.array(SOURCES[c], dtype=float)
    u, v, w = oracle.subdivision_rule(S)
    rho = np.linalg.norm(_lin(u, v) - S[:, None], axis=0)
    V = [np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, 1.0])]
    exact = 0.0
    for a, b in [(0, 1), (1, 2), (2, 0)]:  # sub-triangles with apex S
        if np.linalg.norm(np.cross(V[b] - V[a], S - V[a])) > 1e-12:
            exact += _apex_integral(S, V[a], V[b])
    exact /= np.sqrt(3.0)  # area scale of (u, v) -> (1 - u - v, u, v)
    assert abs((w / rho).sum() / exact - 1) < 1e-4


# --------------------------------------------------------------- connected_sources ---
@pytest.mark.parametrize("mesh", [M.icosphere(2, "linear"), M.icosphere(2, "quadratic"),
                                  M.icosphere(1, "quadratic", physics="linear")],
                         ids=["lin", "quad", "quadgeo_linphys"])
@pytest.mark.parametrize("depth", [1, 2])
def test_connected_sources_match_incidence_products(mesh, depth):
    N, E = mesh.n_sources, mesh.n_elements
    rows = mesh.physics_topology.ravel(order="F")
    cols = np.repeat(np.arange(E), mesh.physics_kind)
    B = sps.csr_matrix((np.ones(rows.size), (rows, cols)), shape=(N, E))
    A = (B @ B.T).tocsr()
    S, EL = (A, B) if depth == 1 else ((A @ A).tocsr(), (A @ B).tocsr())
    els, srcs = oracle.connected_sources(mesh, depth)
    for i in range(N):
        assert np.array_equal(srcs[i], np.sort(S[i].indices))
        assert np.array_equal(els[i], np.sort(EL[i].indices))


def test_connected_sources_rejects_depth_3():
    with pytest.raises(ValueError):
        oracle.connected_sources(M.icosphere(1), 3)


# --------------------------------------------------------- assembled values, k = 0 ---
def _row_sums(I, vals, N):
    return np.bincount(I, weights=vals.real, minlength=N)


@pytest.mark.parametrize("f", [2, 3])
def test_single_layer_row_sums_match_closed_form(f):
    mesh = M.ellipsoid(f, axes=(1.0, 0.8, 1.3))
    I, J, F, G = oracle.sparse_assemble(mesh, 0.0, depth=1)
    X, T = mesh.coordinates, mesh.topology
    exact = np.zeros(mesh.n_sources)
    for e in range(mesh.n_elements):
        P = X[:, T[:, e]]
        for a in range(3):  # the element's integral from each of its vertices
            exact[T[a, e]] += _apex_integral(P[:, a], P[:, (a + 1) % 3], P[:, (a + 2) % 3])
    exact /= 4 * np.pi
    got = _row_sums(I, G, mesh.n_sources)
    assert np.max(np.abs(got / exact - 1)) < 2e-4
    assert np.max(np.abs(G.imag)) == 0.0  # k = 0
    # the source lies in the plane of every depth-1 element: no double-layer term, up to
    # the rounding of n.d (absolute, ~1e-16) divided by r^3 at the smallest cells
    assert np.max(np.abs(F)) < 1e-7 * np.max(np.abs(G))


def test_double_layer_element_sums_match_solid_angles():
    mesh = M.ellipsoid(2, axes=(1.0, 0.8, 1.3))
    X, T, N = mesh.coordinates, mesh.topology, mesh.n_sources
    I, J, F, G = oracle.sparse_assemble(mesh, 0.0, depth=2, gOn=False)
    I1, J1, F1, _ = oracle.sparse_assemble(mesh, 0.0, depth=1, gOn=False)
    els2, _ = oracle.connected_sources(mesh, 2)
    els1, _ = oracle.connected_sources(mesh, 1)
    # the depth-1 elements add nothing to F (flat, in-plane), so each depth-2 row sum is
    # the sum over the row's off-node elements of their solid angles / (4 pi)
    exact = np.zeros(N)
    for i in range(N):
        for e in np.setdiff1d(els2[i], els1[i]):
            P = X[:, T[:, e]]
            exact[i] += _solid_angle(X[:, i], P[:, 0], P[:, 1], P[:, 2]) / (4 * np.pi)
    got = _row_sums(I, F, N)
    assert np.max(np.abs(got - exact)) < 2e-4 * np.max(np.abs(exact))
    assert np.max(np.abs(G)) == 0.0  # gOn = false


# ------------------------------------------------------- against the dense oracle ---
def test_curved_near_field_agrees_with_dense_oracle():
    mesh = M.icosphere(2, "quadratic")
    k = 1.3
    I, J, F, G = oracle.sparse_assemble(mesh, k, depth=2)
    Fd, Gd, _ = oracle.assemble(mesh, k, n=8, m=8)
    _, srcs1 = oracle.connected_sources(mesh, 1)
    # depth-2 entry (i, j) with j a depth-1 neighbour of i holds every element of column
    # j, like the dense entry
    keep = np.array([j in set(srcs1[i]) for i, j in zip(I, J)])
    for S, D in [(F, Fd), (G, Gd)]:
        d = D[I[keep], J[keep]]
        assert np.max(np.abs(S[keep] - d)) < 2e-4 * np.max(np.abs(d))


def test_hypot_variants_agree_off_ties():
    mesh = M.ellipsoid(2, axes=(1.0, 0.8, 1.3), order="quadratic")
    a = oracle.sparse_assemble(mesh, 2.0, depth=1, hypot="julia1.9")
    b = oracle.sparse_assemble(mesh, 2.0, depth=1, hypot="scaled")
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.max(np.abs(a[2] - b[2])) < 1e-12 * np.max(np.abs(a[2]))
