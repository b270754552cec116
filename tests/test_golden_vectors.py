"""The oracle against the committed golden vectors (tests/golden/oracle_vectors.npz, written
by tests/golden/make_fixtures.py): rule tables and rule indices bit-exact, F, G, C to
round-off, and far-field entries of single flat elements against mpmath's adaptive
quadrature (30 digits) - an integration independent of the oracle's rules, which a dropped
factor, a wrong normal sign or a transposed basis fails by orders of magnitude."""
import os
import sys

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")
sys.path.insert(0, GOLD)
import make_fixtures as mf  # noqa: E402

VEC = np.load(os.path.join(GOLD, "oracle_vectors.npz"))


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.mark.parametrize("nm", mf.RULE_NM)
def test_oracle_reproduces_rule_tables(nm):
    n, m = nm
    for c in range(1, 7):
        assert np.array_equal(np.stack(oracle.rule(c, n, m)), VEC[f"rule_{n}{m}_c{c}"])


@pytest.mark.parametrize("name", ["coarse_quad", "coarse_lin", "ico2_lin", "ico1_quad"])
def test_oracle_reproduces_assembly_vectors(name):
    mesh = mf.meshes()[name]
    assert np.array_equal(oracle.select(mesh), VEC[f"{name}_rules"])
    for j, k in enumerate(mf.KS):
        if f"{name}_k{j}_F" not in VEC:
            continue
        F, G, C = oracle.assemble(mesh, k, n=mf.NM[0], m=mf.NM[1])
        assert rel(F, VEC[f"{name}_k{j}_F"]) < 1e-14
        assert rel(G, VEC[f"{name}_k{j}_G"]) < 1e-14
        assert rel(C, VEC[f"{name}_k{j}_C"]) < 1e-14


@pytest.mark.parametrize("i", range(len(mf.FAR)))
def test_oracle_far_entries_converge_to_mpmath(i):
    """(12,12) rules reach the adaptive-quadrature values to ~1e-15; lower orders converge
    towards them (vertex-clustered Duffy rules are not tuned for smooth integrands)."""
    lin = mf.meshes()["coarse_lin"]
    one = mf.single_element(lin, int(VEC[f"far_{i}_e"]))
    x = VEC[f"far_{i}_x"].reshape(3, 1)
    k = complex(VEC[f"far_{i}_k"])
    errs = []
    for n in (3, 4, 12):
        F, G, _ = oracle.assemble(one, k, targets=x, n=n, m=n)
        errs.append(max(rel(F[0], VEC[f"far_{i}_F"]), rel(G[0], VEC[f"far_{i}_G"])))
    assert errs[2] < 5e-15
    assert errs[1] < errs[0] and errs[2] < errs[1]
