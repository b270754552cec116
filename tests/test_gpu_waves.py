"""The single-k slot kernel's workgroup shapes on real k: 8 waves (two per SIMD, the
default for six-node physics) and 12 waves (three per SIMD, the default for three-node
physics, its rows staged after each pair rather than held through it), each forced with
BIE_AMD_WAVES, against the oracle on the rule lengths with unrolled instances (4, 6, 16
points), with and without G."""
import os

import numpy as np
import pytest

import oracle
from bie_amd import mesh as M
from bie_amd.assembly import assemble_parallel
from gpu_helpers import TOL, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("waves", ["8", "12"])
@pytest.mark.parametrize("kind", ["linear", "quadratic"])
@pytest.mark.parametrize("nm", [(2, 2), (3, 2), (4, 4)])
def test_workgroup_shapes_match_oracle(waves, kind, nm, monkeypatch):
    monkeypatch.setenv("BIE_AMD_WAVES", waves)
    mesh = M.perturb_radial(M.icosphere(6 if kind == "linear" else 4, kind), 0.01, 2)
    n, m = nm
    print(f"[case] waves={waves} {kind} nm={nm} N={mesh.n_sources}", flush=True)
    Fo, Go, Co = oracle.assemble(mesh, 4.0, n=n, m=m)
    for gOn in (True, False):
        F, G, C = assemble_parallel(mesh, 4.0, n=n, m=m, gOn=gOn)
        assert rel(F, Fo) <= TOL and rel(C, Co) <= TOL
        if gOn:
            assert rel(G, Go) <= TOL
