"""Host C++ under AddressSanitizer + UndefinedBehaviorSanitizer.

Builds tests/sanitize/driver.cpp with the library's host sources (slot scheduler, rule
tables, sparse pattern and subdivision rules) and the CPU oracle, all with
-fsanitize=address,undefined, and runs it on a few meshes: any invalid access, leak or
undefined operation aborts the driver.  The driver also checks that the library's
subdivision rules are bit-identical to the oracle's and its sparse pattern equal to the
oracle's connected_sources.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from bie_amd import mesh as M

HERE = os.path.dirname(os.path.abspath(__file__))


def _write_mesh(path, m):
    with open(path, "wb") as f:
        np.array([m.n_elements, m.n_sources, m.geometry_kind, m.physics_kind,
                  m.n_coordinates], dtype=np.int32).tofile(f)
        np.asfortranarray(m.coordinates).ravel(order="F").astype(np.float64).tofile(f)
        m.topology.ravel(order="F").astype(np.int32).tofile(f)
        np.asfortranarray(m.sources).ravel(order="F").astype(np.float64).tofile(f)
        m.physics_topology.ravel(order="F").astype(np.int32).tofile(f)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "driver")
    subprocess.run(["bash", os.path.join(HERE, "sanitize", "build.sh"), exe], check=True,
                   capture_output=True, timeout=600)
    meshes = [M.icosphere(2, "linear"), M.icosphere(2, "quadratic"),
              M.icosphere(1, "quadratic", physics="linear"),
              M.ellipsoid(3, axes=(1.0, 0.6, 1.4))]
    paths = []
    for t, m in enumerate(meshes):
        p = str(tmp_path / f"mesh{t}.bin")
        _write_mesh(p, m)
        paths.append(p)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    r = subprocess.run([exe] + paths, capture_output=True, text=True, timeout=900, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "sanitized host checks passed" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
