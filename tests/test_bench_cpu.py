"""bench.py's host-side contract: one GPU only (a clear exit for --gpus > 1, before any
device work), and the algorithmic FLOP count that roofline.achieved is built on."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_algorithmic_flops_reproduce_the_worked_count():
    """The quadratic 10242-node sphere at (n,m) = (4,4): 52,439,040 pairs, 846,786,384
    quadrature points (16 per vertex-rule pair, 32 per mid-edge pair), 83 FLOPs per point
    (35 + 8 n_s) and 60 per pair (10 n_s selection): 73.4 GFLOP per assembly."""
    from bie_amd.rules import rule_size
    lens = [rule_size(c, 4, 4) for c in range(1, 7)]
    assert lens == [16, 16, 16, 32, 32, 32]
    pairs, points = 52_439_040, 846_786_384
    mid = (points - 16 * pairs) // 16          # pairs on the 32-point mid-edge rules
    hist = np.zeros(7, dtype=np.int64)
    hist[1:4] = [(pairs - mid) // 3, (pairs - mid) // 3, pairs - mid - 2 * ((pairs - mid) // 3)]
    hist[4:7] = [mid // 3, mid // 3, mid - 2 * (mid // 3)]
    fl = bench.algorithmic_flops(hist, lens, 6)
    assert fl["pairs"] == pairs and fl["points_dense"] + fl["points_mid"] == points
    assert fl["dense"] + fl["mid"] == points * 83 + pairs * 60
    assert abs((fl["dense"] + fl["mid"]) / 1e9 - 73.4) < 0.05
