"""CPU checks of bench.py: the one-GPU contract, the roofline constants, and
the CPU-baseline leg's report."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from wind_amd import networks as W  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_roofline_constants():
    # hhe6 per solve: 2N^2 + 2/3 N^3 + 2N^2 + 2N at N = 6; per attempt two
    # network evaluations (674 flops each) and three solves over 3 sub-steps
    assert bench.F_SOLVE == 72 + 144 + 72 + 12 == 300
    assert bench.F_SUBSTEP == (2 * 674 + 3 * 300) / 3
    # state in and out (2N floats) and one constant per system
    assert bench.BYTES_PER_SYSTEM == 4 * (2 * 6 + 1) == 52
    # mass action at N = 157: 64N + 4N^2 + 2/3 N^3 + 2N
    assert bench.dense_flops(157) == 64 * 157 + 4 * 157**2 + (2 * 157**3) // 3 + 2 * 157 == 2688886


def test_cpu_baseline_report():
    p = W.hhe6_params()
    y0, nH = W.hhe6_systems(4096, seed=1)
    out = bench.cpu_baseline(p, y0, nH, 0.4, 1e-3, 0)
    assert out["kind"] == "port" and out["unit"] == "systems/s"
    assert out["cores"] >= 1 and out["value"] > 0
    assert out["cores"] <= out["host_cpus_visible"] == len(os.sched_getaffinity(0))
    assert out["single_core"]["cores"] == 1 and out["single_core"]["value"] > 0
    assert "median of 3" in out["sample"]
    assert np.isfinite(out["value"])
