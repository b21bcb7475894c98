"""The parts of bench.py's output contract that need no GPU: the refusal of more than
one GPU, the algorithmic FLOP counting convention (FMA = 2; per attempt n_state *
stage + rhs_calls * F_f + 12 for the adaptive controller), and which launches carry a
counted HBM-traffic figure."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "--gpus 2 is not supported" in r.stderr
    assert r.stdout == ""  # no JSON line


def test_flops_per_attempt():
    lorenz = bench.LORENZ_FLOPS
    assert bench.algorithmic_flops(1, 3, lorenz, 1, 0, 0) == 261      # ATsit5: 3x67 + 6x8 + 12
    assert bench.algorithmic_flops(3, 3, lorenz, 1, 0, 0) == 435      # AVern7: 3x117 + 9x8 + 12
    assert bench.algorithmic_flops(3, 3, lorenz, 0, 1, 0) == lorenz   # its k1 per accepted step
    assert bench.algorithmic_flops(5, 3, lorenz, 1, 0, 0) == 765      # AVern9: 3x211 + 15x8 + 12
    assert bench.algorithmic_flops(5, 16, 528, 1, 0, 0) == 11308      # AVern9 on GLV-16
    # fixed-step Tsit5: 3x48 + 6x8 per step, and the FSAL seed once per trajectory
    assert bench.algorithmic_flops(0, 3, lorenz, 100, 100, 1) == 19208


def test_default_line_flops():
    # the default line's deterministic attempt count gives its algorithmic FLOPs
    assert bench.algorithmic_flops(1, 3, bench.LORENZ_FLOPS, 283952550, 0,
                                   1 << 22) == 74145169982


def test_counted_traffic_only_on_the_counted_kernel():
    key = ("f64", 1 << 20, 3)
    counted = bench.MEASURED_TRAFFIC[key]
    assert bench.measured_traffic(key, {"waves": 0, "per_lane": 1}) == counted
    assert bench.measured_traffic(key, {"waves": 4, "per_lane": 1}) is None
    assert bench.measured_traffic(key, {"waves": 0, "per_lane": 2}) is None
    assert bench.measured_traffic(("f64", 1 << 20, 1), {"waves": 0, "per_lane": 1}) is None
    assert bench.measured_traffic(("f64", 1 << 22, 2, "costmodel"),
                                  {"waves": 0, "per_lane": 1}) == 4.896e9
