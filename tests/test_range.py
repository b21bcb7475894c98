"""The fixed-step time grid t0:dt:tf (Julia Float64 range rules,
gpuatsit5.jl:92): the library's restatement, the oracle's, and the Python
twin's agree, and all reproduce known Julia results."""
from fractions import Fraction

import numpy as np
import pytest

from oracle import pyoracle, pytwin
from simplediffeq_hip.runtime import fixed_grid


def lib_grid(a, s, b, cap=100000):
    n, el = fixed_grid(a, s, b, cap)
    return n, el


def test_known_julia_ranges():
    # length(0:0.01:1) == 101, elements are the correctly rounded i/100
    n, el = lib_grid(0.0, 0.01, 1.0)
    assert n == 101
    assert all(el[i] == float(Fraction(i, 100)) for i in range(101))
    # collect(0:0.1:0.3) == [0.0, 0.1, 0.2, 0.3] (the rational path: last element
    # is exactly 0.3, not 3 * 0.1 = 0.30000000000000004)
    n, el = lib_grid(0.0, 0.1, 0.3)
    assert n == 4 and list(el) == [0.0, 0.1, 0.2, 0.3]
    # the reference's Lorenz test grid 0:0.1:100 (gpusimpleatsit5_tests.jl:52-57)
    n, el = lib_grid(0.0, 0.1, 100.0)
    assert n == 1001 and el[-1] == 100.0 and el[37] == 3.7
    # Float64(0.1f0) = 0.10000000149011612 has no small rational form: the literal
    # fallback; 10 steps would overshoot 1.0, so the range has 10 elements
    n, el = lib_grid(0.0, float(np.float32(0.1)), 1.0)
    assert n == 10
    # empty and single-point ranges
    assert lib_grid(1.0, 0.1, 0.0)[0] == 0
    assert lib_grid(0.5, 0.1, 0.5)[0] == 1
    assert fixed_grid(0.0, 0.0, 1.0)[0] == -1


@pytest.mark.parametrize("seed", range(4))
def test_three_restatements_agree(seed):
    rng = np.random.default_rng(seed)
    cases = []
    for _ in range(150):
        kind = rng.integers(4)
        if kind == 0:
            a, s, b = 0.0, 1.0 / rng.integers(1, 400), float(rng.integers(1, 20))
        elif kind == 1:
            a = float(rng.integers(-50, 50)) / 10
            s = float(rng.integers(1, 50)) / 100
            b = a + float(rng.integers(0, 300)) * s * 1.0000001
        elif kind == 2:
            a, s, b = float(rng.uniform(-1, 1)), float(rng.uniform(0.001, 0.3)), float(rng.uniform(1, 3))
        else:
            a, s, b = 0.0, float(np.float32(rng.uniform(0.001, 0.2))), float(rng.integers(1, 10))
        cases.append((a, s, b))
    for a, s, b in cases:
        n1, e1 = lib_grid(a, s, b)
        n2, e2, _ = pyoracle.julia_range(a, s, b)
        e3 = pytwin.julia_range(a, s, b)
        assert n1 == n2 == len(e3), (a, s, b)
        assert np.array_equal(e1, e2), (a, s, b)
        assert np.array_equal(e1, np.array(e3, dtype=float)), (a, s, b)
