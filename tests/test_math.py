"""The controller's portable pow (sdeh_math.h pow_ctrl, shared by the kernels
and the oracle) against 120-bit mpmath: below 0.6 ulp on the controller's
range, with the special values the controller relies on."""
import math

import mpmath
import numpy as np

from oracle import pyoracle


def test_pow_ctrl_accuracy():
    lib = pyoracle.load()
    mpmath.mp.prec = 120
    rng = np.random.default_rng(0)
    xs = np.concatenate([10.0 ** rng.uniform(-300, 300, 3000), rng.uniform(0, 2, 3000),
                         10.0 ** rng.uniform(-8, 3, 3000), [5e-324, 2.2250738585072014e-308]])
    worst = 0.0
    for y in (7 / 50, 2 / 25):
        for x in xs:
            got = lib.oracle_pow_ctrl(float(x), y)
            ex = mpmath.power(mpmath.mpf(float(x)), mpmath.mpf(y))
            err = abs(float((mpmath.mpf(got) - ex) / math.ulp(float(ex))))
            worst = max(worst, err)
    assert worst < 0.6, worst


def test_pow_ctrl_special_values():
    lib = pyoracle.load()
    assert lib.oracle_pow_ctrl(0.0, 0.14) == 0.0
    assert lib.oracle_pow_ctrl(1.0, 0.14) == 1.0
    assert lib.oracle_pow_ctrl(math.inf, 0.14) == math.inf
    assert math.isnan(lib.oracle_pow_ctrl(math.nan, 0.14))
    assert math.isnan(lib.oracle_pow_ctrl(-1.0, 0.14))
    # agrees with the C library pow in the overwhelming majority of cases
    rng = np.random.default_rng(1)
    xs = 10.0 ** rng.uniform(-10, 3, 20000)
    same = np.mean([lib.oracle_pow_ctrl(float(x), 0.14) == math.pow(float(x), 0.14) for x in xs])
    assert same > 0.98
