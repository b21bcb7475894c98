"""The controller's portable pow (sdeh_math.h pow_ctrl, shared by the kernels
and the oracle) against high-precision mpmath: the reduction tables hold what
the algorithm needs (exact reduction, double-double table values), the
double-double logarithm is accurate to far below a binary64 ulp, and pow is
below 0.51 ulp with the special values the controller relies on."""
import math
import os
import re
from fractions import Fraction

import mpmath
import numpy as np
import pytest

from oracle import pyoracle

HERE = os.path.dirname(os.path.abspath(__file__))
POWTAB = os.path.join(HERE, "..", "simplediffeq.jl_amd", "csrc", "sdeh_powtab.h")


def _powtab():
    text = open(POWTAB).read()
    consts = {m.group(1): float.fromhex(m.group(2)) if "0x" in m.group(2) else int(m.group(2))
              for m in re.finditer(r"constexpr \w+(?: \w+)* (\w+) = (-?0x[0-9a-fp.+-]+|-?\d+)", text)}
    consts["LOG_OFF"] = int(re.search(r"LOG_OFF = (0x[0-9a-f]+)ULL", text).group(1), 16)

    def rows(name):
        body = re.search(name + r"\[\d+\]\[\d\] = \{(.*?)\n\};", text, re.S).group(1)
        return [[float.fromhex(v.strip()) for v in r.split(",")]
                for r in re.findall(r"\{([^{}]*)\}", body)]
    return consts, rows("LOG_TAB"), rows("EXP_TAB")


def _bits(x):
    return int(np.float64(x).view(np.uint64))


def _double(b):
    return float(np.uint64(b).view(np.float64))


def test_log_table_reduction_is_exact():
    c, log_tab, _ = _powtab()
    assert len(log_tab) == 128
    rng = np.random.default_rng(0)
    for i, (invc, lh, ll, _) in enumerate(log_tab):
        # invc has at most 8 significant bits
        m, e = math.frexp(invc)
        assert (m * 256).is_integer(), i
        lo = c["LOG_OFF"] + (i << c["LOG_SHIFT"])
        hi = lo + (1 << c["LOG_SHIFT"]) - 1
        for b in [lo, hi] + list(rng.integers(lo, hi, 50)):
            z = Fraction(_double(int(b)))
            r = z * Fraction(invc) - 1
            assert abs(r) < Fraction(1, 128), (i, b)
            assert Fraction(float(r)) == r, (i, b)  # r = fma(z, invc, -1) is exact
        mpmath.mp.prec = 200
        lc = -mpmath.log(mpmath.mpf(invc))
        assert abs(mpmath.mpf(lh) + mpmath.mpf(ll) - lc) < mpmath.mpf(2) ** -104 * max(abs(lc), 1e-300) + mpmath.mpf(2) ** -200
    # the subintervals around 1 reduce by 1 exactly, so log(1) is exactly 0
    assert log_tab[74][:3] == [1.0, 0.0, 0.0] and log_tab[75][:3] == [1.0, 0.0, 0.0]


def test_exp_table_and_split_constants():
    c, _, exp_tab = _powtab()
    mpmath.mp.prec = 200
    assert len(exp_tab) == 128
    for j, (th, tl) in enumerate(exp_tab):
        v = mpmath.power(2, mpmath.mpf(j) / 128)
        assert abs(mpmath.mpf(th) + mpmath.mpf(tl) - v) < mpmath.mpf(2) ** -104
    ln2 = mpmath.log(2)
    assert abs(mpmath.mpf(c["LN2_HI"]) + mpmath.mpf(c["LN2_LO"]) - ln2) < mpmath.mpf(2) ** -95
    assert abs(mpmath.mpf(c["LN2N_HI"]) + mpmath.mpf(c["LN2N_LO"]) - ln2 / 128) < mpmath.mpf(2) ** -94
    # k * LN2_HI exact for |k| < 2^11, n * LN2N_HI exact for |n| < 2^18
    for hi, bits in ((c["LN2_HI"], 42), (c["LN2N_HI"], 35)):
        m, _ = math.frexp(hi)
        assert (m * 2 ** bits).is_integer()


def test_log_dd_accuracy():
    lib = pyoracle.load()
    mpmath.mp.prec = 200
    rng = np.random.default_rng(2)
    xs = np.concatenate([10.0 ** rng.uniform(-307, 307, 1500), rng.uniform(0.98, 1.02, 1500),
                         10.0 ** rng.uniform(-12, 2, 1500),
                         [5e-324, 1e-310, 2.2250738585072014e-308, 1.7976931348623157e308,
                          1.0, 1.0 - 2 ** -53, 1.0 + 2 ** -52, 0.5, 2.0]])
    out = np.zeros(2)
    worst = mpmath.mpf(0)
    for x in xs:
        lib.oracle_log_dd(float(x), out.ctypes.data)
        ex = mpmath.log(mpmath.mpf(float(x)))
        err = abs(mpmath.mpf(out[0]) + mpmath.mpf(out[1]) - ex)
        worst = max(worst, err)
        assert abs(out[1]) <= math.ulp(out[0]) / 2 or out[0] == 0.0
    assert worst < mpmath.mpf(2) ** -68, float(worst)
    lib.oracle_log_dd(1.0, out.ctypes.data)
    assert out[0] == 0.0 and out[1] == 0.0


@pytest.mark.parametrize("y", [7 / 50, 2 / 25, 1 / 3, 0.5, 0.9])
def test_pow_ctrl_accuracy(y):
    lib = pyoracle.load()
    mpmath.mp.prec = 160
    rng = np.random.default_rng(0)
    xs = np.concatenate([10.0 ** rng.uniform(-300, 300, 1500), rng.uniform(0, 2, 1500),
                         10.0 ** rng.uniform(-20, 3, 3000), [5e-324, 2.2250738585072014e-308]])
    worst = 0.0
    for x in xs:
        got = lib.oracle_pow_ctrl(float(x), y)
        ex = mpmath.power(mpmath.mpf(float(x)), mpmath.mpf(y))
        if ex < mpmath.mpf(2) ** -1022:
            continue
        err = abs(float((mpmath.mpf(got) - ex) / math.ulp(float(ex))))
        worst = max(worst, err)
    assert worst < 0.51, worst


def test_pow_ctrl_special_values():
    lib = pyoracle.load()
    assert lib.oracle_pow_ctrl(0.0, 0.14) == 0.0
    assert lib.oracle_pow_ctrl(-0.0, 0.14) == 0.0
    assert lib.oracle_pow_ctrl(1.0, 0.14) == 1.0
    assert lib.oracle_pow_ctrl(1.0, 0.08) == 1.0
    assert lib.oracle_pow_ctrl(math.inf, 0.14) == math.inf
    assert math.isnan(lib.oracle_pow_ctrl(math.nan, 0.14))
    assert math.isnan(lib.oracle_pow_ctrl(-1.0, 0.14))
    assert math.isnan(lib.oracle_pow_ctrl(-math.inf, 0.14))
    # exp_dd saturates outside its range
    assert lib.oracle_exp_dd(710.0, 0.0) == math.inf
    assert lib.oracle_exp_dd(-746.0, 0.0) == 0.0
    assert lib.oracle_exp_dd(0.0, 0.0) == 1.0
    # agrees with the C library pow in the overwhelming majority of cases
    rng = np.random.default_rng(1)
    xs = 10.0 ** rng.uniform(-10, 3, 20000)
    same = np.mean([lib.oracle_pow_ctrl(float(x), 0.14) == math.pow(float(x), 0.14) for x in xs])
    assert same > 0.98


def test_controller_domain_pow_equals_general():
    """The kernels evaluate the controller's pow with the cheaper pow_log_ctrl /
    pow_from_ctrl, valid for x in {0, +Inf, NaN} or normal and 0 < y <= 0.9; there
    it must return pow_ctrl's bits."""
    lib = pyoracle.load()
    rng = np.random.default_rng(5)
    # normal x across the whole exponent range, plus the controller's typical range
    bits = rng.integers(0x0010000000000000, 0x7FF0000000000000, 200000, dtype=np.int64)
    xs = np.concatenate([bits.view(np.float64), 10.0 ** rng.uniform(-20, 3, 100000),
                         [2.2250738585072014e-308, 1.7976931348623157e308, 1.0, 1e-4, 0.5,
                          2.0, 0.0, math.inf]])
    for y in (7 / 50, 2 / 25, 0.5, 0.9, float(np.float32(7 / 50))):
        for x in xs:
            a = lib.oracle_pow_ctrl_domain(float(x), y)
            b = lib.oracle_pow_ctrl(float(x), y)
            assert a == b or (math.isnan(a) and math.isnan(b)), (x, y, a, b)
    assert math.isnan(lib.oracle_pow_ctrl_domain(math.nan, 0.14))
    assert math.isnan(lib.oracle_pow_ctrl_domain(-1.0, 0.14))


def _same(a, b):
    return (math.isnan(a) and math.isnan(b)) or (a == b and math.copysign(1, a) == math.copysign(1, b))


def test_reduced_min_max_equal_julia_on_their_domains():
    """The kernels replace Julia's NaN-propagating min/max by cheaper forms where
    the arguments allow: *_nonneg for absolute values (no -0.0), a plain
    comparison select (fast_*) where neither argument can be NaN."""
    lib = pyoracle.load()
    rng = np.random.default_rng(6)
    special = [0.0, math.inf, math.nan, 1e-300, 5e-324, 1.0, 5.0, 1e-4]
    vals = list(np.abs(rng.standard_normal(2000))) + list(10.0 ** rng.uniform(-20, 5, 2000)) + special
    for x in vals:
        for y in special + [float(v) for v in rng.choice(vals, 20)]:
            x, y = float(x), float(y)
            assert _same(lib.oracle_minmax(2, x, y), lib.oracle_minmax(0, x, y)), (x, y)
            assert _same(lib.oracle_minmax(3, x, y), lib.oracle_minmax(1, x, y)), (x, y)
            if not (math.isnan(x) or math.isnan(y)):
                assert _same(lib.oracle_minmax(4, x, y), lib.oracle_minmax(0, x, y)), (x, y)
                assert _same(lib.oracle_minmax(5, x, y), lib.oracle_minmax(1, x, y)), (x, y)
    # where the -0.0 rule matters, the reduced forms are not used
    assert math.copysign(1, lib.oracle_minmax(1, 0.0, -0.0)) == -1.0
