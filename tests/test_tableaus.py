"""The Runge-Kutta coefficient header against the reference text and against
what the mathematics fixes: row sums (sum_j a_ij = c_i), the quadrature and
first tree order conditions of each method and of its embedded partner, and
the dense-output weights at θ = 0 and θ = 1.  A transposed, dropped or mistyped
coefficient breaks at least one of these."""
import os
import sys
from fractions import Fraction

import pytest

from tableaus import ROOT, load, tsit5_method, vern7_method, vern9_method

REF = "/root/reference"
TAB = load()


def test_counts_match_reference_tables():
    # Appendix counts of the reference tables: Tsit5 56, Vern7 58+60+79, Vern9 108+135+153.
    assert len(TAB["tsit5"]) == 56
    assert (len(TAB["vern7"]), len(TAB["vern7x"]), len(TAB["vern7i"])) == (58, 60, 79)
    assert (len(TAB["vern9"]), len(TAB["vern9x"]), len(TAB["vern9i"])) == (108, 135, 153)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")
def test_header_matches_reference_text():
    sys.path.insert(0, os.path.join(ROOT, "simplediffeq.jl_amd", "csrc"))
    import gen_tableaus as gt
    ts = gt.parse_tsit5(os.path.join(REF, "src/tsit5/atsit5_cache.jl"))
    vv = gt.parse_verner(os.path.join(REF, "src/verner/verner_tableaus.jl"))
    pairs = {"tsit5": ts, "vern7": vv["Vern7Tableau"], "vern7x": vv["Vern7ExtraStages"],
             "vern7i": vv["Vern7InterpolationCoefficients"], "vern9": vv["Vern9Tableau"],
             "vern9x": vv["Vern9ExtraStages"], "vern9i": vv["Vern9InterpolationCoefficients"]}
    for ns, lst in pairs.items():
        assert len(lst) == len(TAB[ns])
        for name, lit in lst:
            assert TAB[ns][name] == float(lit), (ns, name)


METHODS = {"tsit5": (tsit5_method, 5), "vern7": (vern7_method, 7), "vern9": (vern9_method, 9)}


@pytest.mark.parametrize("name", list(METHODS))
def test_row_sums_equal_nodes(name):
    c, a, b, bt, s = METHODS[name][0]()
    for i in range(2, s + 1):
        row = sum(v for (ii, j), v in a.items() if ii == i)
        assert abs(row - c[i]) < 5e-15 * max(1.0, max(abs(v) for (ii, j), v in a.items() if ii == i)), (name, i)


@pytest.mark.parametrize("name", list(METHODS))
def test_weights_sum_to_one_and_error_weights_to_zero(name):
    c, a, b, bt, s = METHODS[name][0]()
    assert abs(sum(b.values()) - 1.0) < 1e-14
    assert abs(sum(bt.values())) < 1e-14


@pytest.mark.parametrize("name", list(METHODS))
def test_quadrature_conditions(name):
    # sum_i b_i c_i^k = 1/(k+1) for k < order; the embedded method b - btilde
    # of order p-1 (Tsit5 4, Vern7 6, Vern9 8) for k < p - 1.
    mk, p = METHODS[name]
    c, a, b, bt, s = mk()
    bh = {j: b[j] - bt[j] for j in b}
    for k in range(p):
        q = sum(b[i] * c[i] ** k for i in b)
        assert abs(q - 1.0 / (k + 1)) < 1e-13, (name, k, q)
    for k in range(p - 1):
        q = sum(bh[i] * c[i] ** k for i in bh)
        assert abs(q - 1.0 / (k + 1)) < 1e-13, (name, "embedded", k, q)


@pytest.mark.parametrize("name", list(METHODS))
def test_bushy_tree_conditions(name):
    # sum_i b_i c_i^m sum_j a_ij c_j^k = 1/((k+1)(k+m+2)) for k + m + 2 <= order
    mk, p = METHODS[name]
    c, a, b, bt, s = mk()
    for k in range(p - 1):
        for m in range(p - 1 - k):
            q = 0.0
            for i in range(1, s + 1):
                inner = sum(v * c[j] ** k for (ii, j), v in a.items() if ii == i)
                q += b[i] * c[i] ** m * inner
            exact = 1.0 / ((k + 1) * (k + m + 2))
            assert abs(q - exact) < 1e-12, (name, k, m, q, exact)


def test_stability_polynomial_matches_exponential_to_order():
    # One step on u' = lambda u multiplies u by R(z) = 1 + sum_k (b^T A^(k-1) 1) z^k.
    # Computed exactly (rationals of the binary64 coefficients): the z^k
    # coefficients equal 1/k! up to the decimal rounding of the published
    # coefficients for k <= p, and differ at k = p + 1 (the method is not of
    # order p + 1).
    import math
    from fractions import Fraction as F
    for name, (mk, p) in METHODS.items():
        c, a, b, bt, s = mk()
        v = {i: F(1) for i in range(1, s + 1)}
        for k in range(1, p + 2):
            coef = sum(F(b[i]) * v[i] for i in b)
            err = abs(float(coef - F(1, math.factorial(k))))
            if k <= p:
                assert err < 1e-14, (name, k, err)
            else:
                assert err > 1e-10, (name, k, err)
            v = {i: sum((F(val) * v[j] for (ii, j), val in a.items() if ii == i), F(0))
                 for i in range(1, s + 1)}


def _dense_weights(which, th):
    from oracle import pyoracle
    return pyoracle.btheta(which, th)


def test_dense_output_endpoints():
    # bθs(0) = 0 and bθs(1) = b (the interpolant meets both ends of the step)
    for which, mk, slots in ((0, tsit5_method, [1, 2, 3, 4, 5, 6, 7]),
                             (1, vern7_method, [1, 4, 5, 6, 7, 8, 9]),
                             (2, vern9_method, [1, 8, 9, 10, 11, 12, 13, 14, 15])):
        c, a, b, bt, s = mk()
        w0 = _dense_weights(which, 0.0)
        assert max(abs(x) for x in w0) == 0.0
        w1 = _dense_weights(which, 1.0)
        for idx, j in enumerate(slots):
            assert abs(w1[idx] - b[j]) < 2e-12, (which, j, w1[idx], b[j])
        # the extra stages carry no weight at θ = 1
        for x in w1[len(slots):]:
            assert abs(x) < 2e-12
