"""The step-size controller's division and square root without range handling
(Ops<double>::div_r / div_in / root_in), which the adaptive kernels use when the
error sum of squares and dt lie in the checked range: on that domain they must
return the bits of IEEE division and square root.  Run on the device through a
compiled right-hand side (u0 = 0, one Euler step of dt = 1 returns f exactly),
against the device's own `/` and `sqrt` and against numpy's correctly rounded
operations, on random operands over the whole domain and its boundaries."""
import numpy as np
import pytest

from helpers import bits_equal
from simplediffeq_hip import _abi, runtime

pytestmark = pytest.mark.gpu

SRC = """
template <class T>
__device__ void probe(T* du, const T* u, const T* p, const T* __restrict__ ps, double t) {
  typedef sdeh::Ops<T> O;
  du[0] = O::div_in(p[0], p[1]);
  du[1] = p[0] / p[1];
  du[2] = O::div_r(p[0], T(3), O::rcp_nr(T(3)));
  du[3] = p[0] / T(3);
  du[4] = O::div_r(p[0], T(0.9), O::rcp_nr(T(0.9)));
  du[5] = p[0] / T(0.9);
  du[6] = O::root_in(p[2]);
  du[7] = __builtin_sqrt(p[2]);
}
"""


def _operands(n, rng):
    """Numerators |a| in [2^-700, 2^700], divisors |b| in [2^-60, 2^60], roots of x
    in [2^-767, max finite]: the domain the controller's operands stay in."""
    def mag(lo, hi, k):
        e = rng.integers(lo, hi, size=k)
        return np.ldexp(rng.uniform(1.0, 2.0, size=k), e)
    a = mag(-700, 700, n) * rng.choice([-1.0, 1.0], size=n)
    b = mag(-60, 60, n) * rng.choice([-1.0, 1.0], size=n)
    x = mag(-767, 1023, n)
    # boundaries of the domain, exact powers of two, the controller's own values
    edge = np.array([2.0 ** -700, 2.0 ** 700, 2.0 ** 699 * (2 - 2.0 ** -52), 1.0, 3.0, 0.9,
                     0.1, 5.0, 1e-14, 1e-7, 0.47, 1.0 - 2.0 ** -53, 2.0 ** -700 * 1.5])
    k = len(edge)
    a[:k] = edge
    b[:k] = np.clip(edge[::-1], 2.0 ** -60, 2.0 ** 60)
    x[:k] = edge
    x[k:k + 3] = [2.0 ** -767, np.finfo(np.float64).max, 2.0 ** -767 * (2 - 2.0 ** -52)]
    a[k:2 * k] = edge / 3.0 * 2.0
    b[k:2 * k] = 3.0
    # near-halfway quotients: a = b * q for q with a long significand
    q = rng.uniform(1.0, 2.0, size=4096) * 2.0 ** rng.integers(-600, 600, size=4096)
    a[2 * k:2 * k + 4096] = b[2 * k:2 * k + 4096] * q
    a = np.sign(a) * np.clip(np.abs(a), 2.0 ** -700, 2.0 ** 700)  # the checked range
    return a, b, x


def test_fast_division_and_root_are_ieee(ctx):
    n = 1 << 21
    rng = np.random.default_rng(77)
    a, b, x = _operands(n, rng)
    h = ctx.compile_rhs(SRC, "probe", 8, 3)
    d = runtime.make_desc(_abi.ALG_EULER, rhs=h, n_traj=n, tspan=(0.0, 1.0), dt=1.0)
    buf = runtime.EnsembleBuffers(d, 8, np.zeros((8, n)), np.stack([a, b, x]))
    assert ctx.solve(d, buf) == 0
    out = buf.u_end
    for k in (0, 2, 4, 6):
        assert bits_equal(out[k], out[k + 1]), k  # same bits as the device's `/` and sqrt
    with np.errstate(all="ignore"):
        assert bits_equal(out[1], a / b)
        assert bits_equal(out[3], a / 3.0)
        assert bits_equal(out[5], a / 0.9)
        assert bits_equal(out[7], np.sqrt(x))


SRC_NORM = """
template <class T>
__device__ void probe(T* du, const T* u, const T* p, const T* __restrict__ ps, double t) {
  const T bare = sdeh::Ops<T>::div_in(p[0], p[1]), ieee = p[0] / p[1];
  du[0] = bare;
  du[1] = ieee;
  du[2] = bare * bare;
  du[3] = ieee * ieee;
}
"""


def test_error_norm_division_cases(ctx):
    """The error norm's terms e / b with b = m * reltol + abstol in [2^-420, 2^1022)
    (what sdeh_kargs::err_bare guarantees) and e any binary64: the bare quotient is
    the IEEE one where |e / b| is in [2^-540, 2^350]; both square to +0 below; above
    (and for e Inf or NaN) the bare square exceeds 2^700 or is NaN, so the kernels'
    range check on the sum sends the lane to the IEEE norm."""
    n = 1 << 20
    rng = np.random.default_rng(78)
    b = np.ldexp(rng.uniform(1.0, 2.0, n), rng.integers(-420, 1022, n))
    q = np.ldexp(rng.uniform(1.0, 2.0, n), rng.integers(-700, 500, n))
    k = 1 << 16
    with np.errstate(over="ignore"):  # products past the largest binary64 are Inf numerators
        e = b * q * rng.choice([-1.0, 1.0], n)  # quotients across every case and its edges
        e[:k] = np.ldexp(rng.uniform(1.0, 2.0, k), rng.integers(-1074, 1024, k))  # any numerator
        b[k:2 * k] = np.ldexp(rng.uniform(1.0, 2.0, k), rng.integers(1015, 1022, k))  # top
        e[k:2 * k] = b[k:2 * k] * np.ldexp(rng.uniform(1.0, 2.0, k),
                                           rng.integers(-545, 355, k))
    edge = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, np.finfo(np.float64).max,
                     2.0 ** -540, 2.0 ** 350, 2.0 ** -540 * (1 - 2.0 ** -53), 2.0 ** 350 * 1.5])
    e[2 * k:2 * k + len(edge)] = edge
    b[2 * k:2 * k + len(edge)] = 1.0
    b[2 * k + len(edge):2 * k + 2 * len(edge)] = 2.0 ** 1022 * (1 - 2.0 ** -53)
    b[2 * k + 2 * len(edge):2 * k + 3 * len(edge)] = 2.0 ** -420
    h = ctx.compile_rhs(SRC_NORM, "probe", 4, 2)
    d = runtime.make_desc(_abi.ALG_EULER, rhs=h, n_traj=n, tspan=(0.0, 1.0), dt=1.0)
    buf = runtime.EnsembleBuffers(d, 4, np.zeros((4, n)), np.stack([e, b]))
    assert ctx.solve(d, buf) == 0
    bare, ieee, bare2, ieee2 = buf.u_end
    with np.errstate(all="ignore"):
        ref = 0.0 + 1.0 * (e / b)  # the Euler step u0 + dt f with u0 = 0, dt = 1 (-0 -> +0)
    num = ~np.isnan(ref)
    bad = np.flatnonzero(num & (ieee.view(np.int64) != ref.view(np.int64)))
    assert len(bad) == 0 and np.isnan(ieee[~num]).all(), (
        len(bad), [(e[i], b[i], ieee[i], ref[i]) for i in bad[:4]])
    mag = np.abs(ref)
    mid = (mag >= 2.0 ** -540) & (mag <= 2.0 ** 350)
    lo = mag < 2.0 ** -540
    hi = ~(mid | lo)  # above 2^350, Inf or NaN
    assert mid.sum() > n // 4 and lo.sum() > n // 8 and hi.sum() > n // 8
    bad = np.flatnonzero(mid & (bare.view(np.int64) != ieee.view(np.int64)))
    assert len(bad) == 0, (len(bad), [(e[i], b[i], bare[i], ieee[i]) for i in bad[:4]])
    assert (bare2[lo] == 0).all() and (ieee2[lo] == 0).all()
    assert not (bare2[hi] <= 2.0 ** 700).any()
