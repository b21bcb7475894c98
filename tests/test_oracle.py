"""Pins of the CPU oracle (the checker of the HIP kernels) against what does not
depend on it: a second restatement written separately (the Python twin), closed
forms, convergence orders, high-accuracy fixtures, and the controller behaviour
the reference's code fixes.  Also measures how far the two readings of the
reference's arithmetic (fused @muladd vs separately rounded; portable pow vs
libm pow) move results, against the parity bars (fixed step 1e-12 relative;
adaptive: identical accepted-step counts on >= 99.9% of trajectories)."""
import json
import math
import os

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import pyoracle, pytwin
from simplediffeq_hip import _abi, inputs, runtime

HERE = os.path.dirname(os.path.abspath(__file__))
F32_01 = float(np.float32(0.1))


def solve(alg, rhs, u0, p, ps=None, tspan=(0.0, 1.0), dt=0.1, abstol=1e-6, reltol=1e-3,
          arith=0, pow_kind=0, tspan_each=None, dt0=None, **kw):
    n = u0.shape[1]
    d = runtime.make_desc(alg, rhs=rhs, n_traj=n, tspan=tspan, dt=dt, abstol=abstol,
                          reltol=reltol, **kw)
    b = runtime.EnsembleBuffers(d, u0.shape[0], u0, p, ps, tspan_each, dt0)
    st = pyoracle.solve(d, b, arith=arith, pow_kind=pow_kind, nthreads=8)
    return st, b


# ---- the Python twin (separately written restatement) ----

@pytest.mark.parametrize("alg,fn", [(_abi.ALG_ATSIT5, pytwin.atsit5), (_abi.ALG_AVERN7, pytwin.avern7),
                                    (_abi.ALG_AVERN9, pytwin.avern9)], ids=["ATsit5", "AVern7", "AVern9"])
def test_python_twin_matches_oracle_bitwise(alg, fn):
    n = 48
    u0, p = inputs.lorenz_wide(n, seed=1)
    st, b = solve(alg, _abi.RHS_LORENZ, u0, p, tspan=(0.0, 10.0), dt=F32_01, abstol=1e-6,
                  reltol=1e-3, arith=pyoracle.ARITH_PLAIN, pow_kind=pyoracle.POW_LIBM)
    assert st == 0
    for i in range(n):
        u, t, rc, na, nr = fn(pytwin.lorenz, list(u0[:, i]), list(p[:, i]), 0.0, 10.0, F32_01,
                              1e-6, 1e-3)
        assert (rc, na, nr) == (b.retcode[i], b.naccept[i], b.nreject[i]), i
        assert t == b.t_end[i] and u == list(b.u_end[:, i]), i


def test_python_twin_fixed_tsit5_bitwise():
    n = 32
    u0, p = inputs.lorenz_small(n, seed=0)
    st, b = solve(_abi.ALG_TSIT5, _abi.RHS_LORENZ, u0, p, dt=0.01, arith=pyoracle.ARITH_PLAIN)
    for i in range(n):
        u, t, ns = pytwin.tsit5_fixed(pytwin.lorenz, list(u0[:, i]), list(p[:, i]), 0.0, 1.0, 0.01)
        assert ns == 100 and u == list(b.u_end[:, i]) and t == b.t_end[i]


# ---- the two arithmetic readings against the parity bars ----

def _rel(a, b):
    """Largest per-trajectory normwise relative difference max_c|a-b| / max_c|b|
    (arrays shaped (n_state, n_traj))."""
    return np.max(np.max(np.abs(a - b), axis=0) / np.maximum(np.max(np.abs(b), axis=0), 1e-300))


def test_fused_vs_plain_fixed_step_within_1e12():
    u0, p = inputs.lorenz_small(1024, seed=0)
    _, f = solve(_abi.ALG_TSIT5, _abi.RHS_LORENZ, u0, p, dt=0.01, arith=0)
    _, q = solve(_abi.ALG_TSIT5, _abi.RHS_LORENZ, u0, p, dt=0.01, arith=1)
    assert _rel(f.u_end, q.u_end) < 1e-12
    assert np.array_equal(f.t_end, q.t_end)


@pytest.mark.parametrize("variant", [(1, 0), (0, 1), (1, 1)], ids=["plain", "libm-pow", "plain+libm"])
def test_adaptive_readings_keep_step_counts(variant):
    arith, pw = variant
    u0, p = inputs.lorenz_small(1024, seed=0)
    _, f = solve(_abi.ALG_ATSIT5, _abi.RHS_LORENZ, u0, p, dt=0.1, abstol=1e-6, reltol=1e-3)
    _, q = solve(_abi.ALG_ATSIT5, _abi.RHS_LORENZ, u0, p, dt=0.1, abstol=1e-6, reltol=1e-3,
                 arith=arith, pow_kind=pw)
    same = np.mean((f.naccept == q.naccept) & (f.nreject == q.nreject))
    assert same >= 0.999
    # endpoints agree far inside the requested tolerance
    sc = 1e-6 + 1e-3 * np.abs(q.u_end)
    assert np.max(np.abs(f.u_end - q.u_end) / sc) < 1e-3


def test_wide_ensemble_readings_keep_step_counts():
    u0, p = inputs.lorenz_wide(8192, seed=1)
    kw = dict(tspan=(0.0, 10.0), dt=F32_01, abstol=1e-6, reltol=1e-3)
    _, f = solve(_abi.ALG_ATSIT5, _abi.RHS_LORENZ, u0, p, **kw)
    _, q = solve(_abi.ALG_ATSIT5, _abi.RHS_LORENZ, u0, p, arith=1, pow_kind=1, **kw)
    assert np.mean(f.naccept == q.naccept) >= 0.999


# ---- closed forms ----

@pytest.mark.parametrize("alg", range(6))
def test_expdecay_closed_form(alg):
    n = 16
    rng = np.random.default_rng(alg)
    u0 = rng.uniform(0.5, 2.0, (3, n))
    lam = rng.uniform(-2.0, 0.5, (1, n))
    adaptive = alg in _abi.ADAPTIVE_ALGS
    st, b = solve(alg, _abi.RHS_EXPDECAY3, u0, lam, tspan=(0.0, 2.0),
                  dt=0.1 if adaptive else 0.01, abstol=1e-11, reltol=1e-11)
    assert st == 0 and (b.retcode == 0).all()
    ex = u0 * np.exp(lam * b.t_end)
    tol = {0: 1e-10, 1: 1e-9, 2: 1e-13, 3: 1e-9, 4: 1e-13, 5: 1e-9}[alg]
    assert np.max(np.abs(b.u_end - ex) / np.abs(ex)) < tol


@pytest.mark.parametrize("alg", range(6))
def test_affine_nonautonomous_closed_form(alg):
    # u' = a u + c t, u(0) = u0:  u(t) = (u0 + c/a^2) e^{at} - c t / a - c / a^2
    n = 16
    rng = np.random.default_rng(10 + alg)
    u0 = rng.uniform(-1.0, 1.0, (1, n))
    a = rng.uniform(-2.0, -0.2, n)
    c = rng.uniform(-1.0, 1.0, n)
    adaptive = alg in _abi.ADAPTIVE_ALGS
    st, b = solve(alg, _abi.RHS_AFFINE1, u0, np.stack([a, c]), tspan=(0.0, 1.5),
                  dt=0.1 if adaptive else 0.01, abstol=1e-11, reltol=1e-11)
    t = b.t_end
    ex = (u0[0] + c / a ** 2) * np.exp(a * t) - c * t / a - c / a ** 2
    assert np.max(np.abs(b.u_end[0] - ex)) < 1e-9


def test_linear_system_matrix_exponential():
    # The 2x2 system of test/simpleatsit5_tests.jl:87-112, M = [2 3; 4 5], scaled
    # to keep the solution moderate, against expm.
    M = np.array([[2.0, 3.0], [4.0, 5.0]]) * 0.1
    u0 = np.array([[1.0], [0.5]])
    p = M.reshape(4, 1)
    for alg in range(6):
        adaptive = alg in _abi.ADAPTIVE_ALGS
        st, b = solve(alg, _abi.RHS_LINEAR2, u0, p, tspan=(0.0, 1.0),
                      dt=0.1 if adaptive else 0.01, abstol=1e-12, reltol=1e-12)
        ex = expm(M * b.t_end[0]) @ u0[:, 0]
        assert np.max(np.abs(b.u_end[:, 0] - ex) / np.abs(ex)) < 2e-10, alg


@pytest.mark.parametrize("alg,order,dts", [(_abi.ALG_TSIT5, 5, (0.5, 0.25)),
                                           (_abi.ALG_VERN7, 7, (0.5, 0.25)),
                                           (_abi.ALG_VERN9, 9, (1.0, 0.5))])
def test_fixed_step_convergence_order(alg, order, dts):
    # u' = -u on (0, 8); step pairs chosen above the rounding floor (Vern9 reaches
    # 1e-13 relative already at dt = 0.5).
    u0 = np.ones((3, 1))
    lam = np.full((1, 1), -1.0)
    errs = []
    for dt in dts:
        _, b = solve(alg, _abi.RHS_EXPDECAY3, u0, lam, tspan=(0.0, 8.0), dt=dt)
        errs.append(abs(b.u_end[0, 0] - math.exp(-8.0)))
    rate = math.log2(errs[0] / errs[1])
    assert order - 0.7 < rate < order + 1.5, rate


def test_dense_output_closed_form_and_orders():
    # saveat values of u' = -u against e^{-t}; the interpolants are 4th (Tsit5),
    # 7th (Vern7) and 9th (Vern9) order accurate.
    u0 = np.ones((3, 1))
    lam = np.full((1, 1), -1.0)
    saveat = np.linspace(0.0, 1.0, 37)
    for alg, bound in ((_abi.ALG_TSIT5, 3e-7), (_abi.ALG_VERN7, 1e-11), (_abi.ALG_VERN9, 1e-13)):
        _, b = solve(alg, _abi.RHS_EXPDECAY3, u0, lam, dt=0.05, saveat=saveat, quirks=False)
        err = np.max(np.abs(b.u_saveat[:, 0, 0] - np.exp(-saveat)))
        assert err < bound, (alg, err)


def test_fixed_vern9_saveat_quirk_is_reproduced():
    # Reference quirk (gpuvern9.jl:216-331): the fixed-step Vern9 dense output
    # multiplies k2..k9 where the renamed k8..k15 belong.  With the quirk the
    # interpolant is only O(dt^2) accurate (the reference's own test bar for it
    # is 2e-4 at dt = 0.01, test/gpu_ode_regression.jl:57); corrected it is
    # accurate to rounding.
    u0 = np.ones((3, 1))
    lam = np.full((1, 1), -1.0)
    saveat = np.array([0.0, 0.4, 0.55, 0.9])
    _, q = solve(_abi.ALG_VERN9, _abi.RHS_EXPDECAY3, u0, lam, dt=0.01, saveat=saveat, quirks=True)
    _, c = solve(_abi.ALG_VERN9, _abi.RHS_EXPDECAY3, u0, lam, dt=0.01, saveat=saveat, quirks=False)
    ex = np.exp(-saveat)
    eq = np.max(np.abs(q.u_saveat[:, 0, 0] - ex))
    ec = np.max(np.abs(c.u_saveat[:, 0, 0] - ex))
    assert ec < 1e-14
    assert 1e-9 < eq < 2e-4


# ---- high-accuracy fixtures (DOP853 at 1e-13) ----

def _fixtures():
    with open(os.path.join(HERE, "golden", "dop853_fixtures.json")) as f:
        return json.load(f)


def test_lorenz_small_against_fixture():
    fx = _fixtures()["lorenz_small"]
    u0, p = inputs.lorenz_small(32, seed=0)
    ref = np.array(fx["u_end"]).T
    _, b = solve(_abi.ALG_AVERN9, _abi.RHS_LORENZ, u0, p, dt=0.1, abstol=1e-13, reltol=1e-13)
    assert np.max(np.abs(b.u_end - ref)) < 1e-10
    _, b = solve(_abi.ALG_ATSIT5, _abi.RHS_LORENZ, u0, p, dt=0.1, abstol=1e-6, reltol=1e-3)
    # global error of a local-error controller: a modest multiple of the
    # per-step tolerance scale (normwise per trajectory)
    sc = 1e-6 + 1e-3 * np.max(np.abs(ref), axis=0)
    assert np.max(np.max(np.abs(b.u_end - ref), axis=0) / sc) < 20
    _, b = solve(_abi.ALG_TSIT5, _abi.RHS_LORENZ, u0, p, dt=0.01)
    assert np.max(np.abs(b.u_end - ref)) < 1e-6


def test_lorenz_wide_against_fixture():
    fx = _fixtures()["lorenz_wide"]
    u0, p = inputs.lorenz_wide(16, seed=1)
    ref = np.array(fx["u_end"]).T
    for alg, tol in ((_abi.ALG_AVERN7, 1e-10), (_abi.ALG_AVERN9, 1e-10)):
        _, b = solve(alg, _abi.RHS_LORENZ, u0, p, tspan=(0.0, 10.0), dt=F32_01, abstol=1e-12,
                     reltol=1e-12)
        assert np.max(np.abs(b.u_end - ref)) < tol * 100, alg


def test_reference_lorenz_saveat_against_fixture():
    # test/gpusimpleatsit5_tests.jl:73-82: Lorenz u0 = 10 ones, tspan (0, 10), tol 1e-9,
    # agrees with Tsit5 to atol 1e-5; here with saveat 0:0.1:10 against DOP853.
    fx = _fixtures()["lorenz_reference"]
    u0 = np.array(fx["u0"]).reshape(3, 1)
    p = np.array(fx["p"]).reshape(3, 1)
    _, b = solve(_abi.ALG_ATSIT5, _abi.RHS_LORENZ, u0, p, tspan=(0.0, 10.0), dt=F32_01,
                 abstol=1e-9, reltol=1e-9, saveat=fx["saveat"])
    ref = np.array(fx["u"])
    assert np.max(np.abs(b.u_saveat[:, :, 0] - ref)) < 1e-4
    assert b.t_end[0] == 10.0


def test_glv16_against_fixture():
    fx = _fixtures()["glv16"]
    u0, r, A = inputs.glv16(4, seed=3)
    _, b = solve(_abi.ALG_AVERN9, _abi.RHS_GLV16, u0, r, A, tspan=(0.0, 50.0), dt=0.1,
                 abstol=1e-10, reltol=1e-10, saveat=fx["saveat"])
    ref = np.array(fx["u"]).transpose(1, 2, 0)  # (points, species, traj)
    assert np.max(np.abs(b.u_saveat - ref)) < 1e-7


# ---- controller behaviour fixed by the reference code ----

def test_zero_error_grows_dt_ninefold():
    # EEst == 0 gives q = max(1/qmax, min(1/qmin, (1/qmax)/gamma)) = 1/9, so dt
    # grows 9x (not 10x): from dt0 = 0.1 on (0, 1) the second step lands on tf.
    u0 = np.ones((3, 1))
    st, b = solve(_abi.ALG_ATSIT5, _abi.RHS_EXPDECAY3, u0, np.zeros((1, 1)), dt=0.1)
    assert st == 0 and b.naccept[0] == 2 and b.nreject[0] == 0 and b.t_end[0] == 1.0
    st, b = solve(_abi.ALG_ATSIT5, _abi.RHS_EXPDECAY3, u0, np.zeros((1, 1)), dt=0.01)
    assert b.naccept[0] == 4  # steps 0.01, 0.09, 0.81, then the remaining 0.09


def test_rejections_shrink_dt_and_recover():
    u0 = np.ones((3, 1))
    st, b = solve(_abi.ALG_ATSIT5, _abi.RHS_EXPDECAY3, u0, np.full((1, 1), -30.0), dt=1.0,
                  abstol=1e-8, reltol=1e-8)
    assert st == 0 and b.nreject[0] >= 2
    assert abs(b.u_end[0, 0] - math.exp(-30.0)) < 1e-8


def test_dtmin_is_a_per_trajectory_retcode():
    u0 = np.ones((3, 2))
    lam = np.full((1, 2), -1.0)
    st, b = solve(_abi.ALG_ATSIT5, _abi.RHS_EXPDECAY3, u0, lam, dt0=np.array([1e-15, 0.1]))
    assert st == _abi.E_DTMIN
    assert list(b.retcode) == [_abi.RET_DTLESSTHANMIN, _abi.RET_DEFAULT]
    assert b.t_end[0] == 0.0 and np.array_equal(b.u_end[:, 0], u0[:, 0])
    # AVern9's dtmin is the Float32 literal 1f-7 (gpuvern9.jl:466): 5e-8 already fails
    st, b = solve(_abi.ALG_AVERN9, _abi.RHS_EXPDECAY3, u0, lam, dt0=np.array([5e-8, 0.1]))
    assert list(b.retcode) == [_abi.RET_DTLESSTHANMIN, _abi.RET_DEFAULT]
    st, b = solve(_abi.ALG_ATSIT5, _abi.RHS_EXPDECAY3, u0, lam, dt0=np.array([5e-8, 0.1]))
    assert list(b.retcode) == [_abi.RET_DEFAULT, _abi.RET_DEFAULT]


def test_nan_state_ends_with_nan_retcode():
    u0 = np.ones((3, 2))
    u0[1, 0] = np.nan
    for alg in _abi.ADAPTIVE_ALGS:
        st, b = solve(alg, _abi.RHS_EXPDECAY3, u0, np.full((1, 2), -1.0))
        assert list(b.retcode) == [_abi.RET_NAN, _abi.RET_DEFAULT]
        assert b.naccept[0] == 0


def test_empty_span_and_maxiters():
    u0 = np.ones((3, 1))
    st, b = solve(_abi.ALG_ATSIT5, _abi.RHS_EXPDECAY3, u0, np.full((1, 1), -1.0), tspan=(0.5, 0.5))
    assert st == 0 and b.naccept[0] == 0 and b.t_end[0] == 0.5
    assert np.array_equal(b.u_end, u0)
    st, b = solve(_abi.ALG_ATSIT5, _abi.RHS_EXPDECAY3, u0, np.full((1, 1), -1.0), dt=1e-3,
                  maxiters=3)
    assert b.retcode[0] == _abi.RET_MAXITERS and b.naccept[0] + b.nreject[0] == 3


def test_fixed_step_takes_no_partial_final_step():
    # gpuatsit5.jl:92-95: steps over t0:dt:tf only; with dt = 0.3 on (0, 1) the
    # grid is 0, 0.3, 0.6, 0.9 and the solve stops at t = 0.9 (+ rounding).
    u0 = np.ones((3, 1))
    st, b = solve(_abi.ALG_TSIT5, _abi.RHS_EXPDECAY3, u0, np.full((1, 1), -1.0), dt=0.3)
    assert b.naccept[0] == 3 and abs(b.t_end[0] - 0.9) < 1e-15


def test_save_everystep_records_each_step():
    u0, p = inputs.lorenz_small(8, seed=0)
    st, b = solve(_abi.ALG_ATSIT5, _abi.RHS_LORENZ, u0, p, save_everystep=True, max_saves=64)
    _, e = solve(_abi.ALG_ATSIT5, _abi.RHS_LORENZ, u0, p)
    assert np.array_equal(b.n_steps, b.naccept + 1)
    for i in range(8):
        k = b.n_steps[i]
        assert b.t_steps[0, i] == 0.0 and b.t_steps[k - 1, i] == 1.0
        assert np.all(np.diff(b.t_steps[:k, i]) > 0)
        assert np.array_equal(b.u_steps[k - 1, :, i], e.u_end[:, i])
    # same final t with and without save_everystep (gpusimpleatsit5_tests.jl:38-39)
    assert np.array_equal(b.t_end, e.t_end)
    st, b = solve(_abi.ALG_ATSIT5, _abi.RHS_LORENZ, u0, p, save_everystep=True, max_saves=4)
    assert st == _abi.E_SAVE_OVERFLOW and (b.retcode == _abi.RET_SAVE_OVERFLOW).all()
