"""The oracle's GPUSimpleRK4 and GPUSimpleEuler (gpurk4.jl:53-98,
gpueuler.jl:53-90), pinned without the reference: a separately written Python
twin bit for bit, the methods' stability polynomials, exact closed forms of
u' = t that show the end-of-step evaluation time (and its correction),
convergence orders, and the reference's own Lorenz stability run."""
import math

import numpy as np
import pytest

from oracle import pyoracle, pytwin
from simplediffeq_hip import _abi, runtime

RK4, EULER = _abi.ALG_RK4, _abi.ALG_EULER
NAMES = {RK4: "rk4", EULER: "euler"}


def solve(alg, rhs, u0, p, tspan, dt, arith=0, quirks=True, everystep=True, dtype=_abi.F64):
    n = u0.shape[1]
    glen = runtime.fixed_grid(tspan[0], dt, tspan[1])[0]
    d = runtime.make_desc(alg, dtype=dtype, rhs=rhs, n_traj=n, tspan=tspan, dt=dt,
                          save_everystep=everystep, max_saves=max(glen, 1) if everystep else 0,
                          quirks=quirks)
    b = runtime.EnsembleBuffers(d, u0.shape[0], u0, p)
    st = pyoracle.solve(d, b, arith=arith, nthreads=8)
    assert st == 0 and (b.retcode == 0).all()
    return b, glen


@pytest.mark.parametrize("quirks", [True, False], ids=["quirks", "corrected"])
@pytest.mark.parametrize("alg", [RK4, EULER], ids=NAMES.get)
@pytest.mark.parametrize("rhs", [_abi.RHS_LORENZ, _abi.RHS_AFFINE1], ids=["lorenz", "affine1"])
def test_python_twin_every_grid_point_bitwise(rhs, alg, quirks):
    rng = np.random.default_rng(5)
    n = 6
    if rhs == _abi.RHS_LORENZ:
        u0 = rng.uniform(-1.0, 1.0, (3, n))
        p = np.stack([rng.uniform(8, 12, n), rng.uniform(0, 28, n), rng.uniform(2, 3, n)])
        f = pytwin.lorenz
    else:
        u0 = rng.uniform(-1.0, 1.0, (1, n))
        p = np.stack([rng.uniform(-2.0, 0.2, n), rng.uniform(-1.0, 1.0, n)])
        f = pytwin.affine
    dt = 0.03  # not a binary fraction: the grid and stage times round
    b, glen = solve(alg, rhs, u0, p, (0.0, 1.0), dt, arith=pyoracle.ARITH_PLAIN, quirks=quirks)
    twin = pytwin.rk4 if alg == RK4 else pytwin.euler
    for i in range(n):
        grid, us = twin(f, list(u0[:, i]), list(p[:, i]), 0.0, 1.0, dt, quirks)
        assert len(grid) == glen == b.n_steps[i] == 34
        assert list(b.t_steps[:glen, i]) == grid
        for j in range(glen):
            assert list(b.u_steps[j, :, i]) == us[j], (i, j)
        assert b.naccept[i] == glen - 1 and b.t_end[i] == grid[-1]


@pytest.mark.parametrize("alg", [RK4, EULER], ids=NAMES.get)
def test_stability_polynomial(alg):
    # u' = lam u: one step multiplies by R(z), z = lam dt, with
    # R = 1 + z + z^2/2 + z^3/6 + z^4/24 (RK4) or 1 + z (Euler).
    lam = np.array([[-3.0, -1.0, -0.25, 0.5, 1.3]])
    u0 = np.ones((3, lam.shape[1]))
    dt = 0.05
    b, glen = solve(alg, _abi.RHS_EXPDECAY3, u0, lam, (0.0, 2.0), dt)
    z = lam[0] * dt
    R = 1 + z + (z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24 if alg == RK4 else 0)
    for j in (1, 7, glen - 1):
        assert np.allclose(b.u_steps[j, 0], R ** j, rtol=2e-14, atol=0), j


@pytest.mark.parametrize("alg", [RK4, EULER], ids=NAMES.get)
def test_end_of_step_time_closed_form(alg):
    # u' = t, u(0) = 0, dt = 1/8 on (0, 1): every quantity is a short binary
    # fraction, so the results are exact.  Evaluating f from the END of each step
    # (the reference, quirk) gives RK4: u_n = dt^2 n (n + 2) / 2 and Euler:
    # dt^2 n (n + 1) / 2; from its start (corrected) RK4 is exact, dt^2 n^2 / 2,
    # and Euler gives dt^2 n (n - 1) / 2.
    u0 = np.zeros((1, 1))
    p = np.array([[0.0], [1.0]])
    dt, n = 0.125, 8
    for quirks in (True, False):
        b, glen = solve(alg, _abi.RHS_AFFINE1, u0, p, (0.0, 1.0), dt, quirks=quirks)
        assert glen == n + 1
        k = np.arange(n + 1)
        if alg == RK4:
            ex = dt * dt * k * (k + 2) / 2 if quirks else dt * dt * k * k / 2
        else:
            ex = dt * dt * k * (k + 1) / 2 if quirks else dt * dt * k * (k - 1) / 2
        assert np.array_equal(b.u_steps[:, 0, 0], ex), quirks
        assert np.array_equal(b.t_steps[:, 0], k * dt)


def _affine_err(alg, dt, quirks):
    a, c, u0 = -1.0, 1.0, 0.5
    b, _ = solve(alg, _abi.RHS_AFFINE1, np.full((1, 1), u0), np.array([[a], [c]]), (0.0, 2.0),
                 dt, quirks=quirks, everystep=False)
    t = b.t_end[0]
    ex = (u0 + c / a ** 2) * math.exp(a * t) - c * t / a - c / a ** 2
    return abs(b.u_end[0, 0] - ex)


@pytest.mark.parametrize("alg,order", [(RK4, 4), (EULER, 1)], ids=["rk4", "euler"])
def test_convergence_order(alg, order):
    # autonomous u' = -u: the evaluation time does not enter
    errs = []
    for dt in (0.1, 0.05):
        b, _ = solve(alg, _abi.RHS_EXPDECAY3, np.ones((3, 1)), np.full((1, 1), -1.0),
                     (0.0, 2.0), dt, everystep=False)
        errs.append(abs(b.u_end[0, 0] - math.exp(-b.t_end[0])))
    assert order - 0.2 < math.log2(errs[0] / errs[1]) < order + 0.3
    # non-autonomous: full order with the start-of-step time; the reference's
    # end-of-step time drops RK4 to first order
    rate = math.log2(_affine_err(alg, 0.1, False) / _affine_err(alg, 0.05, False))
    assert order - 0.2 < rate < order + 0.3
    rate_q = math.log2(_affine_err(alg, 0.1, True) / _affine_err(alg, 0.05, True))
    assert 0.8 < rate_q < 1.3


@pytest.mark.parametrize("alg", [RK4, EULER], ids=NAMES.get)
def test_lorenz_ten_thousand_steps_stay_finite(alg):
    # simplerk4_tests.jl:44-62 / simpleeuler_tests.jl:44-62: u0 = 10 ones(3),
    # p = (10, 28, 8/3), dt = 0.01, 10,000 steps without NaN.
    u0 = np.full((3, 1), 10.0)
    p = np.array([[10.0], [28.0], [8.0 / 3.0]])
    b, glen = solve(alg, _abi.RHS_LORENZ, u0, p, (0.0, 100.0), 0.01, everystep=False)
    assert glen == 10001 and b.naccept[0] == 10000 and b.t_end[0] == 100.0
    assert np.isfinite(b.u_end).all()
    # and on the attractor, not diverging
    assert np.abs(b.u_end).max() < 60.0


def test_endpoint_matches_last_saved_point():
    u0 = np.full((3, 4), 1.0)
    p = np.array([[10.0] * 4, [28.0] * 4, [8.0 / 3.0] * 4])
    for alg in (RK4, EULER):
        e, _ = solve(alg, _abi.RHS_LORENZ, u0, p, (0.0, 1.0), 0.01, everystep=False)
        s, glen = solve(alg, _abi.RHS_LORENZ, u0, p, (0.0, 1.0), 0.01)
        assert np.array_equal(e.u_end, s.u_steps[glen - 1])
        assert np.array_equal(e.t_end, s.t_steps[glen - 1])


def test_binary32_close_to_binary64():
    u0 = np.full((3, 2), 1.0)
    p = np.array([[10.0] * 2, [28.0] * 2, [8.0 / 3.0] * 2])
    for alg in (RK4, EULER):
        d64, _ = solve(alg, _abi.RHS_LORENZ, u0, p, (0.0, 0.5), 0.01, everystep=False)
        d32, _ = solve(alg, _abi.RHS_LORENZ, u0, p, (0.0, 0.5), 0.01, everystep=False,
                       dtype=_abi.F32)
        assert d32.u_end.dtype == np.float32
        assert np.max(np.abs(d32.u_end - d64.u_end) / np.abs(d64.u_end)) < 1e-4


def test_dt_is_required():
    import simplediffeq_hip as sd
    prob = sd.ODEProblem(lambda u, p, t: -u, 1.0, (0.0, 1.0))
    for alg in (sd.GPUSimpleRK4(), sd.GPUSimpleEuler()):
        with pytest.raises(ValueError, match="dt is required for this algorithm"):
            sd.solve(prob, alg)
