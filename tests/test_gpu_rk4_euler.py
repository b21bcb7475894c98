"""GPUSimpleRK4 and GPUSimpleEuler (gpurk4.jl:53-98, gpueuler.jl:53-90) on the
GPU against the CPU oracle, bit for bit: endpoint and every grid point, both
evaluation-time conventions (the reference's end-of-step time and the
corrected start-of-step time), both save layouts, binary64 and binary32; and
the package's solve() surface (dt required, every grid point returned)."""
import numpy as np
import pytest

import simplediffeq_hip as sd
from helpers import assert_bit_identical, bits_equal, problem, run_pair
from simplediffeq_hip import _abi, runtime

pytestmark = pytest.mark.gpu

ALGS = [_abi.ALG_RK4, _abi.ALG_EULER]
IDS = ["rk4", "euler"]
RHS = [_abi.RHS_LORENZ, _abi.RHS_EXPDECAY3, _abi.RHS_AFFINE1, _abi.RHS_LINEAR2]
RHS_IDS = ["lorenz", "expdecay3", "affine1", "linear2"]


@pytest.mark.parametrize("quirks", [True, False], ids=["quirks", "corrected"])
@pytest.mark.parametrize("alg", ALGS, ids=IDS)
@pytest.mark.parametrize("rhs", RHS, ids=RHS_IDS)
def test_endpoint_bitexact(ctx, rhs, alg, quirks):
    n = 777
    u0, p, ps = problem(rhs, n, seed=20 + alg + 10 * rhs)
    d = runtime.make_desc(alg, rhs=rhs, n_traj=n, tspan=(0.0, 2.0), dt=0.013, quirks=quirks)
    st_g, g, st_c, c = run_pair(ctx, d, u0, p, ps)
    assert st_g == st_c == 0
    assert_bit_identical(g, c)
    assert (g.naccept == runtime.fixed_grid(0.0, 0.013, 2.0)[0] - 1).all()


@pytest.mark.parametrize("layout", [_abi.SAVE_POINT_MAJOR, _abi.SAVE_TRAJ_MAJOR],
                         ids=["point-major", "traj-major"])
@pytest.mark.parametrize("quirks", [True, False], ids=["quirks", "corrected"])
@pytest.mark.parametrize("alg", ALGS, ids=IDS)
@pytest.mark.parametrize("rhs", [_abi.RHS_LORENZ, _abi.RHS_AFFINE1], ids=["lorenz", "affine1"])
def test_every_grid_point_bitexact(ctx, rhs, alg, quirks, layout):
    n = 300
    u0, p, ps = problem(rhs, n, seed=31 + alg)
    glen = runtime.fixed_grid(0.0, 0.03, 1.0)[0]
    d = runtime.make_desc(alg, rhs=rhs, n_traj=n, tspan=(0.0, 1.0), dt=0.03,
                          save_everystep=True, max_saves=glen, quirks=quirks, save_layout=layout)
    st_g, g, st_c, c = run_pair(ctx, d, u0, p, ps)
    assert st_g == st_c == 0
    assert_bit_identical(g, c, fields=("u_end", "t_end", "retcode", "naccept", "nreject",
                                       "n_steps", "t_steps", "u_steps"))
    assert (g.n_steps == glen).all()


@pytest.mark.parametrize("alg", ALGS, ids=IDS)
def test_binary32_bitexact(ctx, alg):
    n = 513
    u0, p, _ = problem(_abi.RHS_LORENZ, n, seed=40)
    d = runtime.make_desc(alg, dtype=_abi.F32, rhs=_abi.RHS_LORENZ, n_traj=n, tspan=(0.0, 1.0),
                          dt=float(np.float32(0.01)))
    st_g, g, st_c, c = run_pair(ctx, d, u0, p)
    assert st_g == st_c == 0
    assert g.u_end.dtype == np.float32
    assert_bit_identical(g, c)


def test_save_overflow_and_saveat_rejected(ctx):
    n = 8
    u0, p, _ = problem(_abi.RHS_LORENZ, n)
    d = runtime.make_desc(_abi.ALG_RK4, rhs=_abi.RHS_LORENZ, n_traj=n, tspan=(0.0, 1.0), dt=0.1,
                          save_everystep=True, max_saves=5)
    st_g, g, st_c, c = run_pair(ctx, d, u0, p)
    assert st_g == st_c == _abi.E_SAVE_OVERFLOW
    assert (g.retcode == _abi.RET_SAVE_OVERFLOW).all()
    assert_bit_identical(g, c, fields=("u_end", "t_end", "retcode", "n_steps"))
    d = runtime.make_desc(_abi.ALG_EULER, rhs=_abi.RHS_LORENZ, n_traj=n, tspan=(0.0, 1.0), dt=0.1,
                          saveat=np.array([0.0, 0.5]))
    with pytest.raises(runtime.SolveError, match="no dense output"):
        ctx.solve(d, runtime.EnsembleBuffers(d, 3, u0, p))


def lorenz(u, p, t):
    return [p[0] * (u[1] - u[0]), u[0] * (p[1] - u[2]) - u[1], u[0] * u[1] - p[2] * u[2]]


@pytest.mark.parametrize("alg", [sd.GPUSimpleRK4(), sd.GPUSimpleEuler()], ids=IDS)
def test_solve_returns_every_grid_point(alg):
    # the reference's Lorenz problem of simplerk4_tests.jl:92-100, 1 s of it
    prob = sd.ODEProblem(lorenz, 10.0 * np.ones(3), (0.0, 1.0), np.array([10.0, 28.0, 8 / 3]))
    sol = sd.solve(prob, alg, dt=0.01)
    assert sol.retcode == "Default" and len(sol.t) == len(sol.u) == 101
    assert sol.t[0] == 0.0 and sol.t[-1] == 1.0
    assert bits_equal(sol.u[0], 10.0 * np.ones(3))
    # save_everystep=false and saveat are ignored, as the reference's kwargs...
    for kw in ({"save_everystep": False}, {"saveat": [0.0, 0.5]}):
        s2 = sd.solve(prob, alg, dt=0.01, **kw)
        assert len(s2.u) == 101 and bits_equal(np.array(s2.u), np.array(sol.u))
    # a traced right-hand side (hipRTC) gives the ahead-of-time kernel's bits
    d = runtime.make_desc(alg.code, rhs=_abi.RHS_LORENZ, n_traj=1, tspan=(0.0, 1.0), dt=0.01)
    b = runtime.EnsembleBuffers(d, 3, 10.0 * np.ones((3, 1)), np.array([[10.0], [28.0],
                                                                         [8 / 3]]))
    sd.problem._context(0).solve(d, b)
    assert bits_equal(sol.u[-1], b.u_end[:, 0])


def test_scalar_ensemble():
    prob = sd.ODEProblem(lambda u, p, t: p * u, 1.0, (0.0, 1.0), -1.0)
    eprob = sd.EnsembleProblem(prob, prob_func=lambda q, i, r: sd.remake(q, p=-0.5 * i))
    esol = sd.solve(eprob, sd.GPUSimpleRK4(), sd.EnsembleHIP(), trajectories=4, dt=0.125)
    for i, s in enumerate(esol.u):
        assert len(s.t) == 9 and np.isscalar(s.u[-1])
        z = -0.5 * (i + 1) * 0.125
        assert abs(s.u[-1] - (1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24) ** 8) < 1e-14
