"""The reference's own tests for these solvers, run through this package's
solve() on the GPU.

test/gpu_ode_regression.jl: u' = -u, Float32 u0 = ones(3), tspan (0, 1), every
GPUSimple solver pair; bars against an OrdinaryDiffEq Vern9 solution, here
against the exact solution exp(-t) u0 (what that Vern9 solution approximates).
test/gpusimpleatsit5_tests.jl: Lorenz (10, 28, 8/3), u0 = 10 ones; fixed and
adaptive Tsit5 consistency between save modes.  Plus the traced-RHS path
(hipRTC) against the ahead-of-time kernel, and EnsembleProblem semantics.
"""
import numpy as np
import pytest

import simplediffeq_hip as sd
from helpers import bits_equal

pytestmark = pytest.mark.gpu

F = np.float32


def neg(u, p, t):
    return [-u[0], -u[1], -u[2]]


PAIRS = [(sd.GPUSimpleATsit5(), sd.GPUSimpleTsit5()), (sd.GPUSimpleAVern7(), sd.GPUSimpleVern7()),
         (sd.GPUSimpleAVern9(), sd.GPUSimpleVern9())]


@pytest.mark.parametrize("adaptive_alg,fixed_alg", PAIRS, ids=["Tsit5", "Vern7", "Vern9"])
def test_gpu_ode_regression(adaptive_alg, fixed_alg):
    u0 = np.ones(3, dtype=F)
    prob = sd.ODEProblem(neg, u0, (0.0, 1.0), np.array([10, 28, 8 / 3], dtype=F))
    exact = lambda t: np.exp(-np.asarray(t, dtype=np.float64))[..., None] * np.ones(3)

    sol = sd.solve(prob, fixed_alg, dt=float(F(0.01)))
    asol = sd.solve(prob, adaptive_alg, dt=float(F(0.01)), abstol=float(F(1e-7)),
                    reltol=float(F(1e-7)), save_everystep=False)
    assert sol.retcode == "Default" and asol.retcode == "Default"
    assert np.linalg.norm(exact(1.0) - sol.u[-1]) < 5e-3
    assert np.linalg.norm(exact(asol.t) - np.array(asol.u)) < 5e-4

    saveat = [0.0, 0.4]
    sol = sd.solve(prob, fixed_alg, dt=float(F(0.01)), saveat=saveat)
    asol = sd.solve(prob, adaptive_alg, dt=float(F(0.01)), abstol=float(F(1e-7)),
                    reltol=float(F(1e-7)), saveat=saveat)
    assert np.linalg.norm(asol.u[-1] - sol.u[-1]) < 5e-3
    assert np.linalg.norm(exact(saveat) - np.array(sol.u)) < 2e-4
    assert np.linalg.norm(exact(saveat) - np.array(asol.u)) < 2e-4
    assert len(sol.u) == len(saveat) and len(asol.u) == len(saveat)

    saveat = np.arange(101, dtype=F) * F(0.01)
    sol = sd.solve(prob, fixed_alg, dt=float(F(0.01)), saveat=saveat)
    asol = sd.solve(prob, adaptive_alg, dt=float(F(0.01)), save_everystep=False,
                    abstol=float(F(1e-7)), reltol=float(F(1e-7)), saveat=saveat)
    reached = ~np.isnan(np.array(sol.u)[:, 0])
    assert np.linalg.norm(exact(saveat)[reached] - np.array(sol.u)[reached]) < 2e-3
    assert np.linalg.norm(exact(saveat) - np.array(asol.u)) < 3e-3
    assert len(sol.u) == len(saveat) and len(asol.u) == len(saveat)

    sol = sd.solve(prob, fixed_alg, dt=float(F(0.01)), save_everystep=False)
    assert len(sol.u) == 2 and np.linalg.norm(exact(sol.t) - np.array(sol.u)) < 5e-3


def lorenz(u, p, t):
    return [p[0] * (u[1] - u[0]), u[0] * (p[1] - u[2]) - u[1], u[0] * u[1] - p[2] * u[2]]


def test_gpusimpleatsit5_consistency():
    prob = sd.ODEProblem(sd.rhs.lorenz, 10 * np.ones(3), (0.0, 100.0), [10.0, 28.0, 8 / 3])
    sol2 = sd.solve(prob, sd.GPUSimpleATsit5(), dt=1e-2, abstol=1e-6, reltol=1e-3,
                    max_saves=20000)
    sol3 = sd.solve(prob, sd.GPUSimpleATsit5(), dt=1e-2, abstol=1e-6, reltol=1e-3,
                    save_everystep=False)
    assert sol2.t[-1] == sol3.t[-1] == 100.0
    assert bits_equal(sol2.u[-1], sol3.u[-1])
    # fixed step: every step / endpoint / saveat [5, 100] agree (gpusimpleatsit5_tests.jl:52-67)
    s2 = sd.solve(prob, sd.GPUSimpleTsit5(), dt=0.1, max_saves=2000)
    s3 = sd.solve(prob, sd.GPUSimpleTsit5(), dt=0.1, save_everystep=False)
    s4 = sd.solve(prob, sd.GPUSimpleTsit5(), dt=0.1, saveat=[5.0, 100.0])
    assert len(s2.t) == 1001 and bits_equal(s2.u[-1], s3.u[-1])
    assert np.allclose(s4.u[-1], s3.u[-1], rtol=1e-12, atol=0)
    i5 = int(np.argmin(np.abs(s2.t - 5.0)))
    assert np.allclose(s4.u[0], s2.u[i5], rtol=1e-10)


def test_traced_rhs_matches_builtin_bitwise():
    # The hipRTC path: the traced Lorenz function runs the same operations as the
    # built-in kernel, so results are identical to the bit.
    from simplediffeq_hip import inputs
    u0, p = inputs.lorenz_wide(2048, seed=1)
    for alg in (sd.GPUSimpleATsit5(), sd.GPUSimpleAVern7(), sd.GPUSimpleTsit5()):
        kw = dict(trajectories=2048, save_everystep=False, dt=0.1 if alg.adaptive else 0.01)
        a = sd.solve(sd.EnsembleProblem(sd.ODEProblem(lorenz, u0[:, 0], (0.0, 5.0), p[:, 0]),
                                        u0s=u0, ps=p), alg, sd.EnsembleHIP(), **kw)
        b = sd.solve(sd.EnsembleProblem(sd.ODEProblem(sd.rhs.lorenz, u0[:, 0], (0.0, 5.0),
                                                      p[:, 0]), u0s=u0, ps=p),
                     alg, sd.EnsembleHIP(), **kw)
        assert bits_equal(a.raw.u_end, b.raw.u_end) and bits_equal(a.raw.naccept, b.raw.naccept)


def oracle_ends(u0s, ps, tspans=None, tspan=(0.0, 1.0), alg=1, dt=0.1, abstol=1e-6,
                reltol=1e-3):
    """The CPU oracle's endpoints (n_state, n) of a Lorenz ensemble, the checker of
    the ensemble surface's per-trajectory results."""
    from oracle import pyoracle
    from simplediffeq_hip import _abi, runtime
    u0s = np.asarray(u0s, dtype=np.float64)
    n = u0s.shape[1]
    d = runtime.make_desc(alg, rhs=_abi.RHS_LORENZ, n_traj=n, tspan=tspan, dt=dt,
                          abstol=abstol, reltol=reltol)
    b = runtime.EnsembleBuffers(d, 3, u0s, np.asarray(ps, dtype=np.float64), tspan=tspans)
    assert pyoracle.solve(d, b, nthreads=8) == 0
    return b.u_end, b.t_end


def test_ensemble_problem_prob_func_output_func_reduction():
    base = sd.ODEProblem(sd.rhs.lorenz, [1.0, 0.0, 0.0], (0.0, 1.0), [10.0, 0.0, 8 / 3])
    rhos = np.linspace(0.0, 21.0, 64)
    eprob = sd.EnsembleProblem(
        base, prob_func=lambda prob, i, repeat: sd.remake(prob, p=[10.0, rhos[i - 1], 8 / 3]),
        output_func=lambda sol, i: (sol.u[-1][2], False),
        reduction=lambda u, data, I: ((u or 0.0) + sum(data), True), u_init=0.0)
    esol = sd.solve(eprob, sd.GPUSimpleATsit5(), sd.EnsembleHIP(), trajectories=64, dt=0.1,
                    abstol=1e-6, reltol=1e-3, save_everystep=False)
    single = [sd.solve(sd.remake(base, p=[10.0, r, 8 / 3]), sd.GPUSimpleATsit5(), dt=0.1,
                       abstol=1e-6, reltol=1e-3, save_everystep=False).u[-1][2] for r in rhos]
    assert esol.converged and esol.u == sum(single)
    # each trajectory the reduction saw is the oracle's, summed in the same order
    ps = np.stack([np.full(64, 10.0), rhos, np.full(64, 8 / 3)])
    ue, _ = oracle_ends(np.tile([[1.0], [0.0], [0.0]], 64), ps)
    assert esol.u == sum(float(z) for z in ue[2])
    # per-trajectory tspans from prob_func
    eprob = sd.EnsembleProblem(base, prob_func=lambda prob, i, r: sd.remake(prob, tspan=(0.0, 0.5 * i)))
    esol = sd.solve(eprob, sd.GPUSimpleATsit5(), sd.EnsembleHIP(), trajectories=4,
                    save_everystep=False)
    assert [s.t[-1] for s in esol.u] == [0.5, 1.0, 1.5, 2.0]
    spans = np.array([[0.0] * 4, [0.5, 1.0, 1.5, 2.0]])
    ue, te = oracle_ends(np.tile([[1.0], [0.0], [0.0]], 4), np.tile([[10.0], [0.0], [8 / 3]], 4),
                         tspans=spans, dt=float(F(0.1)), abstol=float(F(1e-6)),
                         reltol=float(F(1e-3)))
    for i, s in enumerate(esol.u):
        assert bits_equal(s.u[-1], ue[:, i]) and float(s.t[-1]) == float(F(te[i]))


def test_ensemble_batches_and_reruns_match_one_solve():
    base = sd.ODEProblem(sd.rhs.lorenz, [1.0, 0.0, 0.0], (0.0, 1.0), [10.0, 0.0, 8 / 3])
    rhos = np.linspace(0.0, 21.0, 100)

    def prob_func(prob, i, repeat):
        # a rerun solves the same trajectory with rho halved
        return sd.remake(prob, p=[10.0, rhos[i - 1] / 2 ** (repeat - 1), 8 / 3])

    kw = dict(trajectories=100, dt=0.1, abstol=1e-6, reltol=1e-3, save_everystep=False)
    one = sd.solve(sd.EnsembleProblem(base, prob_func=prob_func), sd.GPUSimpleATsit5(),
                   sd.EnsembleHIP(), **kw)
    many = sd.solve(sd.EnsembleProblem(base, prob_func=prob_func), sd.GPUSimpleATsit5(),
                    sd.EnsembleHIP(), batch_size=7, **kw)
    assert len(many.raw) == 15
    assert bits_equal(np.array([s.u[-1] for s in one.u]), np.array([s.u[-1] for s in many.u]))
    ps = np.stack([np.full(100, 10.0), rhos, np.full(100, 8 / 3)])
    ue, _ = oracle_ends(np.tile([[1.0], [0.0], [0.0]], 100), ps)
    assert bits_equal(np.array([s.u[-1] for s in many.u]), ue.T)
    # output_func asks for reruns while z(1) > 5: those end with the rho that
    # brought z below 5, exactly as a direct solve with that rho
    ep = sd.EnsembleProblem(base, prob_func=prob_func,
                            output_func=lambda sol, i: (sol, bool(sol.u[-1][2] > 5.0)))
    res = sd.solve(ep, sd.GPUSimpleATsit5(), sd.EnsembleHIP(), batch_size=32, **kw)
    reran = 0
    final_rho = []
    for i, s in enumerate(res.u):
        assert s.u[-1][2] <= 5.0
        rho = rhos[i]
        while True:
            d = sd.solve(sd.remake(base, p=[10.0, rho, 8 / 3]), sd.GPUSimpleATsit5(), dt=0.1,
                         abstol=1e-6, reltol=1e-3, save_everystep=False)
            if d.u[-1][2] <= 5.0:
                break
            rho /= 2
            reran += 1
        assert bits_equal(s.u[-1], d.u[-1])
        final_rho.append(rho)
    assert reran > 0
    # and every final solution is the oracle's for the rho its reruns ended with
    n = len(final_rho)
    ue, _ = oracle_ends(np.tile([[1.0], [0.0], [0.0]], n),
                        np.stack([np.full(n, 10.0), final_rho, np.full(n, 8 / 3)]))
    assert bits_equal(np.array([s.u[-1] for s in res.u]), ue.T)


def test_dtmin_raises_like_the_reference():
    prob = sd.ODEProblem(sd.rhs.expdecay3, np.ones(3), (0.0, 1.0), [-1.0])
    with pytest.raises(sd.DtLessThanMin, match="dt<dtmin"):
        sd.solve(prob, sd.GPUSimpleATsit5(), dt=1e-15, save_everystep=False)


def test_scalar_problem_keeps_scalar_states():
    """interface_tests.jl:128-132: a Float64-scalar problem keeps scalar states."""
    prob = sd.ODEProblem(lambda u, p, t: p * u, 1.0, (0.0, 1.0), -1.0)
    sol = sd.solve(prob, sd.GPUSimpleATsit5(), dt=0.01, abstol=1e-10, reltol=1e-10)
    assert all(np.ndim(x) == 0 for x in sol.u)
    assert abs(float(sol.u[-1]) - np.exp(-1.0)) < 1e-8
    sol = sd.solve(prob, sd.GPUSimpleTsit5(), dt=0.01, save_everystep=False)
    assert len(sol.u) == 2 and np.ndim(sol.u[-1]) == 0
    assert abs(float(sol.u[-1]) - np.exp(-1.0)) < 1e-9


def test_solution_times_take_dt_element_type():
    # ts = Vector{eltype(dt)} and push!(ts, t) (gpuatsit5.jl:72,226,303,325;
    # gpuvern7.jl:72,321; gpuvern9.jl:72,432): with the default dt = 0.1f0 the six
    # Tsit5/Verner solvers return Float32 times, the binary64 t rounded; a binary64
    # dt keeps binary64 times; saveat returns the caller's times; RK4/Euler return
    # their Float64 range (gpurk4.jl:66).
    prob = sd.ODEProblem(sd.rhs.lorenz, np.ones(3), (0.0, 0.7), [10.0, 28.0, 8 / 3])
    for alg in (sd.GPUSimpleATsit5(), sd.GPUSimpleAVern7(), sd.GPUSimpleAVern9()):
        s32 = sd.solve(prob, alg, save_everystep=False)
        s64 = sd.solve(prob, alg, dt=float(F(0.1)), save_everystep=False)
        assert s32.t.dtype == np.float32 and s64.t.dtype == np.float64
        assert s64.t[-1] == 0.7 and s32.t[-1] == F(0.7) and float(s32.t[-1]) != 0.7
        assert bits_equal(s32.u[-1], s64.u[-1])  # the same solve, only ts differs
        e32 = sd.solve(prob, alg)  # save_everystep = true
        e64 = sd.solve(prob, alg, dt=float(F(0.1)))
        assert e32.t.dtype == np.float32 and np.array_equal(e32.t, e64.t.astype(np.float32))
        sf = sd.solve(prob, alg, dt=F(0.1), save_everystep=False)  # a Float32 dt given
        assert sf.t.dtype == np.float32
    sa = sd.solve(prob, sd.GPUSimpleATsit5(), saveat=[0.0, 0.35, 0.7])
    assert sa.t.dtype == np.float64 and list(sa.t) == [0.0, 0.35, 0.7]
    for alg in (sd.GPUSimpleTsit5(), sd.GPUSimpleVern9()):
        assert sd.solve(prob, alg, dt=F(0.1), save_everystep=False).t.dtype == np.float32
    r = sd.solve(prob, sd.GPUSimpleRK4(), dt=F(0.1))
    assert r.t.dtype == np.float64


def test_binary32_avern9_below_1e9_runs_in_binary64():
    # the precision policy: binary32 states with AVern9 at a tolerance below 1e-9 are
    # integrated in binary64 and returned in binary32; at 1e-9 and above, binary32
    prob = sd.ODEProblem(sd.rhs.lorenz, np.ones(3, dtype=F), (0.0, 2.0),
                         np.array([10.0, 28.0, 8 / 3], dtype=F))
    tight = sd.solve(prob, sd.GPUSimpleAVern9(), abstol=1e-11, reltol=1e-11, save_everystep=False)
    p64 = sd.ODEProblem(sd.rhs.lorenz, np.ones(3), (0.0, 2.0),
                        np.array([10.0, 28.0, 8 / 3], dtype=F).astype(np.float64))
    ref = sd.solve(p64, sd.GPUSimpleAVern9(), abstol=1e-11, reltol=1e-11, save_everystep=False)
    assert tight.u[-1].dtype == np.float32
    assert np.array_equal(tight.u[-1], ref.u[-1].astype(np.float32))
    assert tight.stats == ref.stats
    loose = sd.solve(prob, sd.GPUSimpleAVern9(), abstol=1e-6, reltol=1e-6, save_everystep=False)
    assert loose.u[-1].dtype == np.float32
