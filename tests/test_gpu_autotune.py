"""The autotuner's kernel dimensions: the register cap (kernels compiled for 4, 6
and 8 waves per SIMD: amdgpu_waves_per_eu, at most 512 / waves VGPRs per lane) and
trajectories per lane (a variant whose lanes each integrate two trajectories,
their stages interleaved).  Every variant runs the same per-trajectory arithmetic
as the default kernel, so it gives the same bits, and the binary64 ones equal the
oracle; the tuner times them after the launch shape and later solves use its
choice (reported in stats.waves and stats.per_lane).  So does the persistent
schedules' refill batch (stats.refill)."""
import numpy as np
import pytest

from helpers import assert_bit_identical
from oracle import pyoracle
from simplediffeq_hip import _abi, inputs, runtime

pytestmark = pytest.mark.gpu

F32_01 = float(np.float32(0.1))
CASES = [(_abi.ALG_ATSIT5, _abi.F64, _abi.PREC_PLAIN), (_abi.ALG_ATSIT5, _abi.F32, _abi.PREC_PLAIN),
         (_abi.ALG_ATSIT5, _abi.F32, _abi.PREC_KAHAN), (_abi.ALG_AVERN7, _abi.F64, _abi.PREC_PLAIN),
         (_abi.ALG_AVERN7, _abi.F32, _abi.PREC_PLAIN)]
IDS = ["atsit5-f64", "atsit5-f32", "atsit5-f32kahan", "avern7-f64", "avern7-f32"]


def _run(ctx, alg, dtype, prec, waves, sched, n=30_001, seed=3, per_lane=0, refill=0):
    u0, p, ts = inputs.lorenz_uneven(n, seed=seed)
    ts[1] = np.minimum(ts[1], 5.0)
    d = runtime.make_desc(alg, dtype=dtype, rhs=_abi.RHS_LORENZ, n_traj=n, dt=F32_01,
                          abstol=1e-7, reltol=1e-6, precision=prec, schedule=sched, waves=waves,
                          per_lane=per_lane, refill=refill)
    g = runtime.EnsembleBuffers(d, 3, u0, p, tspan=ts, backend="torch", device="cuda:0")
    ctx.solve(d, g, sync=True)
    return d, g, ctx.stats(), (u0, p, ts)


@pytest.fixture(scope="module")
def fresh_ctx():
    c = runtime.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("alg,dtype,prec", CASES, ids=IDS)
@pytest.mark.parametrize("sched", [_abi.SCHED_ONE_LAUNCH, _abi.SCHED_PERSISTENT, _abi.SCHED_STAGED],
                         ids=["one_launch", "persistent", "staged"])
def test_register_cap_variants_give_the_same_bits(fresh_ctx, alg, dtype, prec, sched):
    ctx = fresh_ctx  # no tuned shape: waves = 0 is the default kernel
    _, ref, st0, inp = _run(ctx, alg, dtype, prec, 0, sched)
    assert st0["waves"] == 0
    for w in (4, 6, 8):
        _, g, st, _ = _run(ctx, alg, dtype, prec, w, sched)
        assert st["waves"] == w
        assert_bit_identical(g, ref)
    if dtype == _abi.F64 and sched == _abi.SCHED_PERSISTENT:
        d, _, _, (u0, p, ts) = _run(ctx, alg, dtype, prec, 6, sched, n=4096, seed=4)
        c = runtime.EnsembleBuffers(d, 3, u0, p, tspan=ts)
        assert pyoracle.solve(d, c, nthreads=8) == 0
        _, g, _, _ = _run(ctx, alg, dtype, prec, 6, sched, n=4096, seed=4)
        assert_bit_identical(g, c)


@pytest.mark.parametrize("alg,dtype,prec", CASES, ids=IDS)
@pytest.mark.parametrize("sched", [_abi.SCHED_PERSISTENT, _abi.SCHED_STAGED],
                         ids=["persistent", "staged"])
def test_refill_batches_give_the_same_bits(fresh_ctx, alg, dtype, prec, sched):
    """The persistent schedules' refill batch (a wave claims new trajectories once at
    least `refill` of its slots are free) changes which lane runs a trajectory and
    when, never its arithmetic: every batch gives the default's bits, and stats report
    the batch used (at most the wave's 64 slots; 4 by default)."""
    ctx = fresh_ctx
    _, ref, st0, _ = _run(ctx, alg, dtype, prec, 0, sched)
    assert st0["refill"] == 4
    for rf, used in ((1, 1), (2, 2), (8, 8), (64, 64), (1000, 64)):
        _, g, st, _ = _run(ctx, alg, dtype, prec, 0, sched, refill=rf)
        assert st["refill"] == used
        assert_bit_identical(g, ref)
    _, g, st, _ = _run(ctx, alg, dtype, prec, 0, _abi.SCHED_ONE_LAUNCH, refill=8)
    assert st["refill"] == 0
    assert_bit_identical(g, ref)


MULTI = [(_abi.ALG_ATSIT5, _abi.F64, _abi.PREC_PLAIN), (_abi.ALG_ATSIT5, _abi.F32, _abi.PREC_PLAIN),
         (_abi.ALG_ATSIT5, _abi.F32, _abi.PREC_KAHAN), (_abi.ALG_AVERN7, _abi.F64, _abi.PREC_PLAIN)]


@pytest.mark.parametrize("alg,dtype,prec", MULTI, ids=IDS[:3] + ["avern7-f64"])
@pytest.mark.parametrize("sched", [_abi.SCHED_ONE_LAUNCH, _abi.SCHED_PERSISTENT, _abi.SCHED_STAGED],
                         ids=["one_launch", "persistent", "staged_shape"])
@pytest.mark.parametrize("n", [30_001, 2, 1], ids=["n30001", "n2", "n1"])
def test_two_trajectories_per_lane_give_the_same_bits(fresh_ctx, alg, dtype, prec, sched, n):
    ctx = fresh_ctx
    _, ref, _, _ = _run(ctx, alg, dtype, prec, 0, sched, n=n)
    d, g, st, (u0, p, ts) = _run(ctx, alg, dtype, prec, 0, sched, n=n, per_lane=2, refill=8)
    assert st["per_lane"] == 2 and st["schedule"] != _abi.SCHED_STAGED
    # the two-per-lane kernels claim for each free slot whatever the batch asked for
    assert st["refill"] == (0 if st["schedule"] == _abi.SCHED_ONE_LAUNCH else 1)
    assert_bit_identical(g, ref)
    if dtype == _abi.F64 and n > 2:
        c = runtime.EnsembleBuffers(d, 3, u0, p, tspan=ts)
        assert pyoracle.solve(d, c, nthreads=8) == 0
        assert_bit_identical(g, c)


def test_two_per_lane_traced_rhs(fresh_ctx):
    # a registered (hipRTC) right-hand side gets the variant compiled on demand
    from simplediffeq_hip import tracer
    ctx = fresh_ctx
    src, fn = tracer.trace(lambda u, p, t: [p[0] * (u[1] - u[0]), u[0] * (p[1] - u[2]) - u[1],
                                            u[0] * u[1] - p[2] * u[2]], 3, 3, "lorenz_pair")
    h = ctx.compile_rhs(src, fn, 3, 3)
    n = 5003
    u0, p = inputs.lorenz_wide(n, seed=1)
    out = []
    for pl in (1, 2):
        d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=h, n_traj=n, tspan=(0.0, 3.0), dt=F32_01,
                              per_lane=pl, schedule=_abi.SCHED_PERSISTENT)
        g = runtime.EnsembleBuffers(d, 3, u0, p, backend="torch", device="cuda:0")
        ctx.solve(d, g, sync=True)
        assert ctx.stats()["per_lane"] == pl
        out.append(g)
    assert_bit_identical(out[1], out[0])


def test_traced_rhs_persistent_paths_match_builtin(fresh_ctx):
    """A registered (hipRTC) right-hand side gets the persistent queue's refill
    batches, order cache and staged resume compiled from the same headers: binary32
    compensated ATsit5 with a caller's order, every batch, and the staged schedule all
    give the built-in kernel's bits."""
    from simplediffeq_hip import tracer
    ctx = fresh_ctx
    src, fn = tracer.trace(lambda u, p, t: [p[0] * (u[1] - u[0]), u[0] * (p[1] - u[2]) - u[1],
                                            u[0] * u[1] - p[2] * u[2]], 3, 3, "lorenz_queue")
    h = ctx.compile_rhs(src, fn, 3, 3)
    n = 20_011
    u0, p, ts = inputs.lorenz_uneven(n, seed=5)
    ts[1] = np.minimum(ts[1], 4.0)
    perm = np.random.default_rng(5).permutation(n).astype(np.int64)
    out = {}
    for rhs in (_abi.RHS_LORENZ, h):
        for name, sched, order, refill in (("persistent_r1", _abi.SCHED_PERSISTENT, perm, 1),
                                           ("persistent_r8", _abi.SCHED_PERSISTENT, perm, 8),
                                           ("staged", _abi.SCHED_STAGED, None, 0)):
            d = runtime.make_desc(_abi.ALG_ATSIT5, dtype=_abi.F32, rhs=rhs, n_traj=n, dt=F32_01,
                                  abstol=1e-6, reltol=1e-4, precision=_abi.PREC_KAHAN,
                                  schedule=sched, refill=refill, chunk=16)
            g = runtime.EnsembleBuffers(d, 3, u0, p, tspan=ts, order=order, backend="torch",
                                        device="cuda:0")
            ctx.solve(d, g, sync=True)
            assert ctx.stats()["schedule"] == sched
            out[(rhs, name)] = g
    for name in ("persistent_r1", "persistent_r8", "staged"):
        assert_bit_identical(out[(h, name)], out[(_abi.RHS_LORENZ, name)])
        assert_bit_identical(out[(h, name)], out[(h, "persistent_r1")])


def test_register_cap_refusals(ctx):
    n = 64
    u0, p = inputs.lorenz_wide(n, seed=1)
    with pytest.raises(runtime.SolveError, match="waves must be"):
        d = runtime.make_desc(_abi.ALG_ATSIT5, n_traj=n, waves=5)
        ctx.solve(d, runtime.EnsembleBuffers(d, 3, u0, p, backend="torch", device="cuda:0"))
    # saveat and the fixed-step solvers have no variants
    d = runtime.make_desc(_abi.ALG_ATSIT5, n_traj=n, waves=6, saveat=[0.0, 0.5])
    with pytest.raises(runtime.SolveError, match="register-cap"):
        ctx.solve(d, runtime.EnsembleBuffers(d, 3, u0, p, backend="torch", device="cuda:0"))
    d = runtime.make_desc(_abi.ALG_TSIT5, n_traj=n, dt=0.01, waves=4)
    with pytest.raises(runtime.SolveError, match="register-cap"):
        ctx.solve(d, runtime.EnsembleBuffers(d, 3, u0, p, backend="torch", device="cuda:0"))
    # nor the group and matrix-core layouts of GLV-16
    u0g, r, A = inputs.glv16(n, seed=3)
    d = runtime.make_desc(_abi.ALG_AVERN9, rhs=_abi.RHS_GLV16, n_traj=n, tspan=(0.0, 1.0),
                          waves=4)
    with pytest.raises(runtime.SolveError, match="register-cap"):
        ctx.solve(d, runtime.EnsembleBuffers(d, 16, u0g, r, A, backend="torch", device="cuda:0"))
    # two per lane: adaptive ATsit5/AVern7 endpoint only, not with a register cap
    d = runtime.make_desc(_abi.ALG_AVERN9, n_traj=n, per_lane=2)
    with pytest.raises(runtime.SolveError, match="several-trajectories-per-lane"):
        ctx.solve(d, runtime.EnsembleBuffers(d, 3, u0, p, backend="torch", device="cuda:0"))
    d = runtime.make_desc(_abi.ALG_ATSIT5, n_traj=n, per_lane=2, waves=4)
    with pytest.raises(runtime.SolveError, match="do not combine"):
        ctx.solve(d, runtime.EnsembleBuffers(d, 3, u0, p, backend="torch", device="cuda:0"))
    with pytest.raises(runtime.SolveError, match="per_lane must be"):
        d = runtime.make_desc(_abi.ALG_ATSIT5, n_traj=n, per_lane=3)
        ctx.solve(d, runtime.EnsembleBuffers(d, 3, u0, p, backend="torch", device="cuda:0"))
    with pytest.raises(runtime.SolveError, match="refill must be"):
        d = runtime.make_desc(_abi.ALG_ATSIT5, n_traj=n, refill=-1)
        ctx.solve(d, runtime.EnsembleBuffers(d, 3, u0, p, backend="torch", device="cuda:0"))


def test_autotuner_reports_and_keeps_its_register_cap(ctx):
    n = 100_000
    u0, p = inputs.lorenz_wide(n, seed=1)
    d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=n, tspan=(0.0, 10.0),
                          dt=F32_01, abstol=1e-6, reltol=1e-3)
    g = runtime.EnsembleBuffers(d, 3, u0, p, backend="torch", device="cuda:0")
    best = ctx.autotune(d, g)
    assert best["waves"] in (0, 4, 6, 8) and best["per_lane"] in (1, 2)
    assert best["refill"] in ((0,) if best["schedule"] == _abi.SCHED_ONE_LAUNCH else (1, 4, 8))
    ctx.solve(d, g, sync=True)
    st = ctx.stats()
    assert st["waves"] == best["waves"] and st["schedule"] == best["schedule"]
    assert st["per_lane"] == best["per_lane"] and st["refill"] == best["refill"]
    ref = runtime.EnsembleBuffers(d, 3, u0, p, backend="torch", device="cuda:0")
    d1 = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=n, tspan=(0.0, 10.0),
                           dt=F32_01, abstol=1e-6, reltol=1e-3, schedule=_abi.SCHED_ONE_LAUNCH)
    ctx.solve(d1, ref, sync=True)
    assert_bit_identical(g, ref)
