"""Staged schedule (SDEH_SCHED_STAGED): a first one-launch pass of 32 attempts per
trajectory parks the trajectories still running (state, step size, controller
state, counts), a device counting sort orders them by expected remaining attempts,
and a persistent launch resumes them.  Parking carries the whole per-trajectory
state, so every output is the same bits as the one-launch and persistent schedules
and as the oracle: trajectories that end in the first pass, in the second, at
maxiters on either side of the cap, and with a NaN."""
import numpy as np
import pytest

from helpers import assert_bit_identical, run_pair
from simplediffeq_hip import _abi, inputs, runtime

pytestmark = pytest.mark.gpu

FIELDS = ("u_end", "t_end", "retcode", "naccept", "nreject")
STAGE = 32  # attempts of the first pass


def _uneven(n, seed):
    u0, p, ts = inputs.lorenz_uneven(n, seed=seed)
    ts[1] = ts[0] + np.minimum(ts[1] - ts[0], 6.0)
    ts[1, ::7] = ts[0, ::7] + 0.05  # a few steps: these end in the first pass
    return u0, p, ts


def _solve(ctx, d, u0, p, ts, **kw):
    b = runtime.EnsembleBuffers(d, u0.shape[0], u0, p, tspan=ts, **kw)
    st = ctx.solve(d, b)
    return st, b, ctx.stats()


@pytest.mark.parametrize("alg,dtype,prec", [
    (_abi.ALG_ATSIT5, _abi.F64, _abi.PREC_PLAIN),
    (_abi.ALG_AVERN7, _abi.F64, _abi.PREC_PLAIN),
    (_abi.ALG_AVERN9, _abi.F64, _abi.PREC_PLAIN),
    (_abi.ALG_ATSIT5, _abi.F32, _abi.PREC_PLAIN),
    (_abi.ALG_ATSIT5, _abi.F32, _abi.PREC_KAHAN),
], ids=["atsit5-f64", "avern7-f64", "avern9-f64", "atsit5-f32", "atsit5-f32-kahan"])
def test_staged_same_bits_as_other_schedules(ctx, alg, dtype, prec):
    n = 6000
    u0, p, ts = _uneven(n, seed=31)
    if dtype == _abi.F32:
        u0, p = u0.astype(np.float32), p.astype(np.float32)
    out = {}
    for sched in (_abi.SCHED_ONE_LAUNCH, _abi.SCHED_PERSISTENT, _abi.SCHED_STAGED):
        d = runtime.make_desc(alg, dtype=dtype, rhs=_abi.RHS_LORENZ, n_traj=n, dt=0.1,
                              abstol=1e-6, reltol=1e-4, schedule=sched, precision=prec,
                              chunk=16, lanes=1)
        st, b, s = _solve(ctx, d, u0, p, ts)
        assert st == 0
        assert s["schedule"] == sched
        out[sched] = b
    staged = out[_abi.SCHED_STAGED]
    att = staged.numpy("naccept").astype(np.int64) + staged.numpy("nreject")
    assert (att < STAGE).any() and (att > 2 * STAGE).any()  # both passes end trajectories
    assert (staged.numpy("retcode") >= 0).all()  # no trajectory is left parked
    for sched in (_abi.SCHED_ONE_LAUNCH, _abi.SCHED_PERSISTENT):
        assert_bit_identical(staged, out[sched], FIELDS)


@pytest.mark.parametrize("maxiters", [20, STAGE, STAGE + 1, 45, 0],
                         ids=["below-cap", "at-cap", "cap+1", "above-cap", "default"])
def test_staged_against_oracle(ctx, maxiters):
    """binary64 ATsit5, bit for bit against the oracle, with maxiters below, at and
    above the first pass's cap (MAXITERS in the first pass, at the boundary, and in
    the second)."""
    n = 3000
    u0, p, ts = _uneven(n, seed=32)
    d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=n, dt=0.1,
                          abstol=1e-6, reltol=1e-4, schedule=_abi.SCHED_STAGED, chunk=16,
                          lanes=1, maxiters=maxiters)
    st_g, g, st_c, c = run_pair(ctx, d, u0, p, tspan=ts)
    assert st_g == st_c
    assert ctx.stats()["schedule"] == _abi.SCHED_STAGED
    assert_bit_identical(g, c, FIELDS)
    rc = g.numpy("retcode")
    if maxiters:
        assert (rc == _abi.RET_MAXITERS).any() and (rc == _abi.RET_DEFAULT).any()


def test_staged_dt0_nan_and_device_memory(ctx):
    """Per-trajectory dt0, trajectories that stop with a NaN (states whose products
    overflow), device (torch) buffers."""
    n = 4000
    u0, p, ts = _uneven(n, seed=33)
    u0[:, ::11] = 1.0e155
    dt0 = np.where(np.arange(n) % 3 == 0, 0.01, 0.2)
    d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=n, dt=0.1,
                          abstol=1e-8, reltol=1e-6, schedule=_abi.SCHED_STAGED, lanes=1)
    st_g, g, st_c, c = run_pair(ctx, d, u0, p, tspan=ts, dt0=dt0, backend="torch")
    assert st_g == st_c
    assert ctx.stats()["schedule"] == _abi.SCHED_STAGED
    assert_bit_identical(g, c, FIELDS)
    assert (g.numpy("retcode") == _abi.RET_NAN).any()


def test_staged_falls_back_to_persistent(ctx):
    """Outputs other than the endpoint, or a caller's order, run the staged shape as
    the persistent schedule, with the same results."""
    n = 2000
    u0, p, ts = _uneven(n, seed=34)
    base = dict(rhs=_abi.RHS_LORENZ, n_traj=n, dt=0.1, abstol=1e-6, reltol=1e-4, lanes=1)
    sv = np.linspace(0.0, 0.05, 3)
    for extra, kw in [(dict(saveat=sv), {}),
                      ({}, dict(order=np.arange(n)[::-1].copy()))]:
        res = {}
        for sched in (_abi.SCHED_PERSISTENT, _abi.SCHED_STAGED):
            d = runtime.make_desc(_abi.ALG_ATSIT5, schedule=sched, **base, **extra)
            st, b, s = _solve(ctx, d, u0, p, ts, **kw)
            assert st == 0 and s["schedule"] == _abi.SCHED_PERSISTENT
            res[sched] = b
        fields = FIELDS + (("u_saveat",) if "saveat" in extra else ())
        assert_bit_identical(res[_abi.SCHED_STAGED], res[_abi.SCHED_PERSISTENT], fields)


def test_staged_compiled_rhs(ctx):
    """A right-hand side compiled at run time takes the staged path too."""
    src = """
template <class T>
__device__ void damped(T* du, const T* u, const T* p, const T* __restrict__ ps, double t) {
  du[0] = u[1];
  du[1] = -p[0] * u[0] - p[1] * u[1];
}
"""
    h = ctx.compile_rhs(src, "damped", 2, 2)
    n = 3000
    rng = np.random.default_rng(35)
    u0 = rng.uniform(-1, 1, (2, n))
    p = np.stack([rng.uniform(1, 400, n), rng.uniform(0.0, 0.5, n)])
    ts = np.stack([np.zeros(n), rng.uniform(0.1, 8.0, n)])
    res = {}
    for sched in (_abi.SCHED_PERSISTENT, _abi.SCHED_STAGED):
        d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=h, n_traj=n, dt=0.1, abstol=1e-7,
                              reltol=1e-5, schedule=sched)
        st, b, s = _solve(ctx, d, u0, p, ts)
        assert st == 0 and s["schedule"] == sched
        res[sched] = b
    att = res[_abi.SCHED_STAGED].numpy("naccept").astype(np.int64)
    assert (att > 2 * STAGE).any()
    assert_bit_identical(res[_abi.SCHED_STAGED], res[_abi.SCHED_PERSISTENT], FIELDS)


def test_staged_wide_lorenz(ctx):
    """The wide Lorenz ensemble (rho in U(0,21), tspan (0, 10), 1f-6 / 1f-3) at 2^16
    trajectories: the staged solve matches the persistent one bit for bit (one
    tspan for all, no per-trajectory spans)."""
    n = 1 << 16
    u0, p = inputs.lorenz_wide(n, seed=1)
    res = {}
    for sched in (_abi.SCHED_PERSISTENT, _abi.SCHED_STAGED):
        d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=n,
                              tspan=(0.0, 10.0), dt=float(np.float32(0.1)), schedule=sched)
        b = runtime.EnsembleBuffers(d, 3, u0, p, backend="torch", device="cuda:0")
        assert ctx.solve(d, b) == 0
        assert ctx.stats()["schedule"] == sched
        res[sched] = b
    assert_bit_identical(res[_abi.SCHED_STAGED], res[_abi.SCHED_PERSISTENT], FIELDS)


def test_staged_shape_in_graph_capture(ctx):
    """Graph capture runs a staged shape as the persistent schedule, still handing
    out trajectories longest span first: every replay gives the eager staged
    solve's bits."""
    import torch
    n = 5000
    u0, p, ts = _uneven(n, seed=36)
    d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=n, dt=0.1,
                          abstol=1e-6, reltol=1e-4, schedule=_abi.SCHED_STAGED, lanes=1)
    ref = runtime.EnsembleBuffers(d, 3, u0, p, tspan=ts, backend="torch", device="cuda:0")
    assert ctx.solve(d, ref) == 0
    assert ctx.stats()["schedule"] == _abi.SCHED_STAGED
    g = runtime.EnsembleBuffers(d, 3, u0, p, tspan=ts, backend="torch", device="cuda:0")
    torch.cuda.synchronize()
    h = ctx.graph_capture(d, g)
    for _ in range(2):
        ctx.graph_launch(h)
        ctx.synchronize()
        assert_bit_identical(g, ref, FIELDS)
        for f in ("u_end", "t_end"):
            getattr(g, f).fill_(float("nan"))
        torch.cuda.synchronize()
    ctx.graph_destroy(h)
