"""Every trajectory is handed to a lane exactly once per launch: the kernels' debug
visit counters (sdeh_io.visits) end at 1 everywhere for the one-launch and
persistent schedules (every queue chunk size and refill batch, index,
longest-span-first and caller orders, lane layouts, the allocation-free solve, graph
replays) and, for the staged
schedule, at 1 for trajectories that ended in its first pass and 2 for those it
parked and resumed (more than 32 attempts)."""
import numpy as np
import pytest

from simplediffeq_hip import _abi, inputs, runtime

pytestmark = pytest.mark.gpu

N = 100_003  # not a multiple of any chunk or block
F32_01 = float(np.float32(0.1))


def _counted(ctx, d, g, call="solve"):
    import torch
    v = torch.zeros(int(d.n_traj), dtype=torch.int32, device="cuda:0")
    g.io.visits = v.data_ptr()
    torch.cuda.synchronize()
    if call == "solve":
        ctx.solve(d, g, sync=True)
    elif call == "static":
        ctx.solve_static(d, g, sync=True)
    torch.cuda.synchronize()
    return v.cpu().numpy(), ctx.stats()


def _lorenz(d, spans=False, order=None, seed=1):
    u0, p, ts = inputs.lorenz_uneven(N, seed=seed)
    ts[1] = np.minimum(ts[1], 4.0)
    return runtime.EnsembleBuffers(d, 3, u0, p, tspan=ts if spans else None, order=order,
                                   backend="torch", device="cuda:0")


@pytest.mark.parametrize("sched,chunk", [(_abi.SCHED_ONE_LAUNCH, 0), (_abi.SCHED_PERSISTENT, 1),
                                         (_abi.SCHED_PERSISTENT, 16), (_abi.SCHED_PERSISTENT, 64),
                                         (_abi.SCHED_PERSISTENT, 1000)],
                         ids=["one_launch", "chunk1", "chunk16", "chunk64", "chunk1000"])
@pytest.mark.parametrize("spans", [False, True], ids=["common_span", "span_order"])
def test_each_trajectory_once(ctx, sched, chunk, spans):
    for alg in (_abi.ALG_ATSIT5, _abi.ALG_AVERN7):
        d = runtime.make_desc(alg, rhs=_abi.RHS_LORENZ, n_traj=N, tspan=(0.0, 2.0), dt=F32_01,
                              abstol=1e-6, reltol=1e-4, schedule=sched, chunk=chunk)
        v, _ = _counted(ctx, d, _lorenz(d, spans))
        assert (v == 1).all(), (alg, np.unique(v, return_counts=True))


@pytest.mark.parametrize("refill", [1, 4, 64])
def test_refill_batches_once(ctx, refill):
    for sched in (_abi.SCHED_PERSISTENT, _abi.SCHED_STAGED):
        d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=N, dt=F32_01,
                              abstol=1e-6, reltol=1e-4, schedule=sched, refill=refill)
        g = _lorenz(d, spans=True)
        v, st = _counted(ctx, d, g)
        assert st["refill"] == refill
        if sched == _abi.SCHED_PERSISTENT:
            assert (v == 1).all()
        else:
            att = g.numpy("naccept").astype(np.int64) + g.numpy("nreject")
            assert np.array_equal(v, np.where(att > 32, 2, 1))


@pytest.mark.parametrize("sched", [_abi.SCHED_ONE_LAUNCH, _abi.SCHED_PERSISTENT],
                         ids=["one_launch", "persistent"])
def test_two_per_lane_once(ctx, sched):
    d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=N, dt=F32_01,
                          abstol=1e-6, reltol=1e-4, schedule=sched, per_lane=2)
    v, st = _counted(ctx, d, _lorenz(d, spans=True))
    assert st["per_lane"] == 2 and (v == 1).all()


def test_caller_order_and_fixed_step(ctx):
    perm = np.random.default_rng(0).permutation(N).astype(np.int64)
    d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=N, tspan=(0.0, 1.0),
                          schedule=_abi.SCHED_PERSISTENT)
    v, _ = _counted(ctx, d, _lorenz(d, order=perm))
    assert (v == 1).all()
    d = runtime.make_desc(_abi.ALG_TSIT5, rhs=_abi.RHS_LORENZ, n_traj=N, tspan=(0.0, 1.0), dt=0.01)
    v, _ = _counted(ctx, d, _lorenz(d))
    assert (v == 1).all()


def test_staged_parks_and_resumes_once(ctx):
    d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=N, dt=F32_01,
                          abstol=1e-6, reltol=1e-4, schedule=_abi.SCHED_STAGED)
    g = _lorenz(d, spans=True)
    v, st = _counted(ctx, d, g)
    assert st["schedule"] == _abi.SCHED_STAGED
    att = g.numpy("naccept").astype(np.int64) + g.numpy("nreject")
    want = np.where(att > 32, 2, 1)  # the first pass runs 32 attempts
    assert np.array_equal(v, want)
    assert (want == 2).any() and (want == 1).any()


def test_group_and_matrix_core_layouts(ctx):
    n = 20_011
    u0, r, A = inputs.glv16(n, seed=3)
    for lanes in (4, 16):
        d = runtime.make_desc(_abi.ALG_AVERN9, rhs=_abi.RHS_GLV16, n_traj=n, tspan=(0.0, 2.0),
                              dt=0.1, abstol=1e-8, reltol=1e-8, lanes=lanes,
                              schedule=_abi.SCHED_PERSISTENT)
        g = runtime.EnsembleBuffers(d, 16, u0, r, A, backend="torch", device="cuda:0")
        v, st = _counted(ctx, d, g)
        assert st["lanes"] == lanes and (v == 1).all()
    # the largest refill batch: every slot of the wave (16 lanes: 4 slots; the
    # matrix-core layout, 4 lanes: 16)
    for lanes, slots in ((16, 4), (4, 16)):
        d = runtime.make_desc(_abi.ALG_AVERN9, rhs=_abi.RHS_GLV16, n_traj=n, tspan=(0.0, 2.0),
                              dt=0.1, abstol=1e-8, reltol=1e-8, lanes=lanes,
                              schedule=_abi.SCHED_PERSISTENT, refill=100)
        g = runtime.EnsembleBuffers(d, 16, u0, r, A, backend="torch", device="cuda:0")
        v, st = _counted(ctx, d, g)
        assert st["lanes"] == lanes and st["refill"] == slots and (v == 1).all()


def test_static_solve_and_graph_replays(ctx):
    import torch
    d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=N, tspan=(0.0, 2.0),
                          dt=F32_01, schedule=_abi.SCHED_PERSISTENT)
    v, _ = _counted(ctx, d, _lorenz(d), call="static")
    assert (v == 1).all()
    g = _lorenz(d)
    vis = torch.zeros(N, dtype=torch.int32, device="cuda:0")
    g.io.visits = vis.data_ptr()
    torch.cuda.synchronize()
    h = ctx.graph_capture(d, g)
    try:
        for k in range(1, 4):
            ctx.graph_launch(h)
            ctx.synchronize()
            assert (vis.cpu().numpy() == k).all()  # each replay hands each one out once
    finally:
        ctx.graph_destroy(h)
