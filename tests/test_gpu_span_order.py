"""Longest-span-first hand-out order (the default when trajectories carry their
own tspans): a permutation of the lanes' work only, so every output is the same
bits as in index order, for every schedule, save mode and memory path (eager,
host pipeline, graph replay)."""
import numpy as np
import pytest

from helpers import bits_equal
from simplediffeq_hip import _abi, inputs, runtime

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
FIELDS = ("u_end", "t_end", "retcode", "naccept", "nreject")


def _nan_bufs(d, u0, p, ts, **kw):
    import torch
    b = runtime.EnsembleBuffers(d, 3, u0, p, tspan=ts, backend="torch", device=DEV, **kw)
    for f in ("u_end", "t_end"):
        getattr(b, f).fill_(float("nan"))
    for f in ("retcode", "naccept", "nreject"):
        getattr(b, f).fill_(-7)
    torch.cuda.synchronize()
    return b


@pytest.mark.parametrize("save", ["end", "saveat", "steps", "packed"])
@pytest.mark.parametrize("sched", [_abi.SCHED_ONE_LAUNCH, _abi.SCHED_PERSISTENT])
def test_span_order_leaves_results_unchanged(ctx, sched, save):
    n = 40000
    u0, p, ts = inputs.lorenz_uneven(n, seed=21)
    ts[1] = np.minimum(ts[1], 4.0)
    kw = dict(rhs=_abi.RHS_LORENZ, n_traj=n, dt=0.1, abstol=1e-6, reltol=1e-4, schedule=sched)
    extra = {}
    if save == "saveat":
        kw.update(saveat=np.linspace(0.0, 4.0, 9), save_layout=_abi.SAVE_TRAJ_MAJOR)
    elif save == "steps":
        kw.update(save_everystep=True, max_saves=400, save_layout=_abi.SAVE_TRAJ_MAJOR)
    elif save == "packed":
        kw.update(save_everystep=True, save_layout=_abi.SAVE_PACKED)
        extra = dict(steps_capacity=n * 400)
    out = {}
    for so in (1, 0):
        d = runtime.make_desc(_abi.ALG_ATSIT5, span_order=so, **kw)
        b = _nan_bufs(d, u0, p, ts, **extra)
        ctx.solve(d, b)
        out[so] = b
    fields = FIELDS + {"end": (), "saveat": ("u_saveat",), "steps": ("n_steps",),
                       "packed": ("n_steps", "steps_offset")}[save]
    for f in fields:
        assert bits_equal(out[0].numpy(f), out[1].numpy(f)), f
    assert (out[0].numpy("retcode") == 0).all()
    if save == "packed":
        used = int(out[0].numpy("steps_offset")[n])
        assert bits_equal(out[0].numpy("u_steps")[:used], out[1].numpy("u_steps")[:used])


def test_span_order_host_pipeline_and_graph(ctx):
    import torch
    n = (1 << 17) + 7
    u0, p, ts = inputs.lorenz_uneven(n, seed=22)
    ts[1] = np.minimum(ts[1], 3.0)
    base = dict(rhs=_abi.RHS_LORENZ, n_traj=n, dt=0.1, abstol=1e-6, reltol=1e-4)
    d1 = runtime.make_desc(_abi.ALG_AVERN7, span_order=1, **base)
    ref = _nan_bufs(d1, u0, p, ts)
    ctx.solve(d1, ref)
    d0 = runtime.make_desc(_abi.ALG_AVERN7, **base)
    host = runtime.EnsembleBuffers(d0, 3, u0, p, tspan=ts)  # pipelined, ordered per chunk
    ctx.solve(d0, host)
    for f in FIELDS:
        assert bits_equal(host.numpy(f), ref.numpy(f)), f
    g = _nan_bufs(d0, u0, p, ts)
    h = ctx.graph_capture(d0, g)
    for _ in range(2):
        ctx.graph_launch(h)
        ctx.synchronize()
        for f in FIELDS:
            assert bits_equal(g.numpy(f), ref.numpy(f)), f
        for f in ("u_end", "t_end"):
            getattr(g, f).fill_(float("nan"))
        torch.cuda.synchronize()
    ctx.graph_destroy(h)


def test_span_order_field_validated(ctx):
    u0, p, ts = inputs.lorenz_uneven(8, seed=1)
    d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=8, span_order=2)
    with pytest.raises(runtime.SolveError, match="span_order"):
        ctx.solve(d, runtime.EnsembleBuffers(d, 3, u0, p, tspan=ts))
