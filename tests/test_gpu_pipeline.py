"""Host-memory solves of 2^17 trajectories and more run as a pipeline of
chunks (inputs of one chunk travel while another integrates); device-memory
solves run as one launch.  The per-trajectory results must be the same bits,
for every output layout and input the pipeline moves."""
import numpy as np
import pytest

from helpers import bits_equal
from simplediffeq_hip import _abi, inputs, runtime

pytestmark = pytest.mark.gpu

N = (1 << 17) + 333  # not a multiple of the chunk count


def _pair(ctx, d, n_state, u0, p, ps=None, tspan=None, dt0=None):
    h = runtime.EnsembleBuffers(d, n_state, u0, p, ps, tspan, dt0)
    sh = ctx.solve(d, h, allow=(_abi.E_DTMIN, _abi.E_SAVE_OVERFLOW))
    st_h = ctx.stats()
    g = runtime.EnsembleBuffers(d, n_state, u0, p, ps, tspan, dt0, backend="torch",
                                device="cuda:0")
    sg = ctx.solve(d, g, allow=(_abi.E_DTMIN, _abi.E_SAVE_OVERFLOW))
    assert sh == sg
    assert st_h["attempts"] == ctx.stats()["attempts"]
    return h, g


def _same(h, g, fields):
    for f in fields:
        assert bits_equal(h.numpy(f), g.numpy(f)), f


BASE = ("u_end", "t_end", "retcode", "naccept", "nreject")


@pytest.mark.parametrize("sched", [_abi.SCHED_ONE_LAUNCH, _abi.SCHED_PERSISTENT])
def test_endpoint_with_spans_and_dt0(ctx, sched):
    u0, p, ts = inputs.lorenz_uneven(N, seed=8)
    ts[1] = np.minimum(ts[1], 3.0)
    dt0 = np.full(N, 0.05)
    d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=N, dt=0.1,
                          abstol=1e-6, reltol=1e-4, schedule=sched)
    h, g = _pair(ctx, d, 3, u0, p, tspan=ts, dt0=dt0)
    _same(h, g, BASE)
    assert (h.t_end == ts[1]).all()


@pytest.mark.parametrize("layout", [_abi.SAVE_POINT_MAJOR, _abi.SAVE_TRAJ_MAJOR],
                         ids=["point-major", "traj-major"])
def test_saveat_both_layouts(ctx, layout):
    u0, p = inputs.lorenz_wide(N, seed=9)
    d = runtime.make_desc(_abi.ALG_AVERN7, rhs=_abi.RHS_LORENZ, n_traj=N, tspan=(0.0, 1.0),
                          dt=0.1, abstol=1e-7, reltol=1e-7, saveat=np.linspace(0.0, 1.0, 9),
                          save_layout=layout)
    h, g = _pair(ctx, d, 3, u0, p)
    _same(h, g, BASE + ("u_saveat",))


@pytest.mark.parametrize("alg", [_abi.ALG_TSIT5, _abi.ALG_ATSIT5])
@pytest.mark.parametrize("layout", [_abi.SAVE_POINT_MAJOR, _abi.SAVE_TRAJ_MAJOR],
                         ids=["point-major", "traj-major"])
def test_everystep_fixed_rows(ctx, alg, layout):
    u0, p = inputs.lorenz_wide(N, seed=10)
    d = runtime.make_desc(alg, rhs=_abi.RHS_LORENZ, n_traj=N, tspan=(0.0, 0.5), dt=0.05,
                          abstol=1e-6, reltol=1e-4, save_everystep=True, max_saves=40,
                          save_layout=layout)
    h, g = _pair(ctx, d, 3, u0, p)
    _same(h, g, BASE + ("n_steps",))
    k = h.n_steps
    for i in range(0, N, 977):  # rows beyond n_steps are not written
        if layout == _abi.SAVE_TRAJ_MAJOR:
            a, b = h.u_steps[i, :k[i]], g.u_steps[i, :k[i]].cpu().numpy()
        else:
            a, b = h.u_steps[:k[i], :, i], g.u_steps[:k[i], :, i].cpu().numpy()
        assert bits_equal(a, b), i


def test_binary32_and_fixed_vern9(ctx):
    u0, p = inputs.lorenz_wide(N, seed=11)
    d = runtime.make_desc(_abi.ALG_ATSIT5, dtype=_abi.F32, rhs=_abi.RHS_LORENZ, n_traj=N,
                          tspan=(0.0, 2.0), dt=0.1, abstol=1e-5, reltol=1e-4)
    h, g = _pair(ctx, d, 3, u0, p)
    _same(h, g, BASE)
    d = runtime.make_desc(_abi.ALG_VERN9, rhs=_abi.RHS_LORENZ, n_traj=N, tspan=(0.0, 1.0),
                          dt=0.05)
    h, g = _pair(ctx, d, 3, u0, p)
    _same(h, g, BASE)


@pytest.mark.parametrize("lanes", [16, 4], ids=["group", "matrix-core"])
def test_glv16_shared_block_lane_layouts(ctx, lanes):
    u0, r, A = inputs.glv16(N, seed=12)
    d = runtime.make_desc(_abi.ALG_AVERN9, rhs=_abi.RHS_GLV16, n_traj=N, tspan=(0.0, 2.0),
                          dt=0.1, abstol=1e-8, reltol=1e-8, saveat=np.linspace(0.0, 2.0, 5),
                          save_layout=_abi.SAVE_TRAJ_MAJOR, lanes=lanes)
    h, g = _pair(ctx, d, 16, u0, r, A)
    assert ctx.stats()["lanes"] == lanes
    _same(h, g, BASE + ("u_saveat",))


def test_host_counts_optional(ctx):
    u0, p = inputs.lorenz_wide(N, seed=13)
    d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=N, tspan=(0.0, 1.0))
    h = runtime.EnsembleBuffers(d, 3, u0, p)
    h.io.naccept = h.io.nreject = None
    ctx.solve(d, h)
    g = runtime.EnsembleBuffers(d, 3, u0, p, backend="torch", device="cuda:0")
    ctx.solve(d, g)
    _same(h, g, ("u_end", "t_end", "retcode"))


def test_pinned_and_pageable_host_arrays(ctx):
    """Pinned host inputs (page-locked by torch) and pageable host outputs in one
    pipelined solve: same bits as the device-memory solve."""
    torch = pytest.importorskip("torch")
    u0, p, ts = inputs.lorenz_uneven(N, seed=14)
    ts[1] = np.minimum(ts[1], 2.0)
    pins = []
    for a in (u0, p, ts):
        t = torch.empty(a.shape, dtype=torch.float64, pin_memory=True)
        t.numpy()[...] = a
        pins.append(t)
    d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=N, dt=0.1,
                          abstol=1e-6, reltol=1e-4)
    h = runtime.EnsembleBuffers(d, 3, pins[0].numpy(), pins[1].numpy(),
                                tspan=pins[2].numpy())
    assert h.io.u0 == pins[0].data_ptr()  # the pinned memory itself, not a copy
    ctx.solve(d, h)
    g = runtime.EnsembleBuffers(d, 3, u0, p, tspan=ts, backend="torch", device="cuda:0")
    ctx.solve(d, g)
    _same(h, g, BASE)


def test_host_pipeline_after_async_device_solve(ctx):
    """An asynchronous device-memory solve with per-trajectory spans on the context's
    stream (its longest-span-first order lives in the context's scratch), then at once
    a pipelined host-memory solve with saveat and p_shared, whose chunk uploads go to
    that same scratch on the pipeline's own streams: the pipeline must wait for the
    earlier solve.  Both against the oracle."""
    import ctypes
    from helpers import assert_bit_identical
    from oracle import pyoracle
    torch = pytest.importorskip("torch")
    # grow the context's scratch first, so the second call reuses it in place
    big = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=2 * N, tspan=(0.0, 0.5))
    u0b, pb = inputs.lorenz_wide(2 * N, seed=21)
    ctx.solve(big, runtime.EnsembleBuffers(big, 3, u0b, pb))
    # 1. device solve, queued on the context's stream, not waited for
    m = 1 << 17
    u0, p, ts = inputs.lorenz_uneven(m, seed=22)
    ts[1] = np.minimum(ts[1], 4.0)
    d1 = runtime.make_desc(_abi.ALG_AVERN7, rhs=_abi.RHS_LORENZ, n_traj=m, dt=0.1,
                           abstol=1e-8, reltol=1e-8)
    g1 = runtime.EnsembleBuffers(d1, 3, u0, p, tspan=ts, backend="torch", device="cuda:0")
    torch.cuda.synchronize()
    st = ctx.lib.sdeh_solve(ctx.h, ctypes.byref(d1), ctypes.byref(g1.io), _abi.MEM_DEVICE, None, 0)
    assert st == 0, ctx.last_error()
    # 2. at once: host solve, pipelined (>= 2^17 trajectories), saveat and p_shared
    u0g, r, A = inputs.glv16(N, seed=23)
    d2 = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_GLV16, n_traj=N, tspan=(0.0, 5.0),
                           dt=0.1, abstol=1e-6, reltol=1e-6, saveat=np.linspace(0.0, 5.0, 6))
    h2 = runtime.EnsembleBuffers(d2, 16, u0g, r, A)
    ctx.solve(d2, h2)
    ctx.synchronize()
    c1 = runtime.EnsembleBuffers(d1, 3, u0, p, tspan=ts)
    assert pyoracle.solve(d1, c1, nthreads=8) == 0
    assert_bit_identical(g1, c1)
    c2 = runtime.EnsembleBuffers(d2, 16, u0g, r, A)
    assert pyoracle.solve(d2, c2, nthreads=8) == 0
    assert_bit_identical(h2, c2, ("u_end", "t_end", "retcode", "naccept", "nreject", "u_saveat"))
