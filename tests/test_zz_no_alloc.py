"""sdeh_solve_static, the allocation-free variant of sdeh_solve, through the C ABI:
its outputs equal sdeh_solve's bit for bit on two calls in a row (the second with
a smaller ensemble, and with fewer saveat points), it allocates no device memory,
and it refuses what would need per-ensemble scratch."""
import ctypes

import numpy as np
import pytest

from simplediffeq_hip import _abi, inputs, runtime

pytestmark = pytest.mark.gpu

FIELDS = ("u_end", "t_end", "retcode", "naccept", "nreject")


def _bufs(d, u0, p, tspan=None):
    return runtime.EnsembleBuffers(d, u0.shape[0], u0, p, tspan=tspan, backend="torch",
                                   device="cuda:0")


def _static(ctx, d, g, sync=1):
    """sdeh_solve_static through the C ABI on the context's stream."""
    import torch
    torch.cuda.synchronize()
    return ctx.lib.sdeh_solve_static(ctx.h, ctypes.byref(d), ctypes.byref(g.io), None, sync)


def _same(a, b, fields):
    for f in fields:
        x, y = a.numpy(f), b.numpy(f)
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), f


@pytest.mark.parametrize("sched", [_abi.SCHED_PERSISTENT, _abi.SCHED_ONE_LAUNCH, _abi.SCHED_STAGED],
                         ids=["persistent", "one_launch", "staged_shape"])
def test_static_endpoint_equals_solve_twice(ctx, sched):
    for n in (100_003, 4_099):  # second call smaller: the shared queue word is re-armed
        u0, p = inputs.lorenz_wide(n, seed=1)
        d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=n, tspan=(0.0, 10.0),
                              dt=float(np.float32(0.1)), abstol=1e-6, reltol=1e-3,
                              schedule=sched)
        ref = _bufs(d, u0, p)
        ctx.solve(d, ref, sync=True)
        att = ctx.stats()["attempts"]
        g = _bufs(d, u0, p)
        assert _static(ctx, d, g) == 0, ctx.last_error()
        _same(g, ref, FIELDS)
        assert ctx.stats()["attempts"] == att


def test_static_saveat_and_tspans_equal_solve_twice(ctx):
    for n, npts in ((20_000, 64), (3_001, 16)):
        u0, p, ts = inputs.lorenz_costmodel(n, seed=7)
        saveat = np.linspace(0.0, 10.0, npts)
        d = runtime.make_desc(_abi.ALG_AVERN9, rhs=_abi.RHS_LORENZ, n_traj=n,
                              dt=float(np.float32(0.1)), abstol=1e-8, reltol=1e-8,
                              saveat=saveat, schedule=_abi.SCHED_PERSISTENT)
        ref = _bufs(d, u0, p, ts)
        ctx.solve(d, ref, sync=True)
        g = _bufs(d, u0, p, ts)
        assert _static(ctx, d, g) == 0, ctx.last_error()
        _same(g, ref, FIELDS + ("u_saveat",))


def test_static_allocates_nothing(ctx):
    import torch
    n = 65_536
    u0, p = inputs.lorenz_wide(n, seed=1)
    d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=n, tspan=(0.0, 10.0),
                          dt=float(np.float32(0.1)), abstol=1e-6, reltol=1e-3,
                          saveat=np.linspace(0.0, 10.0, 8), schedule=_abi.SCHED_PERSISTENT)
    g = _bufs(d, u0, p)
    ctx.solve(d, g, sync=True)  # loads the kernels and the runtime's copy staging
    assert _static(ctx, d, g) == 0
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(3):
        assert _static(ctx, d, g) == 0
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0


def test_static_refuses_what_needs_scratch(ctx):
    n = 256
    u0, p = inputs.lorenz_wide(n, seed=1)
    kw = dict(rhs=_abi.RHS_LORENZ, n_traj=n, tspan=(0.0, 1.0), dt=0.1)
    d = runtime.make_desc(_abi.ALG_ATSIT5, saveat=np.linspace(0.0, 1.0, 1025), **kw)
    assert _static(ctx, d, _bufs(d, u0, p)) == _abi.E_INVALID
    assert "at most 1024 saveat times" in ctx.last_error()
    d = runtime.make_desc(_abi.ALG_ATSIT5, **kw)
    g = runtime.EnsembleBuffers(d, 3, u0, p, order=np.arange(n), backend="torch", device="cuda:0")
    assert _static(ctx, d, g) == _abi.E_UNSUPPORTED
    d = runtime.make_desc(_abi.ALG_ATSIT5, save_everystep=True, max_saves=8, **kw)
    assert _static(ctx, d, _bufs(d, u0, p)) == _abi.E_UNSUPPORTED
    d = runtime.make_desc(_abi.ALG_ATSIT5, **kw)
    g = _bufs(d, u0, p)
    g.io.naccept = None
    assert _static(ctx, d, g) == _abi.E_INVALID
    assert "naccept and nreject" in ctx.last_error()
