"""The measured configurations at their full sizes, on the GPU with device
buffers: bit for bit against the oracle on every trajectory where the oracle
takes seconds (the 2^20 Lorenz ensemble), on strided subsets where it would take
minutes, and size-independent properties over the whole output (every trajectory
finished, end times snapped to tf, saveat rows complete and finite)."""
import numpy as np
import pytest

from helpers import bits_equal
from simplediffeq_hip import _abi, inputs, runtime

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
DT0 = float(np.float32(0.1))  # the reference default dt = 0.1f0


def _oracle(desc, n_state, cols, u0, p, p_shared=None, tspan=None):
    """Oracle solve of the trajectories `cols` (a copy of desc with n_traj = len(cols))."""
    from oracle import pyoracle
    d = runtime.make_desc(desc.alg, dtype=desc.dtype, rhs=desc.rhs, n_traj=len(cols),
                          tspan=(desc.t0, desc.tf), dt=desc.dt, abstol=desc.abstol,
                          reltol=desc.reltol,
                          saveat=getattr(desc, "_saveat_keep", None))
    b = runtime.EnsembleBuffers(d, n_state, u0[:, cols], p[:, cols], p_shared,
                                None if tspan is None else tspan[:, cols])
    assert pyoracle.solve(d, b, nthreads=16) == 0
    return b


def test_lorenz_wide_2p20_binary64_bitexact_all(ctx):
    """Lorenz, 2^20 trajectories, ATsit5 abstol 1e-6 reltol 1e-3, tspan (0, 10)."""
    n = 1 << 20
    u0, p = inputs.lorenz_wide(n, seed=1)
    d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=n, tspan=(0.0, 10.0),
                          dt=DT0, abstol=1e-6, reltol=1e-3)
    g = runtime.EnsembleBuffers(d, 3, u0, p, backend="torch", device=DEV)
    ctx.solve(d, g)
    c = _oracle(d, 3, np.arange(n), u0, p)
    for f in ("u_end", "t_end", "retcode", "naccept", "nreject"):
        assert bits_equal(g.numpy(f), c.numpy(f)), f
    assert (g.numpy("retcode") == 0).all() and (g.numpy("t_end") == 10.0).all()
    st = ctx.stats()
    assert st["attempts"] == int(c.numpy("naccept").sum() + c.numpy("nreject").sum())


def test_lorenz_wide_2p20_binary32_completes(ctx):
    n = 1 << 20
    u0, p = inputs.lorenz_wide(n, seed=1)
    out = {}
    for dtype in (_abi.F64, _abi.F32):
        d = runtime.make_desc(_abi.ALG_ATSIT5, dtype=dtype, rhs=_abi.RHS_LORENZ, n_traj=n,
                              tspan=(0.0, 10.0), dt=DT0, abstol=1e-6, reltol=1e-3)
        b = runtime.EnsembleBuffers(d, 3, u0, p, backend="torch", device=DEV)
        ctx.solve(d, b)
        out[dtype] = (b, ctx.stats())
    b32, st32 = out[_abi.F32]
    b64, st64 = out[_abi.F64]
    assert (b32.numpy("retcode") == 0).all() and (b32.numpy("t_end") == 10.0).all()
    # same work to within a few percent, and endpoints of the same accuracy class
    assert abs(st32["attempts"] - st64["attempts"]) < 0.05 * st64["attempts"]
    u64 = b64.numpy("u_end")
    rel = np.abs(b32.numpy("u_end").astype(np.float64) - u64).max(axis=0) / np.abs(u64).max(axis=0)
    assert np.median(rel) < 1e-3


def test_lorenz_uneven_2p21_avern7(ctx):
    """sigma 10, beta 8/3, rho ~ U(0, 60), t_end ~ LogUniform(0.5, 20), AVern7 1e-8/1e-8."""
    n = 1 << 21
    u0, p, ts = inputs.lorenz_uneven(n, seed=2)
    d = runtime.make_desc(_abi.ALG_AVERN7, rhs=_abi.RHS_LORENZ, n_traj=n, dt=DT0, abstol=1e-8,
                          reltol=1e-8, schedule=_abi.SCHED_PERSISTENT)
    g = runtime.EnsembleBuffers(d, 3, u0, p, tspan=ts, backend="torch", device=DEV)
    ctx.solve(d, g)
    assert (g.numpy("retcode") == 0).all()
    # every trajectory ends at its own tf (the end snap), to rounding
    assert (np.abs(g.numpy("t_end") - ts[1]) <= 4 * np.spacing(ts[1])).all()
    cols = np.arange(0, n, 8)
    c = _oracle(d, 3, cols, u0, p, tspan=ts)
    for f in ("u_end", "t_end", "retcode", "naccept", "nreject"):
        assert bits_equal(g.numpy(f)[..., cols], c.numpy(f)), f


@pytest.mark.parametrize("dtype", [_abi.F64, _abi.F32], ids=["avern9-binary64", "atsit5-binary32"])
def test_glv16_2p17_saveat256(ctx, dtype):
    """GLV-16, 2^17 trajectories, tspan (0, 50), 256 saveat points."""
    import torch
    n = 1 << 17
    u0, r, A = inputs.glv16(n, seed=3)
    saveat = np.linspace(0.0, 50.0, 256)
    f64 = dtype == _abi.F64
    alg = _abi.ALG_AVERN9 if f64 else _abi.ALG_ATSIT5
    tol = 1e-10 if f64 else 1e-6
    d = runtime.make_desc(alg, dtype=dtype, rhs=_abi.RHS_GLV16, n_traj=n, tspan=(0.0, 50.0),
                          dt=0.1, abstol=tol, reltol=tol, saveat=saveat)
    g = runtime.EnsembleBuffers(d, 16, u0, r, A, backend="torch", device=DEV)
    ctx.solve(d, g)
    assert (g.numpy("retcode") == 0).all() and (g.numpy("t_end") == 50.0).all()
    # every point written and finite; the first row is the initial state itself
    assert bool(torch.isfinite(g.u_saveat).all())
    assert bits_equal(g.u_saveat[0].cpu().numpy(), u0.astype(np.float64 if f64 else np.float32))
    cols = np.arange(0, n, 32)
    sub = g.u_saveat[:, :, torch.as_tensor(cols, device=DEV)].cpu().numpy()
    if f64:
        c = _oracle(d, 16, cols, u0, r, A)
        assert bits_equal(sub, c.numpy("u_saveat"))
        for f in ("u_end", "naccept", "nreject"):
            assert bits_equal(g.numpy(f)[..., cols], c.numpy(f)), f
    else:
        dr = runtime.make_desc(_abi.ALG_AVERN9, rhs=_abi.RHS_GLV16, n_traj=n, tspan=(0.0, 50.0),
                               dt=0.1, abstol=1e-12, reltol=1e-12, saveat=saveat)
        c = _oracle(dr, 16, cols, u0, r, A)
        ref = c.numpy("u_saveat")
        # normwise per (point, trajectory): species near extinction make
        # componentwise relative errors meaningless
        err = np.abs(sub - ref).max(axis=1) / np.abs(ref).max(axis=1)
        assert np.max(err) < 1e-3
