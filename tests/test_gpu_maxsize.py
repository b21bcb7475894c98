"""An ensemble larger than 2^31 trajectories in one solve: 2,147,549,184 binary64
Lorenz trajectories (197.6 GB of inputs and outputs in HBM, generated on the
device).  Every trajectory must finish, and those on both sides of the int32
index boundary and at the very end are bit for bit the oracle's: any 32-bit
index arithmetic in the kernels or the runtime would break exactly there."""
import numpy as np
import pytest

from helpers import bits_equal
from simplediffeq_hip import _abi, runtime

pytestmark = pytest.mark.gpu

N = (1 << 31) + (1 << 16)


def test_lorenz_beyond_int32_trajectories():
    import torch
    from oracle import pyoracle
    torch.cuda.empty_cache()
    free, _ = torch.cuda.mem_get_info()
    need = N * (3 * 8 * 3 + 8 + 3 * 4)
    if free < need + (4 << 30):
        pytest.skip("needs %.0f GB of free device memory, %.0f GB free" % (need / 1e9, free / 1e9))
    dev = "cuda:0"
    g = torch.Generator(device=dev)
    g.manual_seed(11)
    u0 = torch.empty((3, N), dtype=torch.float64, device=dev).uniform_(-1.0, 1.0, generator=g)
    p = torch.empty((3, N), dtype=torch.float64, device=dev)
    p[0].uniform_(8.0, 12.0, generator=g)
    p[1].uniform_(0.0, 21.0, generator=g)
    p[2].uniform_(2.0, 3.0, generator=g)
    ctx = runtime.Context(0)
    try:
        for sched in (_abi.SCHED_ONE_LAUNCH, _abi.SCHED_PERSISTENT):
            d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=N,
                                  tspan=(0.0, 0.3), dt=float(np.float32(0.1)), abstol=1e-6,
                                  reltol=1e-3, schedule=sched)
            b = runtime.EnsembleBuffers(d, 3, u0, p, backend="torch", device=dev)
            ctx.solve(d, b)
            assert int((b.retcode != 0).sum()) == 0
            assert bool((b.t_end == 0.3).all())
            st = ctx.stats()
            assert st["accepts"] == int(b.naccept.sum(dtype=torch.int64))
            # trajectories around 2^31 and at the end, against the oracle
            idx = np.concatenate([np.arange(0, 64), np.arange((1 << 31) - 256, (1 << 31) + 256),
                                  np.arange(N - 64, N)])
            ti = torch.as_tensor(idx, device=dev)
            cu0 = u0[:, ti].cpu().numpy()
            cp = p[:, ti].cpu().numpy()
            dc = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=len(idx),
                                   tspan=(0.0, 0.3), dt=float(np.float32(0.1)), abstol=1e-6,
                                   reltol=1e-3)
            c = runtime.EnsembleBuffers(dc, 3, cu0, cp)
            assert pyoracle.solve(dc, c, nthreads=8) == 0
            for f in ("u_end", "t_end", "naccept", "nreject", "retcode"):
                got = getattr(b, f)[..., ti].cpu().numpy()
                assert bits_equal(got, c.numpy(f)), (sched, f)
            del b
    finally:
        ctx.close()
        del u0, p
        torch.cuda.empty_cache()
