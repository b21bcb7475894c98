"""sdeh_plan_order: the cost model's keys (FP32, -log predicted cost) to the
hand-out order sdeh_io.order takes, counting-sorted in the library on a 1/16
octave grid of the predicted cost.  Checked against numpy's stable argsort of the
bucketed keys: the order is a permutation and its bucket sequence equals the
stable sort's (within a bucket the order is unspecified)."""
import numpy as np
import pytest

from simplediffeq_hip import runtime
from test_mlp_gpu import plan_buckets

pytestmark = pytest.mark.gpu


def _order(ctx, key):
    import torch
    dev = torch.device("cuda:0")
    kd = torch.as_tensor(np.asarray(key, dtype=np.float32), device=dev)
    od = torch.full((len(key),), -1, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    ctx.plan_order(kd.data_ptr(), len(key), od.data_ptr())
    ctx.synchronize()
    return od.cpu().numpy()


@pytest.mark.parametrize("n", [1, 2, 63, 257, 100_003, 1 << 20])
def test_plan_order_matches_stable_bucket_sort(ctx, n):
    rng = np.random.default_rng(n)
    key = -rng.normal(4.0, 1.5, n).astype(np.float32)  # log costs around e^4
    if n > 10:
        key[:: max(1, n // 7)] = np.nan                # NaN keys go last
        key[1] = -np.inf                                # cost +inf: first bucket
        key[2] = np.inf                                 # cost 0: a low bucket
        key[3] = -1000.0                                # beyond the top octave
        key[4] = 40.0                                   # below the bottom octave
        key[5:9] = key[9]                               # ties
    o = _order(ctx, key)
    b = plan_buckets(key)
    assert np.array_equal(np.sort(o), np.arange(n))
    ref = np.argsort(-b, kind="stable")
    assert np.array_equal(b[o], b[ref])
    assert (np.diff(b[o]) <= 0).all()


def test_plan_order_rejects_bad_arguments(ctx):
    import torch
    k = torch.zeros(4, dtype=torch.float32, device="cuda:0")
    o = torch.zeros(4, dtype=torch.int64, device="cuda:0")
    ctx.plan_order(k.data_ptr(), 0, o.data_ptr())  # empty: nothing to do
    with pytest.raises(runtime.SolveError, match="not a device allocation|do not fit"):
        ctx.plan_order(k.data_ptr(), 1 << 30, o.data_ptr())  # far beyond the buffers
    with pytest.raises(runtime.SolveError, match="below 2\\^32"):
        ctx.plan_order(k.data_ptr(), 1 << 32, o.data_ptr())


@pytest.mark.parametrize("case", ["one_launch", "persistent", "two_per_lane", "fixed"])
def test_planned_order_is_trusted_and_read_in_bounds(ctx, case):
    """The order sdeh_plan_order last wrote goes to the solve without the permutation
    check and gives the same bits as no order; overwritten with entries outside
    [0, n), it still runs (the kernels never dereference such an entry: the two
    trajectories whose slots were overwritten stay unsolved, every other one is
    unchanged); once another buffer has been planned, the overwritten one is checked
    and refused."""
    import torch
    from simplediffeq_hip import _abi, inputs
    n = 20_000
    u0, p = inputs.lorenz_wide(n, seed=3)
    kw = dict(rhs=_abi.RHS_LORENZ, n_traj=n, tspan=(0.0, 3.0), dt=0.1)
    if case == "fixed":
        d = runtime.make_desc(_abi.ALG_TSIT5, **kw)
    else:
        sched = _abi.SCHED_ONE_LAUNCH if case != "persistent" else _abi.SCHED_PERSISTENT
        d = runtime.make_desc(_abi.ALG_ATSIT5, abstol=1e-6, reltol=1e-4, schedule=sched,
                              per_lane=2 if case == "two_per_lane" else 0, **kw)
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(11)
    kd = torch.as_tensor(-rng.normal(4.0, 1.5, n).astype(np.float32), device=dev)
    od = torch.full((n,), -1, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    ctx.plan_order(kd.data_ptr(), n, od.data_ptr())
    ctx.synchronize()
    o = od.cpu().numpy()
    assert np.array_equal(np.sort(o), np.arange(n))

    def run(order):
        b = runtime.EnsembleBuffers(d, 3, u0, p, order=order, backend="torch", device=dev)
        b.retcode.fill_(-7)
        torch.cuda.synchronize()
        ctx.solve(d, b)
        return b

    ref = run(None)
    got = run(od)
    assert got.order.data_ptr() == od.data_ptr()  # the planned buffer itself
    for f in ("u_end", "t_end", "naccept", "nreject", "retcode"):
        assert np.array_equal(got.numpy(f), ref.numpy(f)), f
    # overwrite two slots with entries that name no trajectory
    lost = o[[0, n // 2]]
    od[0] = n + 5
    od[n // 2] = -3
    torch.cuda.synchronize()
    bad = run(od)
    keep = np.ones(n, dtype=bool)
    keep[lost] = False
    rc = bad.numpy("retcode")
    assert (rc[lost] == -7).all()  # untouched
    for f in ("u_end", "t_end", "naccept", "nreject", "retcode"):
        assert np.array_equal(bad.numpy(f)[..., keep], ref.numpy(f)[..., keep]), f
    # planning into another buffer ends the trust in the first one
    od2 = torch.empty_like(od)
    ctx.plan_order(kd.data_ptr(), n, od2.data_ptr())
    ctx.synchronize()
    with pytest.raises(runtime.SolveError, match="permutation"):
        run(od)
