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
