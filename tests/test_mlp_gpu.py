"""The HIP MLP kernels (forward, loss + gradient, Adam) against the float64
reference restatement, through the C ABI."""
import numpy as np
import pytest

from oracle import mlp_ref
from simplediffeq_hip import costmodel

pytestmark = pytest.mark.gpu


def test_forward_loss_grad_adam_match_reference(ctx):
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(1)
    w = costmodel.init_params(3)
    n = 3001  # not a multiple of 64: exercises the tail of the last block
    x = rng.uniform(-1, 1, (n, 6)).astype(np.float32)
    y = rng.normal(0, 1, (n, 2)).astype(np.float32)
    cm = costmodel.CostModel(ctx, params=w)
    xd, yd = cm._dev(x), cm._dev(y)
    out = torch.empty((n, 2), dtype=torch.float32, device=dev)
    ctx.mlp_forward(cm.params.data_ptr(), xd.data_ptr(), out.data_ptr(), n)
    torch.cuda.synchronize()  # the library's kernels run on its own stream
    ref, _ = mlp_ref.forward(w.astype(np.float64), x.astype(np.float64))
    assert np.max(np.abs(out.cpu().numpy() - ref)) < 2e-5
    loss, grad = cm.loss_grad(xd, yd)
    torch.cuda.synchronize()
    rl, rg = mlp_ref.loss_grad(w.astype(np.float64), x, y)
    assert abs(loss.item() - rl) < 1e-5 * max(1.0, rl)
    g = grad.cpu().numpy()
    assert np.max(np.abs(g - rg)) < 1e-5 * max(1.0, np.max(np.abs(rg)))
    # one Adam step
    m = np.zeros(4738)
    v = np.zeros(4738)
    wn, _, _ = mlp_ref.adam(w.astype(np.float64), g.astype(np.float64), m, v, 1e-3, 0.9, 0.999, 1e-8, 1)
    ctx.mlp_adam(cm.params.data_ptr(), cm.grad.data_ptr(), cm.m.data_ptr(), cm.v.data_ptr(),
                 1e-3, 0.9, 0.999, 1e-8, 1)
    torch.cuda.synchronize()
    assert np.max(np.abs(cm.params.cpu().numpy() - wn)) < 1e-6


def test_training_reduces_loss(ctx):
    pytest.importorskip("torch")
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, (20000, 6)).astype(np.float32)
    y = np.stack([np.sin(2 * x[:, 0]) + x[:, 3] ** 2, x[:, 1] * x[:, 2]], axis=1).astype(np.float32)
    cm = costmodel.CostModel(ctx, seed=0)
    hist = cm.fit(x, y, steps=300, batch=8192, lr=3e-3, log_every=50)
    assert hist[-1][1] < 0.1 * hist[0][1]
    pred = cm.predict(x[:1000])
    assert np.mean((pred - y[:1000]) ** 2) < 0.1 * np.var(y[:1000])


def plan_buckets(key):
    """sdeh_plan_order's bucket of each FP32 key (-log predicted cost), in the
    library's float32 arithmetic: floor(clamp(-key * 16/ln 2)) + 480, at least 1;
    NaN 0."""
    key = np.asarray(key, dtype=np.float32)
    v = (-key) * np.float32(23.083120654223414)
    c = np.clip(v, np.float32(-480.0), np.float32(543.0))
    b = np.floor(np.nan_to_num(c, nan=0.0)).astype(np.int64) + 480
    b = np.maximum(b, 1)
    b[np.isnan(v)] = 0
    return b


def test_fused_lorenz_plan_against_reference(ctx):
    """sdeh_mlp_plan_lorenz (features, forward pass on the f16 matrix cores with FP32
    accumulation, de-normalisation, sort key, dt0 in one launch) against the float64
    reference network on the numpy features: predictions within the f16 rounding of
    weights and activations; key = -pred_0 and dt0 = exp(pred_1) exactly as the
    kernel's own FP32 arithmetic gives them; and the library's hand-out order
    (sdeh_plan_order) against a stable argsort of the keys on its 1/16-octave
    buckets."""
    import json
    torch = pytest.importorskip("torch")
    from simplediffeq_hip import inputs, runtime
    n = 100_003  # not a multiple of a block's 128 samples
    u0, p, ts = inputs.lorenz_costmodel(n, seed=9)
    w = costmodel.init_params(5)
    cm = costmodel.CostModel(ctx, params=w)
    cm.y_mean = np.array([3.1, -2.2], dtype=np.float32)
    cm.y_std = np.array([0.7, 1.3], dtype=np.float32)
    dev = torch.device("cuda:0")
    pd, u0d, tsd = (torch.as_tensor(a, device=dev) for a in (p, u0, ts))
    stream = torch.cuda.Stream()
    pred_f, order_f, dt0_f = cm.plan_lorenz(pd, u0d, tsd, stream)
    torch.cuda.synchronize()
    pf = pred_f.cpu().numpy()
    ref, _ = mlp_ref.forward(w.astype(np.float64), costmodel.features(p, u0, ts).astype(np.float64))
    ref = ref * cm.y_std + cm.y_mean
    dev_abs = np.abs(pf - ref)
    print(json.dumps({"plan_vs_float64": {"max": float(dev_abs.max()),
                                          "median": float(np.median(dev_abs)),
                                          "p99": float(np.percentile(dev_abs, 99))}}))
    assert np.median(dev_abs) < 5e-3 and dev_abs.max() < 5e-2
    # the kernel's own arithmetic: dt0 = exp(pred_1) in FP32, widened
    assert np.array_equal(dt0_f.cpu().numpy(), np.exp(pf[:, 1]).astype(np.float64)) or \
        np.max(np.abs(dt0_f.cpu().numpy() / np.exp(pf[:, 1].astype(np.float64)) - 1)) < 1e-6
    b = plan_buckets(-pf[:, 0])
    of = order_f.cpu().numpy()
    assert np.array_equal(np.sort(of), np.arange(n))
    assert np.array_equal(b[of], b[np.argsort(-b, kind="stable")])
    # validation: a zero feature scale, and an empty ensemble
    key = torch.empty(4, dtype=torch.float32, device=dev)
    dt0 = torch.empty(4, dtype=torch.float64, device=dev)
    bad = list(costmodel.FEAT_LORENZ)
    bad[3] = 0.0
    with pytest.raises(runtime.SolveError, match="feature scale 1"):
        ctx.mlp_plan_lorenz(cm.params.data_ptr(), pd.data_ptr(), u0d.data_ptr(), tsd.data_ptr(),
                            4, bad, (0, 0, 1, 1), key.data_ptr(), dt0.data_ptr())
    ctx.mlp_plan_lorenz(cm.params.data_ptr(), pd.data_ptr(), u0d.data_ptr(), tsd.data_ptr(), 0,
                        costmodel.FEAT_LORENZ, (0, 0, 1, 1), key.data_ptr(), dt0.data_ptr())
