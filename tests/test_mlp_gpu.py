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
