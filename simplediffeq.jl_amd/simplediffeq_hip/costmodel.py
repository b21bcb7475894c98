"""Learned cost model for adaptive Lorenz ensembles.

A 6 -> 64 -> 64 -> 2 MLP (FP32, ReLU hidden layers) predicts, per trajectory,
  y0 = log(accepted steps)    and    y1 = log(mean accepted step size)
from six features of (sigma, rho, beta, u0, t_end).  It is trained on step
logs (the naccept counters) of the plain adaptive solver, entirely on the GPU
with the library's forward, backward and Adam kernels (sdeh_mlp_*).

Its predictions are used two ways, neither of which lets the network decide
anything the controller decides:
  * binning: trajectories are handed to lanes in decreasing predicted cost
    (the `order` permutation), so a wave holds trajectories of similar length and
    the persistent queue starts the longest first.  Each trajectory's arithmetic
    is unchanged, so results are bit-identical to the unsorted run.
  * dt0 seeding: exp(y1) as the initial step instead of the default 0.1f0.  This
    changes results (within tolerance); the accept/reject decisions are still the
    PI controller's.
"""
import ctypes
import math

import numpy as np

from . import _abi

NIN, NH, NOUT = _abi.MLP_NIN, _abi.MLP_NH, _abi.MLP_NOUT
NPARAM = _abi.MLP_NPARAM

# feature normalisation for the cost-model distribution (sigma ~ U(8,12),
# rho ~ U(0,28), beta ~ U(2,3), u0 ~ U(-1,1)^3, t_end ~ LogUniform(0.5,10))
_LOG_T_MID = 0.5 * (math.log(0.5) + math.log(10.0))
_LOG_T_HALF = 0.5 * (math.log(10.0) - math.log(0.5))
# offset, scale of the features (p_0, p_1, p_2, log span) as features() applies them
FEAT_LORENZ = (10.0, 2.0, 14.0, 14.0, 2.5, 0.5, _LOG_T_MID, _LOG_T_HALF)


def features(p, u0, tspan):
    """(n, 6) float32: sigma, rho, beta, log t_end, |u0|, u0_3, each scaled to ~[-1, 1].
    The Lorenz flow is symmetric under (x, y, z) -> (-x, -y, z), so the cost depends
    on u0 through |u0| and z0; sigma, rho, beta and the horizon enter directly."""
    n = p.shape[1]
    x = np.empty((n, NIN), dtype=np.float32)
    x[:, 0] = (p[0] - 10.0) / 2.0
    x[:, 1] = (p[1] - 14.0) / 14.0
    x[:, 2] = (p[2] - 2.5) / 0.5
    x[:, 3] = (np.log(tspan[1] - tspan[0]) - _LOG_T_MID) / _LOG_T_HALF
    x[:, 4] = np.sqrt((u0 ** 2).sum(axis=0)) / math.sqrt(3.0) * 2.0 - 1.0
    x[:, 5] = u0[2]
    return x


def features_device(p, u0, tspan):
    """features() for device-resident ensemble inputs (torch tensors p (3, n),
    u0 (3, n), tspan (2, n), binary64): the same six scaled features as an (n, 6)
    float32 tensor, computed on the device (no host round trip)."""
    import torch
    n = p.shape[1]
    x = torch.empty((n, NIN), dtype=torch.float32, device=p.device)
    x[:, 0] = (p[0] - 10.0) / 2.0
    x[:, 1] = (p[1] - 14.0) / 14.0
    x[:, 2] = (p[2] - 2.5) / 0.5
    x[:, 3] = (torch.log(tspan[1] - tspan[0]) - _LOG_T_MID) / _LOG_T_HALF
    x[:, 4] = torch.sqrt((u0 ** 2).sum(dim=0)) / math.sqrt(3.0) * 2.0 - 1.0
    x[:, 5] = u0[2]
    return x


def targets(naccept, tspan):
    n = np.maximum(naccept.astype(np.float64), 1.0)
    y = np.empty((naccept.shape[0], NOUT), dtype=np.float32)
    y[:, 0] = np.log(n)
    y[:, 1] = np.log((tspan[1] - tspan[0]) / n)
    return y


def init_params(seed=0):
    """Uniform weights, zero biases (layout of include/simplediffeq_hip.h): He's
    limit sqrt(6 / fan_in) for the two ReLU layers, Glorot's sqrt(6 / (fan_in +
    fan_out)) for the linear output layer."""
    rng = np.random.default_rng(seed)
    w = np.zeros(NPARAM, dtype=np.float32)
    o = 0
    for fan_in, fan_out in ((NIN, NH), (NH, NH), (NH, NOUT)):
        relu_layer = fan_out == NH
        lim = math.sqrt(6.0 / (fan_in if relu_layer else fan_in + fan_out))
        w[o:o + fan_in * fan_out] = rng.uniform(-lim, lim, fan_in * fan_out)
        o += fan_in * fan_out + fan_out
    return w


class CostModel:
    """Device-resident MLP parameters plus Adam state; all math on the GPU."""

    def __init__(self, ctx, params=None, device="cuda:0", seed=0):
        import torch
        self.ctx = ctx
        self.torch = torch
        self.dev = torch.device(device)
        p = init_params(seed) if params is None else np.asarray(params, dtype=np.float32)
        self.params = torch.as_tensor(p, device=self.dev).contiguous()
        self.m = torch.zeros(NPARAM, dtype=torch.float32, device=self.dev)
        self.v = torch.zeros(NPARAM, dtype=torch.float32, device=self.dev)
        self.grad = torch.zeros(NPARAM, dtype=torch.float32, device=self.dev)
        self.loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        # the library's kernels run on the context's stream: the zero fills above
        # (torch's stream) must be complete before the first Adam step reads m, v
        torch.cuda.synchronize(self.dev)
        self.t = 0
        # y normalisation (kept on the host, applied to targets/predictions)
        self.y_mean = np.zeros(NOUT, dtype=np.float32)
        self.y_std = np.ones(NOUT, dtype=np.float32)

    def _dev(self, a):
        return self.torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32),
                                    device=self.dev).contiguous()

    def loss_grad(self, x, y):
        """Mean squared error and its gradient for device tensors x (n,6), y (n,2)."""
        n = x.shape[0]
        self.ctx.mlp_loss_grad(self.params.data_ptr(), x.data_ptr(), y.data_ptr(), n,
                               self.loss.data_ptr(), self.grad.data_ptr())
        return self.loss, self.grad

    def fit(self, x, y, steps=600, batch=65536, lr=3e-3, seed=0, log_every=0):
        """Adam on minibatches; x, y host arrays.  Returns the loss history."""
        torch = self.torch
        self.y_mean = y.mean(axis=0).astype(np.float32)
        self.y_std = (y.std(axis=0) + 1e-6).astype(np.float32)
        xd = self._dev(x)
        yd = self._dev((y - self.y_mean) / self.y_std)
        n = x.shape[0]
        rng = np.random.default_rng(seed)
        hist = []
        for s in range(steps):
            if batch >= n:
                xb, yb = xd, yd
            else:
                i0 = int(rng.integers(0, n - batch + 1))
                xb, yb = xd[i0:i0 + batch], yd[i0:i0 + batch]
            self.loss_grad(xb, yb)
            self.t += 1
            self.ctx.mlp_adam(self.params.data_ptr(), self.grad.data_ptr(), self.m.data_ptr(),
                              self.v.data_ptr(), lr, 0.9, 0.999, 1e-8, self.t)
            if log_every and (s % log_every == 0 or s == steps - 1):
                torch.cuda.synchronize()
                hist.append((s, float(self.loss.item())))
        torch.cuda.synchronize()
        return hist

    def predict(self, x):
        """(n, 2) predictions in target units."""
        torch = self.torch
        xd = self._dev(x)
        yd = torch.empty((x.shape[0], NOUT), dtype=torch.float32, device=self.dev)
        self.ctx.mlp_forward(self.params.data_ptr(), xd.data_ptr(), yd.data_ptr(), x.shape[0])
        torch.cuda.synchronize()
        return yd.cpu().numpy() * self.y_std + self.y_mean

    def order_by_cost(self, pred):
        """Trajectory hand-out order: decreasing predicted accepted-step count."""
        return np.argsort(-pred[:, 0], kind="stable").astype(np.int64)

    def dt0(self, pred):
        return np.exp(pred[:, 1]).astype(np.float64)

    def plan_device(self, xd, stream):
        """The whole binning step on the device, ordered on the torch stream
        `stream`: predictions for the (n, 6) float32 features `xd` (the library's
        forward kernel), the hand-out order (stable sort by decreasing predicted
        cost) and the seeded dt0.  Returns (pred (n, 2), order int64, dt0 float64),
        all device tensors; nothing travels to the host."""
        torch = self.torch
        n = xd.shape[0]
        with torch.cuda.stream(stream):
            yd = torch.empty((n, NOUT), dtype=torch.float32, device=self.dev)
            self.ctx.mlp_forward(self.params.data_ptr(), xd.data_ptr(), yd.data_ptr(), n,
                                 stream=stream.cuda_stream)
            pred = yd * torch.as_tensor(self.y_std, device=self.dev) + \
                torch.as_tensor(self.y_mean, device=self.dev)
            order = torch.argsort(-pred[:, 0], stable=True)
            dt0 = torch.exp(pred[:, 1]).to(torch.float64)
        return pred, order, dt0

    def plan_lorenz(self, p, u0, tspan, stream, pred=True):
        """The binning step on the raw binary64 inputs of a Lorenz ensemble (device
        tensors p (3, n), u0 (3, n), tspan (2, n)), entirely in the library:
        features, forward pass, de-normalisation, sort key and dt0 in one launch
        (sdeh_mlp_plan_lorenz), then the hand-out order from the keys
        (sdeh_plan_order: counting sort at 1/16 octave of the predicted cost).
        Returns (pred (n, 2) or None, order int64, dt0 float64), device tensors."""
        torch = self.torch
        n = p.shape[1]
        with torch.cuda.stream(stream):
            key = torch.empty(n, dtype=torch.float32, device=self.dev)
            dt0 = torch.empty(n, dtype=torch.float64, device=self.dev)
            order = torch.empty(n, dtype=torch.int64, device=self.dev)
            pr = torch.empty((n, NOUT), dtype=torch.float32, device=self.dev) if pred else None
            norm = (self.y_mean[0], self.y_mean[1], self.y_std[0], self.y_std[1])
            self.ctx.mlp_plan_lorenz(self.params.data_ptr(), p.data_ptr(), u0.data_ptr(),
                                     tspan.data_ptr(), n, FEAT_LORENZ, norm, key.data_ptr(),
                                     dt0.data_ptr(), pr.data_ptr() if pred else None,
                                     stream=stream.cuda_stream)
            self.ctx.plan_order(key.data_ptr(), n, order.data_ptr(), stream=stream.cuda_stream)
        return pr, order, dt0


def wave_efficiency(attempts, order=None, wave=64):
    """Fraction of lane-iterations doing work in a one-lane-per-trajectory launch:
    sum(attempts) / sum over waves of (wave * max attempts in the wave)."""
    a = np.asarray(attempts, dtype=np.float64)
    if order is not None:
        a = a[order]
    pad = (-len(a)) % wave
    if pad:
        a = np.concatenate([a, np.zeros(pad)])
    w = a.reshape(-1, wave)
    return float(a.sum() / (wave * w.max(axis=1)).sum())


def train_lorenz(ctx, n=1 << 20, seed=40, steps=1500, batch=65536, lr=3e-3, abstol=1e-7,
                 reltol=1e-5, model_seed=0, device="cuda:0", log_every=200):
    """Step logs and training of the Lorenz cost model, all on the GPU: n plain
    binary64 GPUSimpleATsit5 solves of the training distribution (sigma ~ U(8,12),
    rho ~ U(0,28), beta ~ U(2,3), u0 ~ U(-1,1)^3, t_end ~ LogUniform(0.5,10); seed
    `seed`) give the accepted-step counts, and `steps` Adam steps of the library's
    MLP kernels fit the network to them.  Returns (model, loss history, seconds of
    training, seconds of the step-log solve)."""
    import time
    import torch
    from . import inputs, runtime
    u0, p, ts = inputs.lorenz_costmodel(n, seed=seed)
    d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=n,
                          dt=float(np.float32(0.1)), abstol=abstol, reltol=reltol)
    g = runtime.EnsembleBuffers(d, 3, u0, p, tspan=ts, backend="torch", device=device)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.solve(d, g, sync=True)
    t_logs = time.perf_counter() - t0
    if not (g.numpy("retcode") == 0).all():
        raise RuntimeError("step-log solves did not all finish")
    x = features(p, u0, ts)
    y = targets(g.numpy("naccept"), ts)
    del g
    cm = CostModel(ctx, device=device, seed=model_seed)
    t0 = time.perf_counter()
    hist = cm.fit(x, y, steps=steps, batch=batch, lr=lr, log_every=log_every)
    return cm, hist, time.perf_counter() - t0, t_logs
