"""The float64 MLP reference (oracle/mlp_ref.py) against central finite
differences: pins the gradient formula the device kernels are checked with."""
import numpy as np

from oracle import mlp_ref


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    w = rng.normal(0, 0.3, 4738)
    x = rng.normal(0, 1, (17, 6))
    y = rng.normal(0, 1, (17, 2))
    _, g = mlp_ref.loss_grad(w, x, y)
    idx = rng.choice(4738, 60, replace=False)
    for j in list(idx) + [0, 383, 384, 447, 448, 4543, 4544, 4671, 4672, 4737]:
        e = np.zeros_like(w)
        e[j] = 1e-6
        fd = (mlp_ref.loss_grad(w + e, x, y)[0] - mlp_ref.loss_grad(w - e, x, y)[0]) / 2e-6
        assert abs(fd - g[j]) < 1e-7 + 1e-5 * abs(g[j]), j
