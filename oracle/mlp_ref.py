"""ORACLE (cost-model MLP) — TEST INFRASTRUCTURE ONLY.

Float64 numpy restatement of the 6 -> 64 -> 64 -> 2 ReLU MLP and its mean
squared error gradient, in the parameter layout of include/simplediffeq_hip.h
(W1[64][6], b1[64], W2[64][64], b2[64], W3[2][64], b3[2]).  Used to check the
device forward/backward kernels and one Adam step.
"""
import numpy as np

NIN, NH, NOUT = 6, 64, 2


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    o = 0
    out = []
    for shape in ((NH, NIN), (NH,), (NH, NH), (NH,), (NOUT, NH), (NOUT,)):
        k = int(np.prod(shape))
        out.append(w[o:o + k].reshape(shape))
        o += k
    assert o == w.size
    return out


def forward(w, x):
    W1, b1, W2, b2, W3, b3 = unpack(w)
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ W3.T + b3, (h1, h2)


def loss_grad(w, x, y):
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(w)
    out, (h1, h2) = forward(w, x)
    e = out - y
    n = x.shape[0]
    loss = np.mean(e ** 2)
    g = 2.0 * e / (n * NOUT)
    gW3 = g.T @ h2
    gb3 = g.sum(axis=0)
    d2 = (g @ W3) * (h2 > 0)  # ReLU derivative, 0 at 0
    gW2 = d2.T @ h1
    gb2 = d2.sum(axis=0)
    d1 = (d2 @ W2) * (h1 > 0)
    gW1 = d1.T @ x
    gb1 = d1.sum(axis=0)
    grad = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, grad


def adam(w, g, m, v, lr, b1, b2, eps, t):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
