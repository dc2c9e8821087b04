"""Float64 numpy restatement of the learned preconditioner's network (test
infrastructure): 8 inputs, three hidden layers of 128 with exact GELU, one output m; the
loss mean_i w_i (p_i (1 + m_i) - e_i)^2 and its gradient by backpropagation, with the
parameters packed as W1 b1 W2 b2 W3 b3 W4 b4 (each W inputs x outputs, row-major)."""
import numpy as np
from scipy.special import erf

IN, H = 8, 128
SHAPES = [(IN, H), (H,), (H, H), (H,), (H, H), (H,), (H, 1), (1,)]


def unpack(params):
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(np.asarray(params[o:o + n], dtype=np.float64).reshape(s))
        o += n
    return out


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(params, X):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(params)
    X = np.asarray(X, dtype=np.float64)
    Z1 = X @ W1 + b1
    H1 = gelu(Z1)
    Z2 = H1 @ W2 + b2
    H2 = gelu(Z2)
    Z3 = H2 @ W3 + b3
    H3 = gelu(Z3)
    out = (H3 @ W4 + b4)[:, 0]
    return out, (X, Z1, H1, Z2, H2, Z3, H3)


def loss(params, X, p, e, w):
    m, _ = forward(params, X)
    d = p * (1.0 + m) - e
    return float(np.mean(w * d * d))


def loss_grad(params, X, p, e, w):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(params)
    m, (X, Z1, H1, Z2, H2, Z3, H3) = forward(params, X)
    p, e, w = (np.asarray(a, dtype=np.float64) for a in (p, e, w))
    n = len(p)
    d = p * (1.0 + m) - e
    dO = (2.0 / n) * w * p * d
    gW4 = H3.T @ dO[:, None]
    gb4 = np.array([dO.sum()])
    dZ3 = (dO[:, None] @ W4.T) * gelu_d(Z3)
    gW3 = H2.T @ dZ3
    gb3 = dZ3.sum(0)
    dZ2 = (dZ3 @ W3.T) * gelu_d(Z2)
    gW2 = H1.T @ dZ2
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2.T) * gelu_d(Z1)
    gW1 = X.T @ dZ1
    gb1 = dZ1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4)])
    return float(np.mean(w * d * d)), g
