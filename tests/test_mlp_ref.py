"""The numpy network of tests/mlp_ref.py against central finite differences: the
backpropagated gradient matches the numerical derivative of the loss for parameters of
every layer, and m = 0 (zero output layer) gives the loss of the plain preconditioner."""
import numpy as np

import mlp_ref


def random_case(seed, n=16):
    rng = np.random.default_rng(seed)
    params = rng.normal(scale=0.2, size=sum(int(np.prod(s)) for s in mlp_ref.SHAPES))
    X = rng.normal(size=(n, mlp_ref.IN))
    p, e = rng.normal(size=n), rng.normal(size=n)
    w = rng.uniform(0.5, 2.0, size=n)
    return params, X, p, e, w


def test_gradient_against_finite_differences():
    params, X, p, e, w = random_case(3)
    _, g = mlp_ref.loss_grad(params, X, p, e, w)
    rng = np.random.default_rng(4)
    offs = np.cumsum([0] + [int(np.prod(s)) for s in mlp_ref.SHAPES])
    # a few parameters of every block (weights and biases of all four layers)
    picks = [o + int(rng.integers(0, offs[i + 1] - o)) for i, o in enumerate(offs[:-1]) for _ in range(3)]
    h = 1e-6
    for k in picks:
        a, b = params.copy(), params.copy()
        a[k] += h
        b[k] -= h
        fd = (mlp_ref.loss(a, X, p, e, w) - mlp_ref.loss(b, X, p, e, w)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_zero_output_layer_is_plain_preconditioner():
    params, X, p, e, w = random_case(5)
    n4 = mlp_ref.H + 1
    params[-n4:] = 0.0
    m, _ = mlp_ref.forward(params, X)
    assert np.all(m == 0.0)
    assert abs(mlp_ref.loss(params, X, p, e, w) - np.mean(w * (p - e) ** 2)) < 1e-14
