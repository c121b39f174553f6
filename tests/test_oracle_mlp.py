"""Pins of the MLP oracle (oracle/mlp.py): backpropagation against central
differences of the loss, and the bfloat16 rounding against its definition."""
import numpy as np


def test_backprop_matches_central_differences():
    import mlp
    rng = np.random.default_rng(0)
    theta = rng.standard_normal(mlp.n_params()) * 0.2
    X = rng.standard_normal((37, 11))
    T = rng.standard_normal((37, 9))
    _, g = mlp.loss_and_grad(theta, X, T, exact=True)
    idx = rng.choice(theta.size, 60, replace=False)
    h = 1e-6
    for k in idx:
        tp = theta.copy(); tp[k] += h
        tm = theta.copy(); tm[k] -= h
        fd = (mlp.loss_and_grad(tp, X, T, exact=True)[0] - mlp.loss_and_grad(tm, X, T, exact=True)[0]) / (2 * h)
        assert abs(fd - g[k]) < 1e-7 * max(1.0, abs(fd))


def test_gelu_prime_is_the_derivative():
    import mlp
    z = np.linspace(-4, 4, 101)
    h = 1e-6
    fd = (mlp.gelu(z + h, np.float64) - mlp.gelu(z - h, np.float64)) / (2 * h)
    assert np.abs(fd - mlp.gelu_prime(z, np.float64)).max() < 1e-8


def test_bf16_rounding_is_nearest_even():
    import mlp
    # 1 + 2^-8 lies halfway between 1 and 1 + 2^-7: ties to even -> 1
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 3.0e38], dtype=np.float32)
    r = mlp.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2.0 ** -6 and r[3] == -2.5
    # every rounded value has a zero low half
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)
