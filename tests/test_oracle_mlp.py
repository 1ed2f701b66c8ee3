"""The FP64 reference of the surrogate MLP (oracle/mlp.py), pinned by what the mathematics fixes:
central finite differences of the loss for every parameter block, the feature encoding against
closed forms, and SiLU's derivative."""
import numpy as np
import pytest

from oracle import mlp


def _theta(rng, scale=0.3):
    return rng.normal(0.0, scale, mlp.n_params())


def test_parameter_count_matches_layout():
    assert mlp.n_params() == 192 * 128 + 128 + 2 * (128 * 128 + 128) + 128 * 16 + 16


def test_encoding_closed_form():
    x = np.array([[0.0, 0.25, 0.5, 1.0, 0.1, 0.2, 0.3, 0.4, 0.6, 0.9]])
    f = mlp.encode(x)
    assert f.shape == (1, 192)
    assert np.all(f[0, 170:] == 0.0)
    # input 1 (x = 0.25): features 17..33 are x, sin(2^b pi x), cos(2^b pi x)
    base = 17
    assert f[0, base] == 0.25
    assert f[0, base + 1] == pytest.approx(np.sin(np.pi / 4))       # b = 0
    assert f[0, base + 2] == pytest.approx(np.cos(np.pi / 4))
    assert f[0, base + 3] == pytest.approx(1.0)                     # b = 1: sin(pi / 2)
    assert f[0, base + 4] == pytest.approx(0.0, abs=1e-15)
    # x = 0: every sine 0, every cosine 1
    assert np.all(f[0, 1:17:2] == 0.0) and np.all(f[0, 2:17:2] == 1.0)


def test_dsilu_is_derivative():
    z = np.linspace(-6, 6, 101)
    h = 1e-6
    num = (mlp.silu(z + h) - mlp.silu(z - h)) / (2 * h)
    assert np.max(np.abs(num - mlp.dsilu(z))) < 1e-8


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    theta = _theta(rng)
    X = mlp.encode(rng.uniform(0, 1, (7, 10)))
    T = rng.normal(0, 1, (7, 16))
    mask = np.zeros((7, 16))
    mask[:, :9] = 1.0
    loss, g = mlp.loss_and_grad(theta, X, T, mask)
    assert loss > 0
    # a few parameters of every block, including the masked-out outputs' W4 columns (gradient 0)
    offsets = np.cumsum([0] + [int(np.prod(s)) for s in mlp.SHAPES])
    idx = []
    for b in range(len(mlp.SHAPES)):
        idx += list(rng.integers(offsets[b], offsets[b + 1], 6))
    w4 = offsets[6]
    idx += [w4 + 5 * 16 + 12, offsets[7] + 15]  # W4[5][12], b4[15]: masked outputs
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        num = (mlp.loss_and_grad(tp, X, T, mask)[0] - mlp.loss_and_grad(tm, X, T, mask)[0]) / (2 * h)
        assert abs(num - g[i]) <= 2e-5 * max(1.0, abs(num)), (i, num, g[i])
    assert g[w4 + 5 * 16 + 12] == 0.0 and g[offsets[7] + 15] == 0.0


def test_planted_mistake_is_caught():
    """A transposed W2 in the forward pass changes the output (the GPU tests rely on this)."""
    rng = np.random.default_rng(6)
    theta = _theta(rng)
    X = mlp.encode(rng.uniform(0, 1, (3, 10)))
    y, _ = mlp.forward(theta, X)
    parts = mlp.unflatten(theta)
    parts[2] = parts[2].T.copy()
    y2, _ = mlp.forward(np.concatenate([p.reshape(-1) for p in parts]), X)
    assert np.max(np.abs(y - y2)) > 1e-3
