"""FP64 CPU reference of the surrogate MLP (checker of the GPU's bf16-MFMA forward/backward kernels).

TEST INFRASTRUCTURE ONLY (imported by tests/).  The network has no counterpart in the reference: it
stands in for scattering orders >= 2 of the LUT precompute.  Layout, encoding and loss mirror
sfsim_amd/csrc/surrogate.hip:

* features (192): for each of 10 inputs x in [0, 1] (4 LUT coordinates, 6 atmosphere parameters)
  x itself, then sin(2^b pi x) and cos(2^b pi x) for b = 0..7, zero padded from 170 to 192;
* layers 192 -> 128 -> 128 -> 128 -> 16, SiLU after the three hidden layers, linear output;
* parameters, flattened in this order: W1 [192 x 128] (row = input feature), b1 [128], W2 [128 x 128],
  b2, W3 [128 x 128], b3, W4 [128 x 16], b4 [16];
* loss = mean over rows of sum over the outputs of mask * (y - target)^2 / 2 (the library uses the
  first 15 outputs: log order sums P_1..P_5, RGB).
"""
import numpy as np

NIN, NFEAT, NH, NOUTP, NBANDS = 10, 192, 128, 16, 8
SHAPES = [(NFEAT, NH), (NH,), (NH, NH), (NH,), (NH, NH), (NH,), (NH, NOUTP), (NOUTP,)]


def n_params():
    return int(sum(np.prod(s) for s in SHAPES))


def unflatten(theta):
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(np.asarray(theta[o:o + n], np.float64).reshape(s))
        o += n
    return out


def encode(x):
    """x: [B, 10] in [0, 1] -> [B, 192] features."""
    x = np.asarray(x, np.float64)
    cols = []
    for i in range(NIN):
        cols.append(x[:, i:i + 1])
        for b in range(NBANDS):
            a = (2.0 ** b) * np.pi * x[:, i:i + 1]
            cols.append(np.sin(a))
            cols.append(np.cos(a))
    f = np.concatenate(cols, axis=1)
    return np.concatenate([f, np.zeros((f.shape[0], NFEAT - f.shape[1]))], axis=1)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(theta, X):
    W1, b1, W2, b2, W3, b3, W4, b4 = unflatten(theta)
    z1 = X @ W1 + b1
    a1 = silu(z1)
    z2 = a1 @ W2 + b2
    a2 = silu(z2)
    z3 = a2 @ W3 + b3
    a3 = silu(z3)
    y = a3 @ W4 + b4
    return y, (X, z1, a1, z2, a2, z3, a3)


def loss_and_grad(theta, X, T, mask):
    """T, mask: [B, 16]; returns (loss, flat gradient)."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unflatten(theta)
    y, (X, z1, a1, z2, a2, z3, a3) = forward(theta, X)
    B = X.shape[0]
    d = (y - T) * mask
    loss = 0.5 * np.sum(d * d) / B
    dy = d / B
    gW4 = a3.T @ dy
    gb4 = dy.sum(0)
    d3 = (dy @ W4.T) * dsilu(z3)
    gW3 = a2.T @ d3
    gb3 = d3.sum(0)
    d2 = (d3 @ W3.T) * dsilu(z2)
    gW2 = a1.T @ d2
    gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * dsilu(z1)
    gW1 = X.T @ d1
    gb1 = d1.sum(0)
    g = np.concatenate([a.reshape(-1) for a in (gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4)])
    return loss, g
