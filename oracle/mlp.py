"""CPU oracle of the per-aggregate parameter network (TEST INFRASTRUCTURE ONLY:
imported by tests/, never by the product path).

MLP 12 -> 64 -> 64 -> 2, GELU (exact, erf form), parameters in one vector with
the layout of include/hazmath_mi355x.h: W1[16][64] (rows 12..15 padding),
b1[64], W2[64][64], b2[64], W3[64][16] (columns 2..15 padding), b3[16].
Loss: mean squared error over the two outputs.  Float64 throughout.
The network has no counterpart in the reference; this restatement is pinned by
finite differences and by PyTorch CPU autograd in tests/test_oracle_golden.py.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

N_IN, N_HID, N_OUT, IN_PAD, OUT_PAD = 12, 64, 2, 16, 16
OFF_W1 = 0
OFF_B1 = OFF_W1 + IN_PAD * N_HID
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_HID * OUT_PAD
N_PARAMS = OFF_B3 + OUT_PAD


def unpack(theta):
    t = np.asarray(theta, dtype=np.float64)
    W1 = t[OFF_W1:OFF_B1].reshape(IN_PAD, N_HID)[:N_IN]
    b1 = t[OFF_B1:OFF_W2]
    W2 = t[OFF_W2:OFF_B2].reshape(N_HID, N_HID)
    b2 = t[OFF_B2:OFF_W3]
    W3 = t[OFF_W3:OFF_B3].reshape(N_HID, OUT_PAD)[:, :N_OUT]
    b3 = t[OFF_B3:OFF_B3 + N_OUT]
    return W1, b1, W2, b2, W3, b3


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(theta, x):
    W1, b1, W2, b2, W3, b3 = unpack(theta)
    z1 = x @ W1 + b1
    h1 = gelu(z1)
    z2 = h1 @ W2 + b2
    h2 = gelu(z2)
    y = h2 @ W3 + b3
    return y, (x, z1, h1, z2, h2)


def loss(theta, x, t):
    y, _ = forward(theta, x)
    return float(np.mean((y - t) ** 2))


def loss_grad(theta, x, t):
    """(loss, gradient in the packed layout; padding entries are zero)."""
    W1, b1, W2, b2, W3, b3 = unpack(theta)
    y, (x, z1, h1, z2, h2) = forward(theta, x)
    B = x.shape[0]
    dy = 2.0 * (y - t) / (B * N_OUT)
    g = np.zeros(N_PARAMS)
    gW3 = h2.T @ dy
    gb3 = dy.sum(0)
    dz2 = (dy @ W3.T) * gelu_grad(z2)
    gW2 = h1.T @ dz2
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2.T) * gelu_grad(z1)
    gW1 = x.T @ dz1
    gb1 = dz1.sum(0)
    g[OFF_W1:OFF_B1].reshape(IN_PAD, N_HID)[:N_IN] = gW1
    g[OFF_B1:OFF_W2] = gb1
    g[OFF_W2:OFF_B2] = gW2.reshape(-1)
    g[OFF_B2:OFF_W3] = gb2
    g[OFF_W3:OFF_B3].reshape(N_HID, OUT_PAD)[:, :N_OUT] = gW3
    g[OFF_B3:OFF_B3 + N_OUT] = gb3
    return float(np.mean((y - t) ** 2)), g
