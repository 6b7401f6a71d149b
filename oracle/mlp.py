"""CPU oracle of the learned saturation guess: the 3 -> 64 -> 64 -> 1 tanh
network, its loss gradient and its bf16 inference, in float64 numpy.

TEST INFRASTRUCTURE ONLY: the checker of the library's training and inference
kernels.  Only ``tests/`` may import it.

Parameter vector layout (FP32 on the device): W1 [64][3], b1 [64],
W2 [64][64] ([out][in]), b2 [64], W3 [64], b3.  Inputs are normalised as
x_n = (x - mu) * isd for (s, qt, p0); the target is (T - mu_T) / sd_T; the
loss is the mean of 0.5 (y - t)^2.
"""
from __future__ import annotations

import numpy as np

NIN, NH = 3, 64
OW1 = 0
OB1 = OW1 + NH * NIN
OW2 = OB1 + NH
OB2 = OW2 + NH * NH
OW3 = OB2 + NH
OB3 = OW3 + NH
NPARAM = OB3 + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OW1:OB1].reshape(NH, NIN), p[OB1:OW2], p[OW2:OB2].reshape(NH, NH), p[OB2:OW3], p[OW3:OB3], p[OB3])


def normalise(data, norm):
    """data [n][4] = (s, qt, p0, T) -> inputs [n][3] and targets [n]."""
    d = np.asarray(data, dtype=np.float64)
    x = (d[:, :3] - np.asarray(norm["mu"], np.float64)) * np.asarray(norm["isd"], np.float64)
    t = (d[:, 3] - norm["mu_T"]) / norm["sd_T"]
    return x, t


def forward(p, x):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ W3 + b3, (h1, h2)


def loss_and_grad(p, x, t):
    """Sum of 0.5 (y - t)^2 and the gradient of its mean."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    y, (h1, h2) = forward(p, x)
    e = y - t
    n = len(t)
    dy = e / n
    g = np.zeros(NPARAM)
    g[OW3:OB3] = dy @ h2
    g[OB3] = dy.sum()
    dz2 = dy[:, None] * W3[None, :] * (1.0 - h2 * h2)
    g[OW2:OB2] = (dz2.T @ h1).ravel()
    g[OB2:OW3] = dz2.sum(axis=0)
    dz1 = (dz2 @ W2) * (1.0 - h1 * h1)
    g[OW1:OB1] = (dz1.T @ x).ravel()
    g[OB1:OW2] = dz1.sum(axis=0)
    return 0.5 * float(np.sum(e * e)), g


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest, ties to even), as float64."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


# tanh(x) = 1 - 2 / (2^(TANH_C x) + 1); the packed hidden-layer weights are
# TANH_C W2 rounded to bf16 (FP32 value of 2 log2(e))
TANH_C = float(np.float32(2.8853900817779268))


def infer_bf16(p, x):
    """What the packed inference computes: W1, TANH_C W2 (then divided by
    TANH_C), W3 and the first hidden activations rounded to bf16, biases FP32,
    exact tanh, float64 sums."""
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, dtype=np.float32))
    W1, W3 = bf16_round(W1), bf16_round(W3)
    W2 = bf16_round(np.float32(TANH_C) * W2.astype(np.float32)) / TANH_C
    h1 = bf16_round(np.tanh(x @ W1.T + b1))
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ W3 + b3
