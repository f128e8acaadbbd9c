"""NumPy reference of the surrogate MLP, forward and backward (test infra).

FP64 arithmetic with the GPU kernels' BF16 roundings emulated at the same
points (inputs, weights, hidden activations and back-propagated gradients
are BF16; products accumulate in higher precision), so the GPU's FP32
accumulation is the only remaining difference.  GELU in its tanh form.
"""

import numpy as np


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    c = 0.7978845608028654
    u = c * (z + 0.044715 * z ** 3)
    t = np.tanh(u)
    d = 0.5 * (1 + t) + 0.5 * z * (1 - t * t) * c * (1 + 3 * 0.044715 * z * z)
    return 0.5 * z * (1 + t), d


def forward_backward(X, weights, biases, Yt, nv):
    """Loss-gradient of mean((yhat - Yt)^2) over the first nv outputs.

    X [M][K0] (bf16 values), weights[l] [N_l][K_l] (bf16 values), biases fp32.
    Returns (yhat, dW list, db list).
    """
    M = X.shape[0]
    hs, zs = [X], []
    h = X
    L = len(weights)
    for l in range(L):
        z = h @ weights[l].T + biases[l]
        if l < L - 1:
            a, _ = gelu(z)
            zs.append(z.astype(np.float32).astype(np.float64))
            h = bf16(a)
            hs.append(h)
        else:
            yhat = z.astype(np.float32).astype(np.float64)
    d = np.zeros_like(yhat)
    d[:, :nv] = 2.0 * (yhat[:, :nv] - Yt[:, :nv]) / (M * nv)
    dZ = bf16(d)
    dW = [None] * L
    db = [None] * L
    for l in range(L - 1, -1, -1):
        dW[l] = dZ.T @ hs[l]
        db[l] = dZ.sum(axis=0)
        if l > 0:
            _, g = gelu(zs[l - 1])
            dZ = bf16((dZ @ weights[l]) * g)
    return yhat, dW, db
