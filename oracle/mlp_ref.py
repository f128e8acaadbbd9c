"""NumPy reference of the surrogate MLP, forward and backward (test infra).

This module is test infrastructure: only tests, the smoke check and the
bench's CPU leg use it, as the checker of the GPU kernels in
create_rate_subs_amd/csrc/mlp.hip.  It has no counterpart in the reference
(whose rates are always evaluated exactly, rate_subs.py:1845-1880).

FP64 arithmetic with the GPU kernels' BF16 roundings emulated at the same
points (inputs, weights, hidden activations and back-propagated gradients
are BF16; products accumulate in higher precision), so the GPU's FP32
accumulation is the only remaining difference.  GELU in its tanh form; the
pre-activations kept for the backward pass are stored as BF16, as the
training kernels store them.
With ``emulate=False`` every rounding is switched off and the function is
the plain FP64 loss gradient, which tests/test_mlp_ref_pins.py checks
against torch.autograd in float64 and against central differences.
"""

import numpy as np


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _f32(x):
    return np.asarray(x).astype(np.float32).astype(np.float64)


def _same(x):
    return np.asarray(x, dtype=np.float64)


_GELU_C = 0.7978845608028654     # sqrt(2 / pi)
_GELU_A = 0.044715


def gelu(z):
    """tanh-form GELU and its derivative: (g(z), g'(z))."""
    u = _GELU_C * (z + _GELU_A * z ** 3)
    t = np.tanh(u)
    d = 0.5 * (1 + t) + 0.5 * z * (1 - t * t) * _GELU_C * \
        (1 + 3 * _GELU_A * z * z)
    return 0.5 * z * (1 + t), d


def bias_grad(dZ):
    """Gradient of the loss with respect to a layer's bias: sum over rows."""
    return dZ.sum(axis=0)


def mse_grad(yhat, Yt, nv):
    """d mean((yhat - Yt)^2) / d yhat over the first nv outputs."""
    M = yhat.shape[0]
    d = np.zeros_like(yhat)
    d[:, :nv] = 2.0 * (yhat[:, :nv] - Yt[:, :nv]) / (M * nv)
    return d


def forward_backward(X, weights, biases, Yt, nv, emulate=True):
    """Loss-gradient of mean((yhat - Yt)^2) over the first nv outputs.

    X [M][K0] (bf16 values), weights[l] [N_l][K_l] (bf16 values), biases fp32.
    Returns (yhat, dW list, db list).  emulate=False: no BF16 / FP32
    rounding anywhere (the exact FP64 gradient of the same network).
    """
    rb = bf16 if emulate else _same
    r32 = _f32 if emulate else _same
    hs, zs = [X], []
    h = X
    L = len(weights)
    for l in range(L):
        z = h @ weights[l].T + biases[l]
        if l < L - 1:
            a, _ = gelu(z)
            zs.append(rb(z))
            h = rb(a)
            hs.append(h)
        else:
            yhat = r32(z)
    dZ = rb(mse_grad(yhat, Yt, nv))
    dW = [None] * L
    db = [None] * L
    for l in range(L - 1, -1, -1):
        dW[l] = dZ.T @ hs[l]
        db[l] = bias_grad(dZ)
        if l > 0:
            _, g = gelu(zs[l - 1])
            dZ = rb((dZ @ weights[l]) * g)
    return yhat, dW, db


def loss(X, weights, biases, Yt, nv):
    """mean((yhat - Yt)^2) over the first nv outputs, FP64, no rounding."""
    yhat, _dW, _db = forward_backward(X, weights, biases, Yt, nv,
                                      emulate=False)
    return float(((yhat[:, :nv] - Yt[:, :nv]) ** 2).mean())
