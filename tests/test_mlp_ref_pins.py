"""Pins of oracle/mlp_ref.forward_backward, the checker of the surrogate's
forward and backward kernels.

With its BF16/FP32 roundings switched off, the oracle's gradient of the
mean squared error must equal

* torch.autograd on the same network written with torch ops in float64
  (GELU from torch.nn.functional.gelu(approximate='tanh'), an independent
  implementation of the tanh form), and
* central differences of the loss itself.

Planted mistakes in the oracle (GELU' dropped, bias gradient dropped or
averaged, a wrong GELU constant, the MSE gradient's factor 2 lost) must
each fail the same comparisons.  The GELU derivative is also checked
against central differences of GELU alone.
"""

import numpy as np
import pytest
import torch

from oracle import mlp_ref


def _net(seed, dims=(7, 12, 10, 9, 5), M=6):
    rng = np.random.default_rng(seed)
    Ws = [rng.normal(0, 1.0 / np.sqrt(dims[i]), size=(dims[i + 1], dims[i]))
          for i in range(len(dims) - 1)]
    bs = [rng.normal(0, 0.3, size=dims[i + 1]) for i in range(len(dims) - 1)]
    X = rng.normal(size=(M, dims[0]))
    Yt = rng.normal(size=(M, dims[-1]))
    return X, Ws, bs, Yt


def _torch_grads(X, Ws, bs, Yt, nv):
    tW = [torch.tensor(w, dtype=torch.float64, requires_grad=True)
          for w in Ws]
    tb = [torch.tensor(b, dtype=torch.float64, requires_grad=True)
          for b in bs]
    h = torch.tensor(X, dtype=torch.float64)
    for l in range(len(Ws)):
        z = h @ tW[l].T + tb[l]
        h = torch.nn.functional.gelu(z, approximate='tanh') \
            if l < len(Ws) - 1 else z
    yt = torch.tensor(Yt, dtype=torch.float64)
    loss = ((h[:, :nv] - yt[:, :nv]) ** 2).mean()
    loss.backward()
    return ([w.grad.numpy() for w in tW], [b.grad.numpy() for b in tb],
            h.detach().numpy())


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _check_against_autograd(fb, seed=1, nv=4):
    X, Ws, bs, Yt = _net(seed)
    yhat, dW, db = fb(X, Ws, bs, Yt, nv, emulate=False)
    gW, gb, ty = _torch_grads(X, Ws, bs, Yt, nv)
    errs = [_rel(yhat, ty)]
    errs += [_rel(a, b) for a, b in zip(dW, gW)]
    errs += [_rel(a, b) for a, b in zip(db, gb)]
    return max(errs)


def test_gelu_derivative_central_differences():
    z = np.linspace(-6.0, 6.0, 241)
    h = 1e-6
    g_p, _ = mlp_ref.gelu(z + h)
    g_m, _ = mlp_ref.gelu(z - h)
    _g, d = mlp_ref.gelu(z)
    assert np.abs((g_p - g_m) / (2 * h) - d).max() < 1e-8


def test_gelu_matches_torch_tanh_form():
    z = np.linspace(-8.0, 8.0, 321)
    g, _ = mlp_ref.gelu(z)
    t = torch.nn.functional.gelu(torch.tensor(z), approximate='tanh').numpy()
    assert np.abs(g - t).max() < 1e-15


@pytest.mark.parametrize('seed', [1, 2, 3])
def test_forward_backward_matches_autograd(seed):
    assert _check_against_autograd(mlp_ref.forward_backward, seed) < 1e-12


def test_forward_backward_central_differences():
    X, Ws, bs, Yt = _net(11, dims=(5, 8, 8, 6, 4), M=5)
    nv = 3
    _y, dW, db = mlp_ref.forward_backward(X, Ws, bs, Yt, nv, emulate=False)
    h = 1e-6
    rng = np.random.default_rng(0)
    for l in range(len(Ws)):
        for _ in range(6):
            i = rng.integers(Ws[l].shape[0])
            j = rng.integers(Ws[l].shape[1])
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[l][i, j] += h
            Wm[l][i, j] -= h
            fd = (mlp_ref.loss(X, Wp, bs, Yt, nv) -
                  mlp_ref.loss(X, Wm, bs, Yt, nv)) / (2 * h)
            assert abs(fd - dW[l][i, j]) <= 1e-7 * max(1.0, abs(fd)), \
                (l, i, j, fd, dW[l][i, j])
        for i in range(bs[l].shape[0]):
            bp = [b.copy() for b in bs]
            bm = [b.copy() for b in bs]
            bp[l][i] += h
            bm[l][i] -= h
            fd = (mlp_ref.loss(X, Ws, bp, Yt, nv) -
                  mlp_ref.loss(X, Ws, bm, Yt, nv)) / (2 * h)
            assert abs(fd - db[l][i]) <= 1e-7 * max(1.0, abs(fd)), \
                (l, i, fd, db[l][i])


def test_emulated_roundings_stay_close_to_exact():
    """The BF16-emulating gradient is the exact one up to BF16 noise."""
    X, Ws, bs, Yt = _net(5, dims=(16, 32, 32, 32, 8), M=64)
    X = mlp_ref.bf16(X)
    Ws = [mlp_ref.bf16(w) for w in Ws]
    _y0, dW0, db0 = mlp_ref.forward_backward(X, Ws, bs, Yt, 8, emulate=False)
    _y1, dW1, db1 = mlp_ref.forward_backward(X, Ws, bs, Yt, 8)
    for a, b in zip(dW1 + db1, dW0 + db0):
        assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b)


def _planted(monkeypatch, what):
    if what == 'gelu_prime_dropped':
        orig = mlp_ref.gelu

        def g(z):
            a, _d = orig(z)
            return a, np.ones_like(z)
        monkeypatch.setattr(mlp_ref, 'gelu', g)
    elif what == 'gelu_constant':
        monkeypatch.setattr(mlp_ref, '_GELU_A', 0.0447)
    elif what == 'bias_sum_dropped':
        monkeypatch.setattr(mlp_ref, 'bias_grad',
                            lambda dZ: np.zeros(dZ.shape[1]))
    elif what == 'bias_mean':
        monkeypatch.setattr(mlp_ref, 'bias_grad', lambda dZ: dZ.mean(axis=0))
    elif what == 'mse_factor':
        orig = mlp_ref.mse_grad
        monkeypatch.setattr(mlp_ref, 'mse_grad',
                            lambda y, t, nv: 0.5 * orig(y, t, nv))


@pytest.mark.parametrize('what', ['gelu_prime_dropped', 'gelu_constant',
                                  'bias_sum_dropped', 'bias_mean',
                                  'mse_factor'])
def test_planted_mistakes_are_caught(monkeypatch, what):
    _planted(monkeypatch, what)
    assert _check_against_autograd(mlp_ref.forward_backward) > 1e-6
