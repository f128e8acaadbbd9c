"""Learned wdot surrogate with exact fallback and audit.

An MLP maps per-state features [ln T, ln P, ln Y_k (k < NSP)] (normalised)
through three GELU hidden layers of 256 to y_k = arcsinh(wdot_k / s_k), with
s_k a per-species scale (the median |wdot_k| of the training set).  It is
trained on states labelled by the exact FP64 kernel, entirely on the GPU by
the hand-written kernels of libcrs_mlp.so (BF16 MFMA products with FP32
accumulation, FP32-master Adam).

Inference: features -> MLP -> decode to FP64 wdot -> gate.  The gate sends a
state to the exact kernel (crs_eval_wdot_subset of the mechanism library)
when it lies outside the training box in T or P, has an invalid composition
or a non-finite prediction, or when its predicted rates violate element or
mass conservation by more than a relative ``rmax``; ``rmax`` is calibrated
on held-out states so that admitted states stay within a target error.  A
random fraction of all states (1 % by default, fixed seed) is re-evaluated
exactly and scored, reporting the error users actually get.

This is torch-free except for device buffers; every kernel is ours.
"""

import ctypes
import hashlib
import json
import math
import os

import numpy as np
import torch

from . import metrics

__all__ = ['MlpLib', 'Surrogate', 'load_mlp', 'mechanism_hash']

HERE = os.path.dirname(os.path.abspath(__file__))
MLP_LIB = os.path.join(HERE, '_build', 'libcrs_mlp.so')

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I = ctypes.c_int
_F = ctypes.c_float


class GateParams(ctypes.Structure):
    _fields_ = [('t_lo', ctypes.c_double), ('t_hi', ctypes.c_double),
                ('p_lo', ctypes.c_double), ('p_hi', ctypes.c_double),
                ('rmax', ctypes.c_double)]


class NetDesc(ctypes.Structure):
    """crs_mlp_net of include/crs_mlp.h."""
    _fields_ = [('nsp', ctypes.c_int), ('kin', ctypes.c_int),
                ('hidden', ctypes.c_int), ('layers', ctypes.c_int),
                ('kout', ctypes.c_int), ('tiles', ctypes.c_void_p),
                ('bias', ctypes.c_void_p), ('mu', ctypes.c_void_p),
                ('sd', ctypes.c_void_p), ('scale', ctypes.c_void_p),
                ('ymu', ctypes.c_void_p), ('ysd', ctypes.c_void_p),
                ('gate', GateParams), ('elem', ctypes.c_void_p),
                ('n_elem', ctypes.c_int), ('mw', ctypes.c_void_p)]


class MlpError(RuntimeError):
    pass


def _p(t):
    return _P(t.data_ptr()) if t is not None else None


def _s():
    return _P(torch.cuda.current_stream().cuda_stream)


class MlpLib(object):
    """ctypes binding of libcrs_mlp.so (include/crs_mlp.h)."""

    SIGS = {
        'crs_mlp_linear_fwd': [_P, _I64, _P, _I64, _P, _I64, _I, _I, _I, _P,
                               _P, _P, _I64, _I, _P, _I64, _P],
        'crs_mlp_linear_bwd_data': [_P, _I64, _P, _I64, _I64, _I, _I, _P, _P,
                                    _I64, _P, _I64, _P],
        'crs_mlp_linear_bwd_weight': [_P, _I64, _P, _I64, _I64, _I, _I, _P,
                                      _P, _P],
        'crs_mlp_features': [_P, _P, _P, _P, _I64, _I64, _I, _P, _P, _P, _I,
                             _P],
        'crs_mlp_targets': [_P, _I64, _I64, _I, _P, _P, _P, _P, _I, _P],
        'crs_mlp_decode': [_P, _I, _P, _I64, _I64, _I, _P, _P, _P, _P, _P],
        'crs_mlp_mse_grad': [_P, _P, _I64, _I, _I, _I, _P, _I, _P, _P],
        'crs_mlp_adam': [_P, _P, _P, _P, _P, _I64, _F, _F, _F, _F, _I, _F,
                         _P],
        'crs_mlp_to_bf16': [_P, _P, _I64, _P],
        'crs_mlp_transpose_bf16': [_P, _I64, _I64, _I, _P, _I64, _P],
        'crs_mlp_gate': [_P, _P, _P, _P, _I64, _I64, _I, GateParams, _P, _I,
                         _P, _P, _P, _P, _P],
        'crs_mlp_pack_tiles': [_P, _I, _I, _I, _P, _P],
        'crs_mlp_infer': [_P, _P, _P, _I64, _I64, _I, _I, _I, _I, _I, _P, _P,
                          _P, _P, _P, _P, _P, _P, GateParams, _P, _I, _P, _P,
                          _P, _P, _P],
        'crs_mlp_eval_gated': [ctypes.POINTER(NetDesc), _P, _P, _P, _P, _I64,
                               _I64, _P, _I, _P, _P, _P, _P],
    }

    def __init__(self, path=MLP_LIB):
        if not os.path.exists(path):
            raise FileNotFoundError('surrogate library not built: ' + path)
        self.lib = ctypes.CDLL(os.path.abspath(path))
        self.lib.crs_mlp_last_error.restype = ctypes.c_char_p
        for name, args in self.SIGS.items():
            fn = getattr(self.lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int

    def __call__(self, name, *args):
        rc = getattr(self.lib, name)(*args)
        if rc != 0:
            raise MlpError('{}: {}'.format(
                name, self.lib.crs_mlp_last_error().decode()))


_LIB = None


def load_mlp():
    global _LIB
    if _LIB is None:
        _LIB = MlpLib()
    return _LIB


def _pad(n, m):
    return (n + m - 1) // m * m


def gate_threshold(err, resid, target):
    """(rmax, admitted fraction) of the residual gate for one error target.

    States are ranked by their balance residual ``resid``; the admitted set
    is the longest low-residual prefix whose worst error ``err`` stays
    within ``target``, and rmax is the residual of its last state (a state
    passes the gate when its residual is <= rmax, so a prefix never ends
    inside a run of equal residuals).  Non-finite errors or residuals are
    never admitted.  (-1.0, 0.0) when no prefix qualifies: a negative rmax
    marks a gate that admits no state, and the gated evaluation then skips
    the network altogether.
    """
    e = torch.where(torch.isfinite(err), err, torch.full_like(err, 1e300))
    r = torch.where(torch.isfinite(resid), resid,
                    torch.full_like(resid, 1e300))
    order = torch.argsort(r, stable=True)
    rs = r[order]
    worst = torch.cummax(e[order], dim=0).values
    # a prefix may only end where the residual steps up: states tied with
    # its last one would pass the gate too
    ends = torch.ones_like(worst, dtype=torch.bool)
    ends[:-1] = rs[1:] > rs[:-1]
    ok = ((worst <= target) & ends & (rs < 1e300)).nonzero()
    if ok.numel() == 0:
        return -1.0, 0.0
    last = int(ok.max().item())
    return float(rs[last].item()), (last + 1) / err.numel()


def mechanism_hash(specs, reacs):
    """SHA-256 of a mechanism's tables, in the library's species order.

    Covers what the rates depend on: species names, molecular weights,
    NASA-7 ranges and coefficients, and every reaction field of ReacInfo
    (stoichiometry, Arrhenius, reversibility, third-body efficiencies,
    falloff, Troe/SRI, PLOG, Chebyshev).  Floats enter by their exact
    IEEE-754 bits (float.hex), so two mechanisms differing in any constant
    hash differently.
    """
    def canon(v):
        if isinstance(v, np.ndarray):
            return [canon(x) for x in v.tolist()]
        if isinstance(v, (list, tuple)):
            return [canon(x) for x in v]
        if isinstance(v, (bool, np.bool_)):
            return bool(v)
        if isinstance(v, (int, np.integer)):
            return int(v)
        if isinstance(v, (float, np.floating)):
            return float(v).hex()
        return str(v)
    doc = dict(
        species=[[sp.name, canon(sp.mw), canon(sp.Trange), canon(sp.lo),
                  canon(sp.hi)] for sp in specs],
        reactions=[{k: canon(v) for k, v in sorted(vars(r).items())}
                   for r in reacs])
    text = json.dumps(doc, sort_keys=True, separators=(',', ':'))
    return hashlib.sha256(text.encode()).hexdigest()


class Surrogate(object):
    """MLP [ln T, ln P, ln Y] -> arcsinh(wdot / s) for one mechanism."""

    def __init__(self, mech_lib, specs, reacs, hidden=256, layers=3, seed=0):
        self.mlp = load_mlp()
        self.mech = mech_lib
        self.specs, self.reacs = specs, reacs
        self.nsp = len(specs)
        self.nin = self.nsp + 2
        self.kin = _pad(self.nin, 64)
        self.hid = hidden
        self.nout = self.nsp
        self.kout = _pad(self.nout, 64)
        dims = [self.kin] + [hidden] * layers + [self.kout]
        self.shapes = [(dims[i + 1], dims[i]) for i in range(len(dims) - 1)]
        self.offsets = []
        off = 0
        for n_out, n_in in self.shapes:
            self.offsets.append((off, off + n_out * n_in))
            off += n_out * n_in + n_out
        self.nparam = off
        dev = 'cuda'
        self.p = torch.zeros(off, dtype=torch.float32, device=dev)
        self.g = torch.zeros_like(self.p)
        self.m1 = torch.zeros_like(self.p)
        self.m2 = torch.zeros_like(self.p)
        self.pb = torch.zeros(off, dtype=torch.bfloat16, device=dev)
        self.step = 0
        rng = np.random.default_rng(seed)
        host = np.zeros(off, dtype=np.float32)
        for li, (n_out, n_in) in enumerate(self.shapes):
            w0, b0 = self.offsets[li]
            valid_in = self.nin if li == 0 else n_in
            valid_out = self.nout if li == len(self.shapes) - 1 else n_out
            w = rng.normal(0.0, math.sqrt(2.0 / valid_in),
                           size=(n_out, n_in)).astype(np.float32)
            w[valid_out:, :] = 0.0
            w[:, valid_in:] = 0.0
            host[w0:w0 + n_out * n_in] = w.ravel()
        self.p.copy_(torch.from_numpy(host))
        self.mlp('crs_mlp_to_bf16', _p(self.p), _p(self.pb), self.nparam, _s())
        # normalisation / scales / gate (set by fit)
        self.mu = torch.zeros(self.nin, dtype=torch.float32, device=dev)
        self.sd = torch.ones(self.nin, dtype=torch.float32, device=dev)
        self.scale = torch.ones(self.nsp, dtype=torch.float64, device=dev)
        self.ymu = torch.zeros(self.nsp, dtype=torch.float32, device=dev)
        self.ysd = torch.ones(self.nsp, dtype=torch.float32, device=dev)
        W = np.array([sp.mw for sp in specs])
        elems = sorted({e.upper() for sp in specs for e, _ in sp.elem})
        a = np.zeros((len(elems), self.nsp))
        for k, sp in enumerate(specs):
            for e, c in sp.elem:
                a[elems.index(e.upper()), k] += c
        self.elem = torch.tensor(a, dtype=torch.float64, device=dev)
        self.mw = torch.tensor(W, dtype=torch.float64, device=dev)
        self.gate = GateParams(0.0, 0.0, 0.0, 0.0, 0.0)
        self._bufs = {}

    # -- parameter views --------------------------------------------------
    def w(self, li, bf16=True):
        w0, b0 = self.offsets[li]
        n_out, n_in = self.shapes[li]
        src = self.pb if bf16 else self.p
        return src[w0:b0].view(n_out, n_in)

    def b(self, li):
        w0, b0 = self.offsets[li]
        return self.p[b0:b0 + self.shapes[li][0]]

    def gw(self, li):
        w0, b0 = self.offsets[li]
        return self.g[w0:b0]

    def gb(self, li):
        w0, b0 = self.offsets[li]
        return self.g[b0:b0 + self.shapes[li][0]]

    def _buf(self, name, shape, dtype):
        t = self._bufs.get(name)
        if t is None or t.shape[0] < shape[0]:
            t = torch.empty(shape, dtype=dtype, device='cuda')
            self._bufs[name] = t
        return t[:shape[0]]

    # -- transforms -------------------------------------------------------
    def features(self, T, P, Y, idx=None, out=None):
        n = idx.numel() if idx is not None else T.numel()
        X = out if out is not None else self._buf('X', (n, self.kin),
                                                  torch.bfloat16)
        self.mlp('crs_mlp_features', _p(T), _p(P), _p(Y), _p(idx), n,
                 Y.stride(0), self.nsp, _p(self.mu), _p(self.sd), _p(X),
                 self.kin, _s())
        return X

    def targets(self, W):
        n = W.shape[1]
        Yt = torch.empty((n, self.nsp), dtype=torch.float32, device='cuda')
        self.mlp('crs_mlp_targets', _p(W), n, W.stride(0), self.nsp,
                 _p(self.scale), _p(self.ymu), _p(self.ysd), _p(Yt), self.nsp,
                 _s())
        return Yt

    # -- network ------------------------------------------------------------
    def forward(self, X, save=False):
        """X [M][kin] bf16 -> yhat [M][kout] fp32 (columns < nsp valid).

        With save (training) the pre-activations Z and the hidden
        activations are kept, the latter also transposed (states
        contiguous, zero padded to a multiple of 32 states) for the weight
        gradients."""
        M = X.shape[0]
        Mr = _pad(M, 32)
        h = X
        zs, hs, hts = [], [X], [None]
        nl = len(self.shapes)
        for li, (n_out, n_in) in enumerate(self.shapes):
            if li < nl - 1:
                A = self._buf('A%d' % li, (M, n_out), torch.bfloat16)
                Z = AT = None
                if save:
                    Z = self._buf('Z%d' % li, (M, n_out), torch.bfloat16)
                    AT = self._buf('AT%d' % li, (n_out * Mr,),
                                   torch.bfloat16)[:n_out * Mr]
                self.mlp('crs_mlp_linear_fwd', _p(h), h.stride(0),
                         _p(self.w(li)), n_in, _p(self.b(li)), M, n_out, n_in,
                         1, _p(A), _p(Z), None, n_out, n_out, _p(AT), Mr,
                         _s())
                zs.append(Z)
                hs.append(A)
                hts.append(AT)
                h = A
            else:
                Yh = self._buf('Yh', (M, n_out), torch.float32)
                self.mlp('crs_mlp_linear_fwd', _p(h), h.stride(0),
                         _p(self.w(li)), n_in, _p(self.b(li)), M, n_out, n_in,
                         0, None, None, _p(Yh), n_out, self.nout, None, 0,
                         _s())
        if save:
            self._saved = (zs, hs, hts)
        return Yh

    def backward(self, Yh, Yt, loss=None):
        """Gradients of the mean squared arcsinh error into self.g.

        The weight gradient of each layer reads dZ^T and X^T (states
        contiguous): the forward pass and the data gradients write them
        transposed as they go; the input features and the output layer's
        dZ (64 wide) are transposed here."""
        zs, hs, hts = self._saved
        M = Yh.shape[0]
        Mr = _pad(M, 32)
        nl = len(self.shapes)
        dZ = self._buf('dZ%d' % (nl - 1), (M, self.kout), torch.bfloat16)
        self.mlp('crs_mlp_mse_grad', _p(Yh), _p(Yt), M, self.nout, self.kout,
                 Yt.stride(0), _p(dZ), self.kout, _p(loss), _s())
        dZT = self._buf('dZT_out', (self.kout * Mr,),
                        torch.bfloat16)[:self.kout * Mr]
        self.mlp('crs_mlp_transpose_bf16', _p(dZ), dZ.stride(0), M, self.kout,
                 _p(dZT), Mr, _s())
        XT0 = self._buf('XT_in', (self.kin * Mr,),
                        torch.bfloat16)[:self.kin * Mr]
        self.mlp('crs_mlp_transpose_bf16', _p(hs[0]), hs[0].stride(0), M,
                 self.kin, _p(XT0), Mr, _s())
        self.g.zero_()
        for li in range(nl - 1, -1, -1):
            n_out, n_in = self.shapes[li]
            XT = hts[li] if li > 0 else XT0
            self.mlp('crs_mlp_linear_bwd_weight', _p(dZT), Mr, _p(XT), Mr, M,
                     n_out, n_in, _p(self.gw(li)), _p(self.gb(li)), _s())
            if li > 0:
                dprev = self._buf('dZ%d' % (li - 1), (M, n_in), torch.bfloat16)
                dprevT = self._buf('dZT%d' % (li - 1), (n_in * Mr,),
                                   torch.bfloat16)[:n_in * Mr]
                WT = self._buf('WT%d' % li, (n_in * n_out,),
                               torch.bfloat16)
                self.mlp('crs_mlp_transpose_bf16', _p(self.w(li)), n_in,
                         n_out, n_in, _p(WT), n_out, _s())
                self.mlp('crs_mlp_linear_bwd_data', _p(dZ), dZ.stride(0),
                         _p(WT), n_out, M, n_out, n_in, _p(zs[li - 1]),
                         _p(dprev), n_in, _p(dprevT), Mr, _s())
                dZ, dZT = dprev, dprevT

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        self.step += 1
        self.mlp('crs_mlp_adam', _p(self.p), _p(self.g), _p(self.m1),
                 _p(self.m2), _p(self.pb), self.nparam, lr, b1, b2, eps,
                 self.step, wd, _s())

    # -- training -----------------------------------------------------------
    def fit_normalisation(self, T, P, Y, W):
        """Feature mean/std, per-species scales s_k and the training box."""
        with torch.no_grad():
            f = torch.empty((self.nin, T.numel()), dtype=torch.float64,
                            device='cuda')
            f[0] = T.log()
            f[1] = P.log()
            ys = Y[:self.nsp - 1]
            f[2:self.nsp + 1] = ys.clamp_min(1e-30).log()
            f[self.nsp + 1] = (1.0 - ys.sum(0)).clamp_min(1e-30).log()
            self.mu.copy_(f.mean(1).float())
            self.sd.copy_(f.std(1).clamp_min(1e-6).float())
            med = W.abs().median(dim=1).values
            self.scale.copy_(torch.where(med > 0, med, torch.ones_like(med)))
            ya = torch.asinh(W / self.scale[:, None])
            self.ymu.copy_(ya.mean(1).float())
            self.ysd.copy_(ya.std(1).clamp_min(1e-3).float())
            del ya
            self.gate.t_lo = float(T.min())
            self.gate.t_hi = float(T.max())
            self.gate.p_lo = float(P.min())
            self.gate.p_hi = float(P.max())

    def mse(self, T, P, Y, W, chunk=1 << 18):
        """Mean squared error of the network in its standardised output
        space over states labelled with exact rates W (the training loss)."""
        n = T.numel()
        tot = 0.0
        for a in range(0, n, chunk):
            b = min(n, a + chunk)
            Tc, Pc = T[a:b].contiguous(), P[a:b].contiguous()
            X = self.features(Tc, Pc, Y[:, a:b].contiguous())
            Yh = self.forward(X)[:, :self.nout]
            Yt = self.targets(W[:, a:b].contiguous())
            tot += ((Yh - Yt) ** 2).sum().item()
        return tot / (n * self.nout)

    def grad_step(self, X, Yt, sel, loss):
        """Forward and backward over the batch X[sel], Yt[sel]: gradients
        into self.g, the batch's squared error added to loss.  Reads and
        writes only buffers of fixed address, so a captured graph of it can
        be replayed with new indices in sel."""
        xb = self._buf('xb', (sel.numel(), self.kin), torch.bfloat16)
        yb = self._buf('yb', (sel.numel(), self.nout), torch.float32)
        torch.index_select(X, 0, sel, out=xb)
        torch.index_select(Yt, 0, sel, out=yb)
        Yh = self.forward(xb, save=True)
        self.backward(Yh, yb, loss)

    def train(self, T, P, Y, W, epochs=10, batch=16384, lr=2e-3,
              lr_final=1e-4, seed=0, log=None, val=None, graph=True):
        """Train on states (T, P, Y) labelled with exact rates W [NSP][n].

        Cosine learning-rate schedule from lr to lr_final.  With graph the
        forward and backward pass of a step is captured once as a HIP graph
        and replayed (the batch's indices are copied into its fixed index
        buffer first; Adam runs eagerly, its step size changing every
        step).  Returns the training MSE per epoch; with val = (T, P, Y, W)
        the validation MSE per epoch is kept in self.val_history.
        """
        n = T.numel()
        self.val_history = []
        self.fit_normalisation(T, P, Y, W)
        X = torch.empty((n, self.kin), dtype=torch.bfloat16, device='cuda')
        self.features(T, P, Y, out=X)
        Yt = self.targets(W)
        gen = torch.Generator(device='cuda')
        gen.manual_seed(seed)
        loss = torch.zeros(1, dtype=torch.float32, device='cuda')
        steps_per_epoch = max(1, n // batch)
        batch = min(batch, n)
        total = epochs * steps_per_epoch
        sel = torch.zeros(batch, dtype=torch.int64, device='cuda')
        g = None
        if graph and epochs:
            # one eager step sizes every buffer, then the capture
            self.grad_step(X, Yt, sel, loss)
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self.grad_step(X, Yt, sel, loss)
        history = []
        k = 0
        for ep in range(epochs):
            perm = torch.randperm(n, device='cuda', generator=gen)
            loss.zero_()
            for s in range(steps_per_epoch):
                sel.copy_(perm[s * batch:(s + 1) * batch])
                if g is not None:
                    g.replay()
                else:
                    self.grad_step(X, Yt, sel, loss)
                frac = k / max(total - 1, 1)
                cur = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(
                    math.pi * frac))
                self.adam(lr=cur)
                k += 1
            mse = loss.item() / (steps_per_epoch * batch * self.nout)
            history.append(mse)
            if val is not None:
                self.val_history.append(self.mse(*val))
            if log:
                log('epoch {} mse {:.4e}{}'.format(
                    ep, mse, ' val {:.4e}'.format(self.val_history[-1])
                    if val is not None else ''))
        return history

    # -- inference ----------------------------------------------------------
    @property
    def fused_ok(self):
        """The fused kernel covers this network's shape."""
        return (self.kin, self.hid, len(self.shapes) - 1, self.kout) == \
            (64, 256, 3, 64)

    def pack(self):
        """Weights in the fused kernel's MFMA tile order, biases in one
        array; redone whenever the parameters changed since the last pack."""
        if getattr(self, '_packed_step', None) == self.step and \
                getattr(self, '_packed_id', None) == id(self.pb):
            return self._packed
        tiles = self._buf('tiles', (sum(a * b for a, b in self.shapes),),
                          torch.bfloat16)
        off = 0
        for li, (n_out, n_in) in enumerate(self.shapes):
            cnt = n_out * n_in
            self.mlp('crs_mlp_pack_tiles', _p(self.w(li)), n_out, n_in,
                     1 if li > 0 else 0, _p(tiles[off:off + cnt]), _s())
            off += cnt
        bias = torch.cat([self.b(li) for li in range(len(self.shapes))])
        self._packed = (tiles, bias.contiguous())
        self._packed_step, self._packed_id = self.step, id(self.pb)
        return self._packed

    def infer(self, T, P, Y, W):
        """Fused surrogate + gate over every state: W [NSP][ld] gets the
        surrogate rates; returns (flags, fallback indices, count tensor)."""
        n = T.numel()
        tiles, bias = self.pack()
        flags = self._buf('flags', (n,), torch.int32)
        idx = self._buf('fb_idx', (n,), torch.int32)
        cnt = torch.zeros(1, dtype=torch.int32, device='cuda')
        self.mlp('crs_mlp_infer', _p(T), _p(P), _p(Y), n, Y.stride(0),
                 self.nsp, self.kin, self.hid, len(self.shapes) - 1,
                 self.kout, _p(tiles), _p(bias), _p(self.mu), _p(self.sd),
                 _p(self.scale), _p(self.ymu), _p(self.ysd), _p(W), self.gate,
                 _p(self.elem), self.elem.shape[0], _p(self.mw), _p(flags),
                 _p(idx), _p(cnt), _s())
        return flags, idx, cnt

    def predict(self, T, P, Y, W, fused=True):
        """Surrogate wdot for every state into W [NSP][ld] (FP64)."""
        if fused and self.fused_ok:
            self.infer(T, P, Y, W)
            return W
        X = self.features(T, P, Y)
        Yh = self.forward(X)
        n = T.numel()
        self.mlp('crs_mlp_decode', _p(Yh), self.kout, None, n, W.stride(0),
                 self.nsp, _p(self.scale), _p(self.ymu), _p(self.ysd), _p(W),
                 _s())
        return W

    def net_desc(self):
        """The network as the C ABI's crs_mlp_net (weights packed first);
        the tensors it points to stay alive as attributes of self."""
        tiles, bias = self.pack()
        return NetDesc(self.nsp, self.kin, self.hid, len(self.shapes) - 1,
                       self.kout, _p(tiles), _p(bias), _p(self.mu),
                       _p(self.sd), _p(self.scale), _p(self.ymu),
                       _p(self.ysd), self.gate, _p(self.elem),
                       self.elem.shape[0], _p(self.mw))

    def exact_fn(self):
        """Address of the mechanism library's crs_eval_wdot_subset_dev (the
        crs_subset_fn the gated entry points call)."""
        return ctypes.cast(self.mech.lib.crs_eval_wdot_subset_dev,
                           ctypes.c_void_p)

    def gate_states(self, T, P, Y, W):
        n = T.numel()
        flags = self._buf('flags', (n,), torch.int32)
        idx = self._buf('fb_idx', (n,), torch.int32)
        cnt = torch.zeros(1, dtype=torch.int32, device='cuda')
        self.mlp('crs_mlp_gate', _p(T), _p(P), _p(Y), _p(W), n, W.stride(0),
                 self.nsp, self.gate, _p(self.elem), self.elem.shape[0],
                 _p(self.mw), _p(flags), _p(idx), _p(cnt), _s())
        return flags, idx, cnt

    def evaluate(self, T, P, Y, W, mode='fp64', fused=True, read_count=True):
        """Gated surrogate: predict, gate, exact fallback.

        With the fused network shape this is one call of the C ABI's
        crs_mlp_eval_gated (inference, gate and the mechanism library's
        exact subset kernel on the current stream, the fallback count kept
        on the device).  Returns the fallback count m, read back after the
        work is enqueued, or with read_count=False the device count tensor
        (no synchronisation at all).
        """
        from .runtime import MODES
        n = T.numel()
        if fused and self.fused_ok:
            flags = self._buf('flags', (n,), torch.int32)
            idx = self._buf('fb_idx', (n,), torch.int32)
            cnt = self._buf('fb_count', (1,), torch.int32)
            desc = self.net_desc()
            self.mlp('crs_mlp_eval_gated', ctypes.byref(desc), _p(T), _p(P),
                     _p(Y), _p(W), n, Y.stride(0), self.exact_fn(),
                     MODES[mode], _p(flags), _p(idx), _p(cnt), _s())
            return int(cnt.item()) if read_count else cnt
        if not self.gate.rmax >= 0.0:
            # the gate admits no state: exact rates, no network
            self.mech.eval_wdot(T, P, Y, mode, out=W)
            cnt = torch.full((1,), n, dtype=torch.int32, device=W.device)
            return n if read_count else cnt
        self.predict(T, P, Y, W, fused=False)
        flags, idx, cnt = self.gate_states(T, P, Y, W)
        m = int(cnt.item())
        if m:
            self.mech.eval_wdot_subset(T, P, Y, W, idx[:m], mode)
        return m if read_count else cnt

    def gross_scale(self, T, P, Y):
        """S_k of the states, from the mechanism's FP64 stage kernels."""
        C = self.mech.eval_conc(T, P, Y)['C']
        fwd, rev = self.mech.eval_rxn_rates(T, P, C)
        pm = self.mech.get_rxn_pres_mod(T, P, C)
        return metrics.gross_scale(self.specs, self.reacs, fwd, rev, pm)

    def audit(self, T, P, Y, W, frac=0.01, seed=1234, tau=5e-2):
        """Exact re-evaluation of a random ``frac`` of the states.

        Returns the conditioned relative error (create_rate_subs_amd.metrics)
        of the delivered rates W on those states: what a user gets.
        """
        n = T.numel()
        m = max(1, int(round(frac * n)))
        gen = torch.Generator(device='cuda')
        gen.manual_seed(seed)
        col = torch.randperm(n, device='cuda', generator=gen)[:m].sort().values
        Ts, Ps = T[col].contiguous(), P[col].contiguous()
        Ys = Y[:, col].contiguous()
        w_ex = self.mech.eval_wdot(Ts, Ps, Ys, 'fp64')
        S = self.gross_scale(Ts, Ps, Ys)
        e = metrics.rel_err(W[:, col], w_ex, S, tau)
        per_state = e.max(dim=0).values
        return dict(states=m, max=e.max().item(), median=e.median().item(),
                    state_p99=per_state.quantile(0.99).item()
                    if m > 1 else per_state.max().item())

    def disagreement(self, other, T, P, Y):
        """Per-state normwise disagreement of two networks' rates,
        max_k |w_k - w'_k| / max(max_k |w_k|, max_k |w'_k|): an error
        estimate for gating (two members of an ensemble trained from
        different initialisations disagree most where both are poor)."""
        Wa = torch.empty((self.nsp, T.numel()), dtype=torch.float64,
                         device='cuda')
        Wb = torch.empty_like(Wa)
        self.predict(T, P, Y, Wa)
        other.predict(T, P, Y, Wb)
        d = (Wa - Wb).abs().amax(dim=0)
        m = torch.maximum(Wa.abs().amax(dim=0), Wb.abs().amax(dim=0))
        out = d / m
        return torch.where(d == 0, torch.zeros_like(out), out)

    def residuals(self, W):
        """Per-state max relative element/mass balance residual of W."""
        a = torch.cat([self.elem, self.mw[None, :]], dim=0)     # [E+1][NSP]
        r = (a @ W).abs() / (a @ W.abs()).clamp_min(1e-300)
        return r.max(dim=0).values

    def calibrate_gate(self, T, P, Y, W_exact, S, target, tau=5e-2,
                       curve=(1e-2, 3e-2, 1e-1, 3e-1, 1.0),
                       metric='conditioned'):
        """Largest rmax whose admitted states all meet ``target`` error.

        States are ranked by their balance residual; the admitted set is the
        longest low-residual prefix whose worst error is within the target:
        the conditioned error (max over species, create_rate_subs_amd.metrics)
        or, with metric='normwise', the per-state normwise error.  Sets
        gate.rmax and returns (rmax, admitted fraction, curve) where curve
        lists the admitted fraction for each error target in ``curve``.
        """
        W = torch.empty_like(W_exact)
        self.predict(T, P, Y, W)
        if metric == 'normwise':
            e = metrics.normwise_err(W, W_exact)
        else:
            e = metrics.rel_err(W, W_exact, S, tau).max(dim=0).values
        r = self.residuals(W)
        table = []
        for t in curve:
            rm, fr = gate_threshold(e, r, t)
            table.append(dict(target=t, rmax=rm, admitted=fr))
        rmax, frac = gate_threshold(e, r, target)
        self.gate.rmax = rmax
        return rmax, frac, table

    # -- checkpoint ---------------------------------------------------------
    def _array_specs(self):
        """Arrays a checkpoint holds: name -> (tensor, dtype, shape)."""
        arrays = dict(p=self.p, m1=self.m1, m2=self.m2, mu=self.mu,
                      sd=self.sd, scale=self.scale, ymu=self.ymu,
                      ysd=self.ysd)
        return {k: (t, str(t.dtype).replace('torch.', ''), list(t.shape))
                for k, t in arrays.items()}

    def save(self, directory):
        os.makedirs(directory, exist_ok=True)
        manifest = dict(nsp=self.nsp, hidden=self.hid,
                        shapes=self.shapes, step=self.step,
                        mechanism=mechanism_hash(self.specs, self.reacs),
                        gate=dict(t_lo=self.gate.t_lo, t_hi=self.gate.t_hi,
                                  p_lo=self.gate.p_lo, p_hi=self.gate.p_hi,
                                  rmax=self.gate.rmax), files={})
        for k, (t, dtype, shape) in self._array_specs().items():
            a = t.detach().cpu().numpy()
            fn = k + '.bin'
            a.astype(a.dtype.newbyteorder('<')).tofile(
                os.path.join(directory, fn))
            manifest['files'][k] = dict(file=fn, dtype=dtype, shape=shape)
        with open(os.path.join(directory, 'manifest.json'), 'w') as fh:
            json.dump(manifest, fh, indent=1)

    def load(self, directory):
        """Restore a checkpoint written by save() for this mechanism.

        Refuses (ValueError) a checkpoint of another mechanism (hash of the
        species, thermo and reaction tables), of another network shape, or
        whose manifest names arrays other than the expected ones or with
        another dtype or shape; nothing is modified then.
        """
        with open(os.path.join(directory, 'manifest.json')) as fh:
            man = json.load(fh)
        if man.get('mechanism') != mechanism_hash(self.specs, self.reacs):
            raise ValueError('checkpoint was trained on another mechanism')
        if man['nsp'] != self.nsp or [tuple(s) for s in man['shapes']] != \
                [tuple(s) for s in self.shapes]:
            raise ValueError('checkpoint does not match this network')
        expect = self._array_specs()
        files = man.get('files', {})
        if set(files) != set(expect):
            raise ValueError('checkpoint arrays {} differ from {}'.format(
                sorted(files), sorted(expect)))
        loaded = {}
        for k, (t, dtype, shape) in expect.items():
            spec = files[k]
            if spec.get('dtype') != dtype or list(spec.get('shape', ())) != \
                    shape or os.path.basename(spec.get('file', '')) != \
                    spec.get('file'):
                raise ValueError('checkpoint array {} has {} {} (expected '
                                 '{} {})'.format(k, spec.get('dtype'),
                                                 spec.get('shape'), dtype,
                                                 shape))
            a = np.fromfile(os.path.join(directory, spec['file']),
                            dtype=np.dtype(dtype).newbyteorder('<'))
            if a.size != int(np.prod(shape)):
                raise ValueError('checkpoint array {} holds {} values, '
                                 'expected {}'.format(k, a.size,
                                                      int(np.prod(shape))))
            loaded[k] = a.reshape(shape).astype(dtype)
        for k, a in loaded.items():
            expect[k][0].copy_(torch.from_numpy(a))
        self.step = man['step']
        g = man['gate']
        self.gate = GateParams(g['t_lo'], g['t_hi'], g['p_lo'], g['p_hi'],
                               g['rmax'])
        self.mlp('crs_mlp_to_bf16', _p(self.p), _p(self.pb), self.nparam,
                 _s())
        self._packed_step = None
