"""Error metrics for species production rates.

wdot_k = sum_j nu_kj q_j is a sum of terms of both signs, and near
equilibrium q_j = f_j - r_j itself cancels, so a pointwise relative error
|dw|/|w| is unbounded even between two correct FP64 evaluations (a one-ulp
difference in exp or log is amplified by the cancellation).  Every result
is therefore scored with a conditioning-aware relative error

    err_k = |w_hat_k - w_k| / max(|w_k|, tau * S_k),

    S_k = sum_j |nu_kj| (|f_j| + |r_j|) |pres_mod_j|   (gross rate scale),

computed against an FP64 reference w.  Where |w_k| >= tau S_k (the
well-conditioned entries) this is the plain relative error; below, errors
are measured relative to tau times the magnitude of the terms being summed.
An entry with w_hat == w scores 0 (this covers species in no reaction,
where w = S = 0).

For the learned surrogate a per-state normwise error is also reported,
max_k |w_hat_k - w_k| / max_k |w_k|: the error relative to the state's
largest rate, the usual figure for vector-valued surrogates.
"""

import numpy as np

__all__ = ['nu_matrix', 'gross_scale', 'rel_err', 'normwise_err',
           'summarize']


def _is_torch(x):
    return type(x).__module__.split('.')[0] == 'torch'


def nu_matrix(specs, reacs):
    """Net stoichiometric matrix [NSP][NR] (products minus reactants)."""
    nu = np.zeros((len(specs), len(reacs)))
    for j, r in enumerate(reacs):
        for k, v in zip(r.reac, r.reac_nu):
            nu[k, j] -= v
        for k, v in zip(r.prod, r.prod_nu):
            nu[k, j] += v
    return nu


def _gross_terms(reacs, fwd, rev, pm, xp):
    g = xp.abs(fwd).clone() if _is_torch(fwd) else np.abs(fwd).copy()
    ri = pi = 0
    for j, r in enumerate(reacs):
        if r.rev:
            g[j] += xp.abs(rev[ri])
            ri += 1
        if r.thd_body or r.pdep:
            g[j] *= xp.abs(pm[pi])
            pi += 1
    return g


def gross_scale(specs, reacs, fwd, rev, pm):
    """S_k for every state; works on numpy arrays or torch tensors."""
    nu = np.abs(nu_matrix(specs, reacs))
    if _is_torch(fwd):
        import torch
        g = _gross_terms(reacs, fwd, rev, pm, torch)
        return torch.as_tensor(nu, dtype=g.dtype, device=g.device) @ g
    g = _gross_terms(reacs, fwd, rev, pm, np)
    return nu @ g


def rel_err(w_hat, w, S, tau):
    """Element-wise conditioned relative error (numpy or torch)."""
    if _is_torch(w):
        import torch
        d = (w_hat - w).abs()
        den = torch.maximum(w.abs(), tau * S)
        e = torch.where(d == 0, torch.zeros_like(d),
                        d / torch.where(den == 0, torch.ones_like(den), den))
        return e
    d = np.abs(np.asarray(w_hat) - np.asarray(w))
    den = np.maximum(np.abs(w), tau * np.asarray(S))
    with np.errstate(divide='ignore', invalid='ignore'):
        e = np.where(d == 0, 0.0, d / np.where(den == 0, 1.0, den))
    return e


def summarize(err):
    """max / median / 99.9th percentile of an error array."""
    if _is_torch(err):
        e = err.flatten().double()
        n = e.numel()
        if n > (1 << 24):
            # quantile on an evenly strided sample keeps the sort cheap
            e_q = e[::(n >> 24) + 1]
        else:
            e_q = e
        k = max(1, int(e_q.numel() * 0.999))
        q999 = e_q.kthvalue(k).values.item() if n else 0.0
        return dict(max=e.max().item(), median=e_q.median().item(),
                    p999=q999)
    e = np.asarray(err).ravel()
    return dict(max=float(e.max()), median=float(np.median(e)),
                p999=float(np.quantile(e, 0.999)))


def normwise_err(w_hat, w):
    """Per-state max_k |w_hat_k - w_k| / max_k |w_k| over rows k ([NSP][n]).

    0 where both are identical (including all-zero states); inf where the
    exact rates are all zero but the estimate is not.
    """
    if _is_torch(w):
        import torch
        d = (w_hat - w).abs().amax(dim=0)
        m = w.abs().amax(dim=0)
        out = d / m
        out = torch.where(d == 0, torch.zeros_like(out), out)
        return torch.where(torch.isnan(out), torch.full_like(out, float('inf')),
                           out)
    d = np.abs(w_hat - w).max(axis=0)
    m = np.abs(w).max(axis=0)
    with np.errstate(divide='ignore', invalid='ignore'):
        out = d / m
    out = np.where(d == 0, 0.0, out)
    return np.where(np.isnan(out), np.inf, out)


def spearman(a, b):
    """Spearman rank correlation of two per-state vectors (torch); non-finite
    values rank above every finite one."""
    import torch

    def rank(x):
        x = x.double()
        x = torch.where(torch.isfinite(x), x, torch.full_like(x, 1e300))
        r = torch.empty_like(x)
        r[torch.argsort(x, stable=True)] = torch.arange(
            x.numel(), device=x.device, dtype=x.dtype)
        return r - r.mean()
    ra, rb = rank(a), rank(b)
    return float((ra * rb).sum() / (ra.norm() * rb.norm()))
