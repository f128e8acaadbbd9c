"""Seeded thermochemical states in device layout (structure of arrays).

Layout shared by every path: ``T[n]`` [K], ``P[n]`` [Pa] and ``Y[NSP][n]``
mass fractions with the bath gas in the last row.  The exact paths read
rows 0..NSP-2 and form the bath-gas fraction as 1 - sum of the others, the
way the reference's ``eval_conc`` does (rate_subs.py:1321-1335); row NSP-1
is filled with that same value so the array is a complete composition.

Random states: T ~ U(800, 2500) K, P log-uniform in [1, 50] atm, the NSP-1
non-bath mass fractions ~ Dirichlet(0.1) scaled to 0.4, bath gas 0.6.

Near-equilibrium states: T ~ U(1800, 2500) K and a composition built from
element potentials, C_k = (PA / (RU T)) exp(smh_k + sum_e a_ek lambda_e),
for which every element-balanced reaction has prod C^nu = Kc exactly; the
non-bath species make up 40 % of the moles, then each mass fraction is
perturbed by a relative U(-1e-4, 1e-4).
"""

import math

import numpy as np

from . import chem

__all__ = ['random_states', 'equilibrium_states', 'species_smh']


def _finish(Y):
    """Set the bath-gas row to 1 - sum(others)."""
    Y[-1] = 1.0 - Y[:-1].sum(axis=0)
    return Y


def random_states(nsp, n, seed, t_range=(800.0, 2500.0)):
    """``n`` random states for a mechanism with ``nsp`` species."""
    rng = np.random.default_rng(seed)
    T = rng.uniform(t_range[0], t_range[1], n)
    P = np.exp(rng.uniform(0.0, math.log(50.0), n)) * chem.PA
    Y = np.empty((nsp, n))
    g = rng.standard_gamma(0.1, size=(nsp - 1, n))
    s = g.sum(axis=0)
    s[s == 0.0] = 1.0
    Y[:-1] = g * (0.4 / s)
    return T, P, _finish(Y)


def species_smh(specs, T):
    """S/R - H/(RT) for every species at every temperature in ``T``.

    Vectorised form of chem_utilities.py:251-293; shape [NSP][n].
    """
    T = np.asarray(T, dtype=np.float64)
    logT = np.log(T)
    out = np.empty((len(specs), T.size))
    for k, sp in enumerate(specs):
        lo = np.asarray(sp.lo, dtype=np.float64)
        hi = np.asarray(sp.hi, dtype=np.float64)
        a = np.where(T <= sp.Trange[1], lo[:, None], hi[:, None])
        out[k] = (a[0] * (logT - 1.0) + a[1] * T / 2.0 + a[2] * T * T / 6.0 +
                  a[3] * T ** 3 / 12.0 + a[4] * T ** 4 / 20.0 - a[5] / T +
                  a[6])
    return out


def equilibrium_states(specs, n, seed, perturb=1e-4, t_range=(1800.0, 2500.0),
                       bath_mole_fraction=0.6, spread=1.0):
    """``n`` element-potential equilibrium states, perturbed by ``perturb``.

    ``specs`` are the parsed species (bath gas last); their element
    compositions and NASA-7 data define the equilibrium.  With
    ``perturb=0`` the states are exactly Kc-consistent up to rounding.
    """
    rng = np.random.default_rng(seed)
    nsp = len(specs)
    T = rng.uniform(t_range[0], t_range[1], n)
    P = np.exp(rng.uniform(0.0, math.log(50.0), n)) * chem.PA

    elems = sorted({e.upper() for sp in specs[:-1] for e, _ in sp.elem})
    a = np.zeros((nsp - 1, len(elems)))
    for k, sp in enumerate(specs[:-1]):
        for e, cnt in sp.elem:
            a[k, elems.index(e.upper())] += cnt
    atoms = a.sum(axis=1)

    smh = species_smh(specs[:-1], T)                       # [NSP-1][n]
    lam = rng.normal(0.0, spread, size=(len(elems), n))
    ctot = P / (chem.RU * T)
    base = np.log(chem.PA / (chem.RU * T)) + smh + a @ lam  # ln C_k at s=0
    target = np.log((1.0 - bath_mole_fraction) * ctot)

    # shift every element potential by s so that sum_k C_k hits the target;
    # ln sum exp(base + atoms*s) is increasing in s, so bisect
    lo = np.full(n, -400.0)
    hi = np.full(n, 400.0)
    for _ in range(64):
        mid = 0.5 * (lo + hi)
        v = base + atoms[:, None] * mid
        m = v.max(axis=0)
        lse = m + np.log(np.exp(v - m).sum(axis=0))
        up = lse > target
        hi = np.where(up, mid, hi)
        lo = np.where(up, lo, mid)
    s = 0.5 * (lo + hi)
    C = np.exp(base + atoms[:, None] * s)
    C_bath = ctot - C.sum(axis=0)

    mw = np.array([sp.mw for sp in specs])
    mass = np.vstack([C * mw[:-1, None], (C_bath * mw[-1])[None, :]])
    Y = mass / mass.sum(axis=0)
    if perturb:
        Y[:-1] *= 1.0 + perturb * rng.uniform(-1.0, 1.0, size=(nsp - 1, n))
    return T, P, _finish(Y)
