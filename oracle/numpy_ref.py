"""Independent NumPy / mpmath evaluator of the wdot path (test infrastructure).

A second statement of the same physics, written from the textbook forms
rather than from the reference's emitted text, used to pin the C oracle:

* concentrations   C_k = rho Y_k / W_k, rho = P Wbar / (RU T)
* rate constants   k = A T^b exp(-E/T)
* equilibrium      Kc = exp(sum_k nu_k (S_k/R - H_k/RT)) (PA/(RU T))^dnu
* third body       [M] = sum_k C_k + sum_listed (eff_k - 1) C_k
* Troe             log F = log Fcent / (1 + ((log Pr + c)/(n - d (log Pr + c)))^2),
                   c = -0.4 - 0.67 log Fcent, n = 0.75 - 1.27 log Fcent, d = 0.14
                   (Gilbert, Luther & Troe, Ber. Bunsenges. Phys. Chem. 87, 1983)
* SRI              F = (a e^{-b/T} + e^{-T/c})^X d T^e, X = 1/(1 + log10(Pr)^2)
                   (Stewart, Larson & Golden, Combust. Flame 75, 1989)
* falloff          pres_mod = F Pr/(1+Pr) (LOW given) or F/(1+Pr) (HIGH given)
* species rates    wdot = nu^T q, a dense matrix product
* PLOG             ln k linear in ln P between the bracketing table rows,
                   the end rows held outside the table
* Chebyshev        log10 k = sum_ij c_ij T_i(T~) T_j(P~) (numpy's chebval2d)
                   with T~ = (2/T - 1/Tmin - 1/Tmax) / (1/Tmax - 1/Tmin),
                   P~ = (2 log10 P - log10 Pmin - log10 Pmax) /
                        (log10 Pmax - log10 Pmin)

It reads the parsed data model (``SpecInfo``/``ReacInfo``), never the
emitted C or the HIP tables, so that it shares nothing with either beyond
the parser.  ``mp_state`` evaluates one state in mpmath at 50 digits.
"""

import math

import numpy as np

RU = 8314.4621
PA = 101325.0


def _smh(specs, T):
    T = np.asarray(T, dtype=np.float64)
    out = np.empty((len(specs), T.size))
    for k, sp in enumerate(specs):
        lo = np.asarray(sp.lo, dtype=np.float64)
        hi = np.asarray(sp.hi, dtype=np.float64)
        use_lo = T <= sp.Trange[1]
        for a, mask in ((lo, use_lo), (hi, ~use_lo)):
            t = T[mask]
            s_over_r = (a[0] * np.log(t) + a[1] * t + a[2] * t ** 2 / 2 +
                        a[3] * t ** 3 / 3 + a[4] * t ** 4 / 4 + a[6])
            h_over_rt = (a[0] + a[1] * t / 2 + a[2] * t ** 2 / 3 +
                         a[3] * t ** 3 / 4 + a[4] * t ** 4 / 5 + a[5] / t)
            out[k, mask] = s_over_r - h_over_rt
    return out


def _k(A, b, E, T):
    return A * T ** b * np.exp(-E / T)


def _kf(r, T, P):
    """Forward rate constant of reaction r at states (T, P), every form."""
    if r.plog:
        rows = r.plog_par
        lk = [np.log(_k(A, b, E, T)) for _p, A, b, E in rows]
        lp = [math.log(row[0]) for row in rows]
        lP = np.log(P)
        out = np.where(P <= rows[0][0], lk[0], lk[-1])
        for i in range(len(rows) - 1):
            inside = (P > rows[i][0]) & (P <= rows[i + 1][0])
            val = lk[i] + (lk[i + 1] - lk[i]) * (lP - lp[i]) / (lp[i + 1] - lp[i])
            out = np.where(inside, val, out)
        return np.exp(out)
    if r.cheb:
        from numpy.polynomial import chebyshev
        tmin, tmax = r.cheb_tlim
        pmin, pmax = r.cheb_plim
        tt = (2.0 / T - 1.0 / tmin - 1.0 / tmax) / (1.0 / tmax - 1.0 / tmin)
        pp = ((2.0 * np.log10(P) - math.log10(pmin) - math.log10(pmax)) /
              (math.log10(pmax) - math.log10(pmin)))
        return 10.0 ** chebyshev.chebval2d(tt, pp, np.asarray(r.cheb_par))
    return _k(r.A, r.b, r.E, T)


def nu_matrix(specs, reacs):
    """Net stoichiometric matrix [NSP][NR] (products minus reactants)."""
    nu = np.zeros((len(specs), len(reacs)))
    for j, r in enumerate(reacs):
        for k, v in zip(r.reac, r.reac_nu):
            nu[k, j] -= v
        for k, v in zip(r.prod, r.prod_nu):
            nu[k, j] += v
    return nu


def evaluate(specs, reacs, T, P, Y):
    """All intermediates for states (T[n], P[n], Y[NSP][n]); dict of arrays.

    ``Y`` row NSP-1 is ignored; the bath gas takes 1 - sum(others).
    """
    T = np.asarray(T, dtype=np.float64)
    P = np.asarray(P, dtype=np.float64)
    nsp = len(specs)
    W = np.array([sp.mw for sp in specs])
    y = np.array(Y[:nsp - 1], dtype=np.float64)
    yN = 1.0 - y.sum(axis=0)
    yall = np.vstack([y, yN[None, :]])
    wbar = 1.0 / (yall / W[:, None]).sum(axis=0)
    rho = P * wbar / (RU * T)
    C = rho * yall / W[:, None]
    smh = _smh(specs, T)

    nr = len(reacs)
    fwd = np.empty((nr, T.size))
    rev = []
    pm = []
    q = np.empty((nr, T.size))
    for j, r in enumerate(reacs):
        kf = _kf(r, T, P)
        cf = np.ones_like(T)
        for k, v in zip(r.reac, r.reac_nu):
            cf = cf * C[k] ** v
        fwd[j] = cf * kf
        net = fwd[j].copy()
        if r.rev:
            dnu = sum(r.prod_nu) - sum(r.reac_nu)
            lnKp = np.zeros_like(T)
            for k, v in zip(r.prod, r.prod_nu):
                lnKp += v * smh[k]
            for k, v in zip(r.reac, r.reac_nu):
                lnKp -= v * smh[k]
            Kc = np.exp(lnKp) * (PA / (RU * T)) ** dnu
            cr = np.ones_like(T)
            for k, v in zip(r.prod, r.prod_nu):
                cr = cr * C[k] ** v
            rr = cr * kf / Kc
            rev.append(rr)
            net = net - rr
        mod = None
        if r.thd_body or r.pdep:
            M = C.sum(axis=0)
            for k, eff in r.thd_body_eff:
                M = M + (eff - 1.0) * C[k]
        if r.thd_body:
            mod = M
        elif r.pdep:
            low = r.low if r.low else [r.A, r.b, r.E]
            high = r.high if r.high else [r.A, r.b, r.E]
            k0 = _k(low[0], low[1], low[2], T)
            kinf = _k(high[0], high[1], high[2], T)
            coll = C[r.pdep_sp] if r.pdep_sp != '' else M
            Pr = k0 * coll / kinf
            if r.troe:
                a, t3, t1 = r.troe_par[0:3]
                fc = (1 - a) * np.exp(-T / t3) + a * np.exp(-T / t1)
                if len(r.troe_par) == 4 and r.troe_par[3] != 0.0:
                    fc = fc + np.exp(-r.troe_par[3] / T)
                lfc = np.log10(np.maximum(fc, 1e-300))
                lpr = np.log10(np.maximum(Pr, 1e-300))
                c = -0.4 - 0.67 * lfc
                nn = 0.75 - 1.27 * lfc
                x = (lpr + c) / (nn - 0.14 * (lpr + c))
                F = 10.0 ** (lfc / (1.0 + x * x))
            elif r.sri:
                a, b, c = r.sri_par[0:3]
                lpr = np.log10(np.maximum(Pr, 1e-300))
                X = 1.0 / (1.0 + lpr * lpr)
                F = (a * np.exp(-b / T) + np.exp(-T / c)) ** X
                if (len(r.sri_par) == 5 and r.sri_par[3] != 1.0 and
                        r.sri_par[4] != 0.0):
                    F = F * r.sri_par[3] * T ** r.sri_par[4]
            else:
                F = np.ones_like(T)
            if r.low:
                mod = F * Pr / (1.0 + Pr)
            else:
                mod = F / (1.0 + Pr)
        if mod is not None:
            pm.append(mod)
            net = net * mod
        q[j] = net
    nu = nu_matrix(specs, reacs)
    wdot = nu @ q
    return dict(y_N=yN, mw=wbar, rho=rho, C=C, fwd=fwd,
                rev=np.array(rev) if rev else np.zeros((0, T.size)),
                pm=np.array(pm) if pm else np.zeros((0, T.size)),
                q=q, wdot=wdot)


def gross_scale(specs, reacs, fwd, rev, pm):
    """S_k = sum_j |nu_kj| (|f_j| + |r_j|) |pm_j|: the size of the terms
    whose sum is wdot_k, for conditioning-aware error metrics."""
    nu = np.abs(nu_matrix(specs, reacs))
    nr = len(reacs)
    g = np.abs(fwd).copy()
    ri = 0
    pi = 0
    for j, r in enumerate(reacs):
        if r.rev:
            g[j] += np.abs(rev[ri])
            ri += 1
        if r.thd_body or r.pdep:
            g[j] *= np.abs(pm[pi])
            pi += 1
    return nu @ g


def mp_state(specs, reacs, T, P, Y, dps=50):
    """One state in mpmath; returns wdot as a list of mpf (closed forms)."""
    import mpmath as mp
    mp.mp.dps = dps
    T = mp.mpf(T)
    P = mp.mpf(P)
    nsp = len(specs)
    y = [mp.mpf(float(v)) for v in Y[:nsp - 1]]
    y.append(1 - mp.fsum(y))
    W = [mp.mpf(sp.mw) for sp in specs]
    wbar = 1 / mp.fsum(y[k] / W[k] for k in range(nsp))
    rho = P * wbar / (mp.mpf(RU) * T)
    C = [rho * y[k] / W[k] for k in range(nsp)]

    def smh(sp):
        a = [mp.mpf(float(x)) for x in (sp.lo if T <= sp.Trange[1] else sp.hi)]
        return (a[0] * (mp.log(T) - 1) + a[1] * T / 2 + a[2] * T ** 2 / 6 +
                a[3] * T ** 3 / 12 + a[4] * T ** 4 / 20 - a[5] / T + a[6])

    def k(A, b, E):
        return mp.mpf(A) * T ** mp.mpf(b) * mp.exp(-mp.mpf(E) / T)

    def kf_of(r):
        if r.plog:
            rows = r.plog_par
            if P <= rows[0][0]:
                return k(*rows[0][1:4])
            for i in range(len(rows) - 1):
                if rows[i][0] < P <= rows[i + 1][0]:
                    l0 = mp.log(k(*rows[i][1:4]))
                    l1 = mp.log(k(*rows[i + 1][1:4]))
                    w = ((mp.log(P) - mp.log(rows[i][0])) /
                         (mp.log(rows[i + 1][0]) - mp.log(rows[i][0])))
                    return mp.exp(l0 + (l1 - l0) * w)
            return k(*rows[-1][1:4])
        if r.cheb:
            tmin, tmax = (mp.mpf(v) for v in r.cheb_tlim)
            pmin, pmax = (mp.mpf(v) for v in r.cheb_plim)
            tt = (2 / T - 1 / tmin - 1 / tmax) / (1 / tmax - 1 / tmin)
            pp = ((2 * mp.log10(P) - mp.log10(pmin) - mp.log10(pmax)) /
                  (mp.log10(pmax) - mp.log10(pmin)))
            acc = mp.mpf(0)
            for i in range(r.cheb_n_temp):
                for j in range(r.cheb_n_pres):
                    acc += (mp.mpf(float(r.cheb_par[i][j])) *
                            mp.chebyt(i, tt) * mp.chebyt(j, pp))
            return mp.power(10, acc)
        return k(r.A, r.b, r.E)

    wdot = [mp.mpf(0)] * nsp
    for r in reacs:
        kf = kf_of(r)
        f = kf
        for s, v in zip(r.reac, r.reac_nu):
            f *= C[s] ** v
        net = f
        if r.rev:
            dnu = sum(r.prod_nu) - sum(r.reac_nu)
            lnkp = (mp.fsum(v * smh(specs[s]) for s, v in zip(r.prod, r.prod_nu)) -
                    mp.fsum(v * smh(specs[s]) for s, v in zip(r.reac, r.reac_nu)))
            Kc = mp.exp(lnkp) * (mp.mpf(PA) / (mp.mpf(RU) * T)) ** dnu
            rr = kf / Kc
            for s, v in zip(r.prod, r.prod_nu):
                rr *= C[s] ** v
            net = f - rr
        if r.thd_body or r.pdep:
            M = mp.fsum(C) + mp.fsum((mp.mpf(e) - 1) * C[s]
                                     for s, e in r.thd_body_eff)
            if r.thd_body:
                net *= M
            else:
                low = r.low if r.low else [r.A, r.b, r.E]
                high = r.high if r.high else [r.A, r.b, r.E]
                coll = C[r.pdep_sp] if r.pdep_sp != '' else M
                Pr = k(*low[0:3]) * coll / k(*high[0:3])
                F = mp.mpf(1)
                if r.troe:
                    a, t3, t1 = (mp.mpf(x) for x in r.troe_par[0:3])
                    fc = (1 - a) * mp.exp(-T / t3) + a * mp.exp(-T / t1)
                    if len(r.troe_par) == 4 and r.troe_par[3] != 0.0:
                        fc += mp.exp(-mp.mpf(r.troe_par[3]) / T)
                    lfc = mp.log10(fc)
                    lpr = mp.log10(Pr)
                    c = -mp.mpf('0.4') - mp.mpf('0.67') * lfc
                    nn = mp.mpf('0.75') - mp.mpf('1.27') * lfc
                    x = (lpr + c) / (nn - mp.mpf('0.14') * (lpr + c))
                    F = mp.power(10, lfc / (1 + x * x))
                elif r.sri:
                    a, b, c = (mp.mpf(x) for x in r.sri_par[0:3])
                    X = 1 / (1 + mp.log10(Pr) ** 2)
                    F = (a * mp.exp(-b / T) + mp.exp(-T / c)) ** X
                    # d T^e, kept where the reference keeps it: only when
                    # d != 1 and e != 0 (rate_subs.py:1033-1037)
                    if (len(r.sri_par) == 5 and r.sri_par[3] != 1.0 and
                            r.sri_par[4] != 0.0):
                        F *= mp.mpf(r.sri_par[3]) * T ** mp.mpf(r.sri_par[4])
                net *= F * (Pr / (1 + Pr) if r.low else 1 / (1 + Pr))
        for s, v in zip(r.reac, r.reac_nu):
            wdot[s] -= v * net
        for s, v in zip(r.prod, r.prod_nu):
            wdot[s] += v * net
    return wdot


def mp_dydt(specs, reacs, T, P, Y, dps=50):
    """Constant-pressure right-hand side of one state in mpmath
    (rate_subs.py:1829-1946): [dT/dt, dY_0/dt, ..., dY_NSP-2/dt], with the
    mass-specific cp_k = RU/W_k (a0 + a1 T + ... + a4 T^4) and
    h_k = RU T/W_k (a0 + a1 T/2 + ... + a4 T^4/5 + a5/T) of each species'
    NASA range; also returns the mass-average cp and the density."""
    import mpmath as mp
    w = mp_state(specs, reacs, T, P, Y, dps)
    T = mp.mpf(T)
    P = mp.mpf(P)
    nsp = len(specs)
    y = [mp.mpf(float(v)) for v in Y[:nsp - 1]]
    y.append(1 - mp.fsum(y))
    W = [mp.mpf(sp.mw) for sp in specs]
    ru = mp.mpf(RU)
    rho = P / (ru * T * mp.fsum(y[k] / W[k] for k in range(nsp)))
    cp, h = [], []
    for k, sp in enumerate(specs):
        a = [mp.mpf(float(x)) for x in (sp.lo if T <= sp.Trange[1] else sp.hi)]
        cp.append(ru / W[k] * (a[0] + T * (a[1] + T * (a[2] + T * (a[3] +
                                                              T * a[4])))))
        h.append(ru * T / W[k] * (a[0] + T * (a[1] / 2 + T * (a[2] / 3 + T * (
            a[3] / 4 + T * a[4] / 5))) + a[5] / T))
    cp_avg = mp.fsum(y[k] * cp[k] for k in range(nsp))
    dT = -mp.fsum(w[k] * h[k] * W[k] for k in range(nsp)) / (rho * cp_avg)
    dY = [w[k] * W[k] / rho for k in range(nsp - 1)]
    return [dT] + dY, h, cp_avg, rho
