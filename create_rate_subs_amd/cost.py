"""Algorithmic cost of the wdot path per state, from the mechanism tables.

Counting convention (fixed once, used for every roofline figure): each
``+ - * /`` and each libm call (exp, log, log10, exp10/pow) in the
expressions the reference emits counts as one flop; fmax, comparisons,
negation and selects count as zero; no FMA credit.  Only the branch a
state takes is counted (one NASA range per Kc group).  Bytes are the HBM
traffic a state cannot avoid: read T, P and NSP-1 mass fractions, write
NSP rates, all FP64.
"""

from .hip_emitter import build_ir

__all__ = ['flops_per_state', 'bytes_per_state']


def _arr_flops(a):
    if a is None:
        return 0
    if not a.has_exp:
        return a.nT
    f = 1                                  # exp
    terms = (a.logA is not None) + (a.b is not None)
    if a.b is not None:
        f += 1                             # b * logT
    if a.E is not None:
        f += 1                             # E / T
        terms += 1
    f += max(terms - 1, 0)                 # joins of the terms
    if a.pre is not None:
        f += 1
    return f


def _prod_flops(terms):
    f = 0
    for _k, nu in terms:
        f += int(nu) if float(nu).is_integer() else 2
    return f


def flops_per_state(ir):
    """dict of flops per state by stage, plus 'total'."""
    nsp = ir.nsp
    conc = (nsp - 1) + nsp + (nsp - 1) + 1 + 3 + 2 * nsp
    rates = 1                              # logT
    pres = 2 if ir.npd else 0              # m
    spec = 0
    for x in ir.rxns:
        rates += _arr_flops(x.arr) + _prod_flops(x.reac)
        if x.rev:
            if x.rev_arr is not None:
                rates += _arr_flops(x.rev_arr) + _prod_flops(x.prod)
            else:
                rates += 12 * len(x.kc_groups) + (len(x.kc_groups) - 1) + 2
                rates += _prod_flops(x.prod) + 1
        if x.pd_index is not None:
            pres += 2 * len(x.eff)
            if not x.thd:
                pres += _arr_flops(x.low) + _arr_flops(x.high) + 2
                if x.falloff == 'troe':
                    pres += 11 if x.troe['t2'] is not None else 8
                    pres += 1 + 3 + 4 + 9
                elif x.falloff == 'sri':
                    pres += 1 + 3 + 2 + 3 + 1 + 1
                    if x.sri['d'] is not None:
                        pres += 3
                    pres += 3
                else:
                    pres += 2
        for _k, nu in x.spec:
            spec += (1 if x.rev else 0) + (1 if abs(nu) != 1 else 0)
            spec += (1 if x.pd_index is not None else 0) + 1
    return dict(conc=conc, rates=rates, pres_mod=pres, spec_rates=spec,
                total=conc + rates + pres + spec)


def bytes_per_state(nsp):
    """Unavoidable HBM bytes per state: 8 (2 NSP + 1)."""
    return 8 * (2 * nsp + 1)


def flops_for(specs, reacs):
    return flops_per_state(build_ir(specs, reacs))


def surrogate_flops_per_state(nin, hidden, layers, nout):
    """2 x multiply-adds of the MLP nin -> hidden^layers -> nout (unpadded):
    317,440 for 55 -> 3 x 256 -> 53."""
    dims = [nin] + [hidden] * layers + [nout]
    return 2 * sum(dims[i] * dims[i + 1] for i in range(len(dims) - 1))
