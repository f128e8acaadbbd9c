"""Mixed-precision path against the FP64 C oracle (GPU).

Kc (the species G_k), the net rate and the species sums in full FP64; the
Arrhenius and falloff transcendentals at about 32 bits (degree-5
exponential, 2.2e-10 relative).  Scored with the same conditioned relative
error as the FP64 path (create_rate_subs_amd.metrics, tau = 5e-2) on random
states, on the full 2^22-state mixed-precision workload (against the FP64
kernel, itself pinned to the oracle at 1e-12) and on near-equilibrium
states (element-potential equilibrium, perturbed by a relative 1e-4),
where f and r of every reversible reaction cancel to 1e-4.  The target is
1e-6 maximum relative error.
"""

import numpy as np
import pytest

from create_rate_subs_amd import metrics, states

from _helpers import hip_lib, mechanism, oracle, to_dev, to_host

pytestmark = pytest.mark.gpu

TAU = 5e-2
# the target: 1e-6 maximum relative error against FP64.  A rate constant
# off by 2.2e-10 relative moves a species sum by at most 2.2e-10 / tau
# under the conditioned metric.
BOUND = 1e-6


def _score(name, T, P, Y):
    specs, reacs, _ = mechanism(name)
    ref = oracle(name).eval_stages(T, P, Y)
    lib = hip_lib(name)
    W = to_host(lib.eval_wdot(*to_dev(T, P, Y), mode='mixed'))
    S = metrics.gross_scale(specs, reacs, ref['fwd'], ref['rev'], ref['pm'])
    s = metrics.summarize(metrics.rel_err(W, ref['wdot'], S, TAU))
    s['tau'] = TAU
    ok = np.abs(ref['wdot']) >= TAU * S
    pw = metrics.rel_err(W[ok], ref['wdot'][ok], S[ok], 0.0)
    s['well_conditioned_pointwise_max'] = float(pw.max()) if pw.size else 0.0
    s['well_conditioned_fraction'] = float(ok.mean())
    return s


@pytest.mark.parametrize('name', ['mech13', 'mech53', 'mech111', 'mech21pc',
                                  'mech17tm', 'mech19fx'])
def test_mixed_random_states(name):
    specs, _r, _f = mechanism(name)
    T, P, Y = states.random_states(len(specs), 65536, 41)
    s = _score(name, T, P, Y)
    print(name, 'random', s)
    assert s['max'] <= BOUND, s


@pytest.mark.parametrize('name', ['mech53'])
def test_mixed_near_equilibrium(name):
    specs, _r, _f = mechanism(name)
    T, P, Y = states.equilibrium_states(specs, 16384, 42, perturb=1e-4)
    s = _score(name, T, P, Y)
    print(name, 'near-eq', s)
    assert s['max'] <= BOUND, s


def test_mixed_at_full_workload_size():
    """The mixed-precision workload itself: 53 species, 2^22 states, every
    state against the FP64 kernel (the oracle would take minutes)."""
    import torch
    name = 'mech53'
    specs, reacs, _ = mechanism(name)
    n = 1 << 22
    T, P, Y = states.random_states(len(specs), n, 2002)
    lib = hip_lib(name)
    Td, Pd, Yd = to_dev(T, P, Y)
    del T, P, Y
    W64 = lib.eval_wdot(Td, Pd, Yd, 'fp64')
    Wmx = lib.eval_wdot(Td, Pd, Yd, 'mixed')
    worst = 0.0
    chunk = 1 << 20
    for a in range(0, n, chunk):
        b = a + chunk
        Tc, Pc = Td[a:b].contiguous(), Pd[a:b].contiguous()
        C = lib.eval_conc(Tc, Pc, Yd[:, a:b].contiguous())['C']
        fwd, rev = lib.eval_rxn_rates(Tc, Pc, C)
        pm = lib.get_rxn_pres_mod(Tc, Pc, C)
        S = metrics.gross_scale(specs, reacs, fwd, rev, pm)
        e = metrics.rel_err(Wmx[:, a:b], W64[:, a:b], S, TAU)
        worst = max(worst, e.max().item())
        del C, fwd, rev, pm, S, e
    print('mixed vs FP64 at 2^22 states: max', worst)
    assert torch.isfinite(Wmx).all()
    assert worst <= BOUND


def test_mixed_mode_reported():
    lib = hip_lib('mech53')
    assert lib.modes & 2


def test_fp32_variant_against_the_oracle():
    """The FP32 variant of the mixed kernel (emitter option mixed_arith=
    'fp32': v_exp_f32 cores, FP32 falloff and third-body blocks), built
    beside the default library: within 5e-6 of the oracle, FP32's ~1e-7
    per rate constant amplified by cancellation up to 1 / tau."""
    import json
    import os
    from create_rate_subs_amd import runtime, tune
    name = 'mech53'
    with open(os.path.join(tune.variant_dir(name, 'fp32'),
                           'opts.json')) as fh:
        assert json.load(fh)['opts'] == {'mixed_arith': 'fp32'}
    specs, reacs, _ = mechanism(name)
    T, P, Y = states.random_states(len(specs), 65536, 41)
    ref = oracle(name).eval_stages(T, P, Y)
    lib = runtime.load(tune.variant_lib(name, 'fp32'))
    W = to_host(lib.eval_wdot(*to_dev(T, P, Y), mode='mixed'))
    S = metrics.gross_scale(specs, reacs, ref['fwd'], ref['rev'], ref['pm'])
    s = metrics.summarize(metrics.rel_err(W, ref['wdot'], S, TAU))
    print('fp32 variant', s)
    assert s['max'] <= 5e-6, s
