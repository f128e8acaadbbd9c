"""Mixed-precision path against the FP64 C oracle (GPU).

FP32 Arrhenius/falloff/third-body arithmetic with FP64 Kc, net rate and
species sums.  Scored with the same conditioned relative error as the FP64
path (create_rate_subs_amd.metrics, tau = 5e-2) on random states and on
near-equilibrium states (element-potential equilibrium, perturbed by a
relative 1e-4), where f and r of every reversible reaction cancel to 1e-4.
"""

import numpy as np
import pytest

from create_rate_subs_amd import metrics, states

from _helpers import hip_lib, mechanism, oracle, to_dev, to_host

pytestmark = pytest.mark.gpu

TAU = 5e-2
# FP32 exponential cores (v_exp_f32, ~1e-7 per rate constant) bound the
# random-state error at ~1e-7 / tau; near equilibrium the FP64 f - r keeps
# the error at FP32 level relative to the net rate.
BOUND_RANDOM = 5e-6
BOUND_NEAR_EQ = 1e-6


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
    assert s['max'] <= BOUND_RANDOM, s
    assert s['p999'] <= 1e-6, s


@pytest.mark.parametrize('name', ['mech53'])
def test_mixed_near_equilibrium(name):
    specs, _r, _f = mechanism(name)
    T, P, Y = states.equilibrium_states(specs, 16384, 42, perturb=1e-4)
    s = _score(name, T, P, Y)
    print(name, 'near-eq', s)
    assert s['max'] <= BOUND_NEAR_EQ, s


def test_mixed_mode_reported():
    lib = hip_lib('mech53')
    assert lib.modes & 2
