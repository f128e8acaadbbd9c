"""The wdot kernels at a non-default block size, against the oracle (GPU).

The block size of k_wdot_f64 / k_wdot_mixed is an emitter option
(``block``); the per-lane LDS rows per wave follow it.  Libraries built
with 128-thread blocks (tune.PARITY_VARIANTS, compiled by build()) must
give the default libraries' accuracy: the full batch and the subset path at
<= 1e-12 conditioned against the oracle, and the subset path bit-identical
to the same library's full batch.
"""

import json
import os

import numpy as np
import pytest
import torch

from create_rate_subs_amd import metrics, runtime, states, tune

from _helpers import mechanism, oracle, to_dev, to_host

pytestmark = pytest.mark.gpu

TAU = 5e-2
VARIANTS = sorted(k for k, o in tune.PARITY_VARIANTS.items() if 'block' in o)


def _lib(mech, variant):
    path = tune.variant_lib(mech, variant)
    with open(os.path.join(tune.variant_dir(mech, variant),
                           'opts.json')) as fh:
        built = json.load(fh)['opts']
    assert built == tune.PARITY_VARIANTS[(mech, variant)], built
    return runtime.load(path)


@pytest.mark.parametrize('mech,variant', VARIANTS)
@pytest.mark.parametrize('n', [100, 4096])
def test_block_variant_matches_oracle(mech, variant, n):
    specs, reacs, _ = mechanism(mech)
    T, P, Y = states.random_states(len(specs), n, 31)
    ref = oracle(mech).eval_stages(T, P, Y)
    lib = _lib(mech, variant)
    Td, Pd, Yd = to_dev(T, P, Y)
    W = to_host(lib.eval_wdot(Td, Pd, Yd))
    S = metrics.gross_scale(specs, reacs, ref['fwd'], ref['rev'], ref['pm'])
    e = metrics.rel_err(W, ref['wdot'], S, TAU).max()
    print(mech, variant, n, 'max conditioned', e)
    assert e <= 1e-12
    # subset path: a permuted half of the states, bit-identical to the
    # full batch of the same library, other columns untouched
    rng = np.random.default_rng(n)
    sel = np.unique(rng.integers(0, n, n // 2))
    idx = torch.tensor(rng.permutation(sel), dtype=torch.int32,
                       device='cuda')
    Ws = torch.full_like(Yd, np.nan)
    lib.eval_wdot_subset(Td, Pd, Yd, Ws, idx)
    Wh = to_host(Ws)
    np.testing.assert_array_equal(Wh[:, sel], W[:, sel])
    assert np.isnan(Wh[:, np.setdiff1d(np.arange(n), sel)]).all()


@pytest.mark.parametrize('mech,variant', VARIANTS)
def test_block_variant_mixed_matches_default(mech, variant):
    """Mixed mode at 128-thread blocks: the same arithmetic as the default
    library, so the same rates to the last bit."""
    specs, _r, _ = mechanism(mech)
    T, P, Y = states.random_states(len(specs), 3000, 32)
    Td, Pd, Yd = to_dev(T, P, Y)
    from create_rate_subs_amd import build
    base = runtime.load(build.lib_path(mech))
    a = to_host(_lib(mech, variant).eval_wdot(Td, Pd, Yd, 'mixed'))
    b = to_host(base.eval_wdot(Td, Pd, Yd, 'mixed'))
    np.testing.assert_array_equal(a, b)
