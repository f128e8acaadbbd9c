"""The surrogate's residual gate threshold (surrogate.gate_threshold), on
the CPU with hand-made errors and residuals: the admitted set is the
longest low-residual prefix whose worst error meets the target."""

import math

import pytest
import torch

from create_rate_subs_amd.surrogate import gate_threshold


def _t(v):
    return torch.tensor(v, dtype=torch.float64)


def test_longest_prefix_within_target():
    # ranked by residual: errors 1e-3, 5e-3, 2e-2, 1e-3
    e = _t([2e-2, 1e-3, 1e-3, 5e-3])
    r = _t([0.3, 0.4, 0.1, 0.2])
    assert gate_threshold(e, r, 1e-2) == (0.2, 0.5)
    assert gate_threshold(e, r, 3e-2) == (0.4, 1.0)
    assert gate_threshold(e, r, 1e-4) == (-1.0, 0.0)


def test_non_finite_never_admitted():
    e = _t([1e-3, math.nan, 1e-3])
    r = _t([0.1, 0.05, 0.2])
    # the NaN-error state ranks first, so nothing is admitted
    assert gate_threshold(e, r, 1e-2) == (-1.0, 0.0)
    e = _t([1e-3, 1e-3, 1e-3])
    r = _t([0.1, math.inf, 0.2])
    assert gate_threshold(e, r, 1e-2) == (0.2, 2 / 3)


def test_prefix_does_not_end_inside_a_tie():
    # two states share residual 0.2; the second of them fails the target,
    # so rmax = 0.2 would admit it: the prefix stops before the tie
    e = _t([1e-3, 1e-3, 5e-2, 1e-3])
    r = _t([0.1, 0.2, 0.2, 0.3])
    assert gate_threshold(e, r, 1e-2) == (0.1, 0.25)


@pytest.mark.parametrize('seed', range(5))
def test_admitted_states_meet_the_target(seed):
    g = torch.Generator().manual_seed(seed)
    n = 500
    e = 10.0 ** (-4 * torch.rand(n, generator=g, dtype=torch.float64))
    r = torch.floor(100 * torch.rand(n, generator=g, dtype=torch.float64))
    for tgt in (1e-3, 1e-2, 1e-1):
        rmax, frac = gate_threshold(e, r, tgt)
        passed = r <= rmax if frac > 0 else torch.zeros(n, dtype=torch.bool)
        assert passed.double().mean().item() == pytest.approx(frac)
        assert (e[passed] <= tgt).all()


def test_spearman_rank_correlation():
    import torch
    from create_rate_subs_amd import metrics
    a = torch.tensor([0.1, 3.0, 2.0, 5.0, 4.0], dtype=torch.float64)
    assert abs(metrics.spearman(a, a ** 3) - 1.0) < 1e-12
    assert abs(metrics.spearman(a, -a) + 1.0) < 1e-12
    # against scipy's definition on random data without ties
    import numpy as np
    from scipy.stats import spearmanr
    rng = np.random.default_rng(3)
    x, y = rng.normal(size=500), rng.normal(size=500)
    y = y + 0.5 * x
    ref = spearmanr(x, y).correlation
    got = metrics.spearman(torch.from_numpy(x), torch.from_numpy(y))
    assert abs(got - ref) < 1e-12
    # non-finite values rank last
    b = torch.tensor([1.0, float('inf'), 2.0, float('nan')],
                     dtype=torch.float64)
    c = torch.tensor([1.0, 4.0, 2.0, 3.0], dtype=torch.float64)
    assert metrics.spearman(b, c) > 0.7
