"""Constant-pressure dydt on the GPU (crs_dydt) against the C oracle.

The kernel fuses the wdot path with the reference's CONP right-hand side
(rate_subs.py:1829-1946): cp_avg, the enthalpies, the energy equation and
the mass-fraction rates.  Scored like wdot, with the conditioned error of
create_rate_subs_amd.metrics carried through the scalings:

* row 1 + k (dY_k/dt = wdot_k W_k / rho): gross scale S_k W_k / rho;
* row 0 (dT/dt = -sum_k wdot_k h_k W_k / (rho cp_avg)): gross scale
  sum_k |h_k W_k| S_k / (rho cp_avg).

mech13co is mech13 with CO moved last, so the implied species reacts and
its term of the energy equation is non-zero.  CRS_MODE_MIXED takes the
rates from the mixed kernel and is held to the mixed kernel's 1e-6.
"""

import numpy as np
import pytest

from create_rate_subs_amd import metrics, states

from _helpers import hip_lib, mechanism, oracle, to_dev, to_host

pytestmark = pytest.mark.gpu

TAU = 5e-2
LIBS = ['mech13', 'mech53', 'mech111', 'mech13co', 'mech21pc', 'mech17tm',
        'mech19fx']


def _inputs(name, n, seed, rows=None, ld=None):
    specs, _r, _f = mechanism(name)
    nsp = len(specs)
    T, P, Y = states.random_states(nsp, n, seed)
    rows = rows or nsp
    ld = ld or n
    y = np.full((rows, ld), np.nan)
    y[0, :n] = T
    y[1:nsp, :n] = Y[:nsp - 1]
    return T, P, Y, y


def _scales(name, T, P, Y):
    specs, reacs, _f = mechanism(name)
    o = oracle(name)
    st = o.eval_stages(T, P, Y)
    S = metrics.gross_scale(specs, reacs, st['fwd'], st['rev'], st['pm'])
    mw = np.array([sp.mw for sp in specs])
    cp, h = o.thermo(T)
    Yall = np.vstack([Y[:-1], 1.0 - Y[:-1].sum(0)])
    cp_avg = (cp * Yall).sum(0)
    rows = S[:-1] * mw[:-1, None] / st['rho']
    trow = (np.abs(h * mw[:, None]) * S).sum(0) / (st['rho'] * np.abs(cp_avg))
    return np.vstack([trow, rows])


def _score(got, ref, scale):
    return float(metrics.rel_err(got, ref, scale, TAU).max())


@pytest.mark.parametrize('name', LIBS)
@pytest.mark.parametrize('n', [1, 4096])
def test_dydt_matches_oracle(name, n):
    T, P, Y, y = _inputs(name, n, 30 + n)
    ref = oracle(name).dydt(P, y)
    lib = hip_lib(name)
    Pd, yd = to_dev(P, y)
    got = to_host(lib.dydt(Pd, yd))
    scale = _scales(name, T, P, Y)
    e_T = _score(got[:1], ref[:1], scale[:1])
    e_Y = _score(got[1:], ref[1:], scale[1:])
    print(name, n, 'dT/dt', e_T, 'dY/dt', e_Y)
    assert e_T <= 1e-12 and e_Y <= 1e-12


def test_dydt_integrator_layout_and_ragged_stride():
    """An NN = NSP + 1 row state (the reference integrator's) with a row
    stride above n: the NN-th row is neither read nor written."""
    import torch
    name = 'mech53'
    n, ld = 1000, 1029
    T, P, Y, y = _inputs(name, n, 44, rows=None, ld=None)
    lib = hip_lib(name)
    nsp = lib.nsp
    big = np.full((nsp + 1, ld), np.nan)
    big[:nsp, :n] = y
    Pd, yb = to_dev(P, big)
    dy = torch.full_like(yb, 5.0)
    lib.dydt(Pd, yb[:, :n], out=dy[:, :n])
    dyh = to_host(dy)
    compact = to_host(lib.dydt(*to_dev(P, y)))
    np.testing.assert_array_equal(dyh[:nsp, :n], compact)
    assert (dyh[nsp] == 5.0).all()
    assert (dyh[:, n:] == 5.0).all()


@pytest.mark.parametrize('name', ['mech13', 'mech53', 'mech111',
                                  'mech21pc'])
def test_mixed_dydt_within_the_mixed_bound(name):
    """CRS_MODE_MIXED: the mixed kernel's rates, then the same FP64 energy
    stage; against the oracle with the scales above, within the mixed
    kernel's 1e-6 target."""
    n = 4096
    T, P, Y, y = _inputs(name, n, 61)
    ref = oracle(name).dydt(P, y)
    lib = hip_lib(name)
    assert lib.modes & 2
    Pd, yd = to_dev(P, y)
    got = to_host(lib.dydt(Pd, yd, mode='mixed'))
    f64 = to_host(lib.dydt(Pd, yd))
    scale = _scales(name, T, P, Y)
    e_T = _score(got[:1], ref[:1], scale[:1])
    e_Y = _score(got[1:], ref[1:], scale[1:])
    print(name, 'mixed dT/dt', e_T, 'dY/dt', e_Y)
    assert e_T <= 1e-6 and e_Y <= 1e-6
    assert not np.array_equal(got, f64)       # the mixed kernel ran


def test_dydt_bad_sizes_and_empty():
    import torch
    lib = hip_lib('mech13')
    P = torch.ones(4, dtype=torch.float64, device='cuda') * 101325.0
    y = torch.ones((lib.nsp, 4), dtype=torch.float64, device='cuda')
    with pytest.raises(ValueError):
        lib.dydt(P, y[:lib.nsp - 1])
    e = torch.empty(0, dtype=torch.float64, device='cuda')
    out = lib.dydt(e, torch.empty((lib.nsp, 0), dtype=torch.float64,
                                  device='cuda'))
    assert out.shape == (lib.nsp, 0)


def test_dydt_interleaved_buffers_and_overlap_refused():
    """y and dy may share an allocation as long as no entry is shared
    (here side by side within each row, row stride 2n); any shared entry
    is refused before a launch, and so is work overlapping y in crs_rk4."""
    import torch
    from create_rate_subs_amd.runtime import CrsError
    name = 'mech53'
    n = 1000
    T, P, Y, y = _inputs(name, n, 45)
    lib = hip_lib(name)
    nsp = lib.nsp
    compact = to_host(lib.dydt(*to_dev(P, y)))
    Pd, yd = to_dev(P, y)
    buf = torch.full((nsp, 2 * n), 7.0, dtype=torch.float64, device='cuda')
    buf[:, :n] = yd
    lib.dydt(Pd, buf[:, :n], out=buf[:, n:])
    got = to_host(buf)
    np.testing.assert_array_equal(got[:, n:], compact)
    np.testing.assert_array_equal(got[:, :n], y)
    for out in (buf[:, :n], buf[:, 1:n + 1], buf[:, n - 1:2 * n - 1]):
        with pytest.raises(CrsError, match='y and dy overlap'):
            lib.dydt(Pd, buf[:, :n], out=out)
    np.testing.assert_array_equal(to_host(buf), got)
    blk = torch.zeros((4 * nsp, n), dtype=torch.float64, device='cuda')
    blk[:nsp] = yd
    with pytest.raises(CrsError, match='y and work overlap'):
        lib.rk4(Pd, blk[:nsp], 0.0, 1e-12, 1, work=blk[nsp - 1:4 * nsp - 1])
    np.testing.assert_array_equal(to_host(blk[:nsp]), y)
