"""FP64 HIP path against the C oracle (GPU).

Stage kernels are compared on the same inputs: eval_conc and
eval_spec_rates must be bit-identical to the oracle (same operations, no
contraction, IEEE division on both sides); eval_rxn_rates and
get_rxn_pres_mod may differ by the last bits of libm's exp/log (glibc on
the host, the ROCm device library on the GPU), bounded at 1e-12 pointwise.
The fused kernel is scored with the conditioning-aware relative error of
create_rate_subs_amd.metrics against the oracle at <= 1e-12.  mech17tm has
two T_mid groups (1000 K and 1300 K), so states on either side of either
T_mid take each species' NASA range by its own T_mid.
"""

import numpy as np
import pytest

from create_rate_subs_amd import metrics, states

from _helpers import hip_lib, mechanism, oracle, to_dev, to_host

pytestmark = pytest.mark.gpu

MECHS = ['mech13', 'mech53', 'mech111', 'mech13co', 'mech21pc',
         'mech17tm', 'mech19fx', 'mech9ns']
# tau of the conditioned error: where |wdot_k| < tau * S_k, errors count
# relative to tau * S_k.  Every rate constant is exp(x) with |x| up to
# ~100 (ln A, E/T and the NASA polynomial of ln Kc), and two correct FP64
# evaluations of x round differently by up to |x| * 2^-52 ~ 2e-14, which
# is then the relative difference of the term; near cancellation that
# floor is 2e-14 S_k.  tau = 5e-2 puts the test's floor at 5e-14 S_k.
TAU = 5e-2


def _states(name, n, seed):
    specs, _r, _f = mechanism(name)
    return states.random_states(len(specs), n, seed)


@pytest.mark.parametrize('name', MECHS)
@pytest.mark.parametrize('n', [1, 257, 4096])
def test_eval_conc_bit_exact(name, n):
    T, P, Y = _states(name, n, 11)
    ref = oracle(name).eval_stages(T, P, Y)
    lib = hip_lib(name)
    out = lib.eval_conc(*to_dev(T, P, Y))
    for key in ('y_N', 'mw', 'rho', 'C'):
        np.testing.assert_array_equal(to_host(out[key]), ref[key], err_msg=key)


@pytest.mark.parametrize('name', MECHS)
def test_rxn_rates_and_pres_mod(name):
    T, P, Y = _states(name, 4096, 12)
    ref = oracle(name).eval_stages(T, P, Y)
    lib = hip_lib(name)
    Td, Pd, Cd = to_dev(T, P, ref['C'])
    fwd, rev = lib.eval_rxn_rates(Td, Pd, Cd)
    pm = lib.get_rxn_pres_mod(Td, Pd, Cd)
    for key, got in (('fwd', fwd), ('rev', rev), ('pm', pm)):
        r = ref[key]
        g = to_host(got)
        rel = np.abs(g - r) / np.maximum(np.abs(r), 1e-300)
        assert rel.max() <= 1e-12, (key, rel.max())


@pytest.mark.parametrize('name', MECHS)
def test_spec_rates_bit_exact(name):
    T, P, Y = _states(name, 4096, 13)
    o = oracle(name)
    ref = o.eval_stages(T, P, Y)
    lib = hip_lib(name)
    fwd, rev, pm = to_dev(ref['fwd'], ref['rev'] if lib.nrev else
                          np.zeros((1, 4096)),
                          ref['pm'] if lib.npd else np.zeros((1, 4096)))
    W = lib.eval_spec_rates(fwd, rev[:lib.nrev], pm[:lib.npd])
    np.testing.assert_array_equal(to_host(W), ref['wdot'])


@pytest.mark.parametrize('name', MECHS)
@pytest.mark.parametrize('n', [1, 100, 4096, 65536])
def test_wdot_fp64_matches_oracle(name, n):
    specs, reacs, _ = mechanism(name)
    T, P, Y = _states(name, n, 14)
    ref = oracle(name).eval_stages(T, P, Y)
    lib = hip_lib(name)
    W = to_host(lib.eval_wdot(*to_dev(T, P, Y)))
    S = metrics.gross_scale(specs, reacs, ref['fwd'], ref['rev'], ref['pm'])
    err = metrics.rel_err(W, ref['wdot'], S, TAU)
    s = metrics.summarize(err)
    s['tau'] = TAU
    s['pointwise_max'] = float(metrics.rel_err(W, ref['wdot'], S, 0.0).max())
    print(name, n, s)
    assert s['max'] <= 1e-12, s


@pytest.mark.parametrize('name', ['mech13co', 'mech53', 'mech17tm'])
def test_wdot_fp64_near_equilibrium(name):
    """Element-potential equilibrium states perturbed by 1e-4: forward and
    reverse rates nearly cancel in every reversible reaction, which the
    fused kernel forms as one FMA on H products (prod H_r - K prod H_p)
    and the oracle as f - r; the conditioned error keeps the same bound."""
    specs, reacs, _ = mechanism(name)
    T, P, Y = states.equilibrium_states(specs, 16384, 43, perturb=1e-4)
    ref = oracle(name).eval_stages(T, P, Y)
    lib = hip_lib(name)
    W = to_host(lib.eval_wdot(*to_dev(T, P, Y)))
    S = metrics.gross_scale(specs, reacs, ref['fwd'], ref['rev'], ref['pm'])
    s = metrics.summarize(metrics.rel_err(W, ref['wdot'], S, TAU))
    cancel = np.abs(ref['wdot']) / np.maximum(S, 1e-300)
    print(name, 'near-eq', s, 'median |w| / S', float(np.median(cancel)))
    assert s['max'] <= 1e-12, s


def test_wdot_fp64_matches_oracle_at_headline_size():
    """The headline configuration itself: 53 species, 2^20 states, every
    state against the oracle (the largest error of a sample grows with
    its size, so the bound is checked on the full batch)."""
    test_wdot_fp64_matches_oracle('mech53', 1 << 20)


@pytest.mark.parametrize('name', ['mech53'])
def test_wdot_subset(name):
    import torch
    T, P, Y = _states(name, 5000, 15)
    lib = hip_lib(name)
    Td, Pd, Yd = to_dev(T, P, Y)
    full = lib.eval_wdot(Td, Pd, Yd)
    idx = torch.tensor([4999, 0, 17, 2500, 17], dtype=torch.int32,
                       device='cuda')
    W = torch.full_like(full, np.nan)
    lib.eval_wdot_subset(Td, Pd, Yd, W, idx)
    Wh, Fh = to_host(W), to_host(full)
    sel = [0, 17, 2500, 4999]
    np.testing.assert_array_equal(Wh[:, sel], Fh[:, sel])
    rest = np.setdiff1d(np.arange(5000), sel)
    assert np.isnan(Wh[:, rest]).all()


@pytest.mark.parametrize('mode', ['fp64', 'mixed'])
@pytest.mark.parametrize('count', [0, 3, 5, 9])
def test_wdot_subset_device_count(mode, count):
    """crs_eval_wdot_subset_dev: the subset's length is read on the device
    and capped by the list's length m_max = 5."""
    import torch
    name = 'mech53'
    T, P, Y = _states(name, 5000, 15)
    lib = hip_lib(name)
    Td, Pd, Yd = to_dev(T, P, Y)
    full = to_host(lib.eval_wdot(Td, Pd, Yd, mode))
    order = [4999, 0, 17, 2500, 3]
    idx = torch.tensor(order, dtype=torch.int32, device='cuda')
    cnt = torch.tensor([count], dtype=torch.int32, device='cuda')
    W = torch.full_like(Yd, np.nan)
    lib.eval_wdot_subset_dev(Td, Pd, Yd, W, idx, cnt, mode)
    Wh = to_host(W)
    sel = sorted(order[:min(count, 5)])
    np.testing.assert_array_equal(Wh[:, sel], full[:, sel])
    rest = np.setdiff1d(np.arange(5000), sel)
    assert np.isnan(Wh[:, rest]).all()


@pytest.mark.parametrize('mode', ['fp64', 'mixed'])
def test_wdot_ragged_leading_dimension(mode):
    """Rows with a stride larger than the state count (views into wider
    buffers): the same rates as the compact layout, padding untouched."""
    import torch
    name = 'mech53'
    n, ld = 1000, 1037
    T, P, Y = _states(name, n, 16)
    lib = hip_lib(name)
    Td, Pd, Yd = to_dev(T, P, Y)
    compact = to_host(lib.eval_wdot(Td, Pd, Yd, mode))
    Ybig = torch.full((lib.nsp, ld), float('nan'), dtype=torch.float64,
                      device='cuda')
    Ybig[:, :n] = Yd
    Wbig = torch.full_like(Ybig, 7.0)
    lib.eval_wdot(Td, Pd, Ybig[:, :n], mode, out=Wbig[:, :n])
    Wh = to_host(Wbig)
    np.testing.assert_array_equal(Wh[:, :n], compact)
    assert (Wh[:, n:] == 7.0).all()


@pytest.mark.parametrize('mode', ['fp64', 'mixed'])
def test_wdot_subset_matches_full_batch(mode):
    import torch
    T, P, Y = _states('mech111', 3000, 17)
    lib = hip_lib('mech111')
    Td, Pd, Yd = to_dev(T, P, Y)
    full = to_host(lib.eval_wdot(Td, Pd, Yd, mode))
    rng = np.random.default_rng(3)
    sel = np.unique(rng.integers(0, 3000, 700))
    idx = torch.tensor(rng.permutation(sel), dtype=torch.int32, device='cuda')
    W = torch.full_like(Yd, np.nan)
    lib.eval_wdot_subset(Td, Pd, Yd, W, idx, mode)
    Wh = to_host(W)
    np.testing.assert_array_equal(Wh[:, sel], full[:, sel])
    rest = np.setdiff1d(np.arange(3000), sel)
    assert np.isnan(Wh[:, rest]).all()


def test_pdep_forms_over_wide_pressures():
    """PLOG (below, inside and above the table) and Chebyshev reactions:
    stage rates pointwise and the fused kernels conditioned, against the
    oracle at pressures from 0.005 to 500 atm."""
    name = 'mech21pc'
    specs, reacs, _ = mechanism(name)
    T, P, Y = _states(name, 8192, 18)
    rng = np.random.default_rng(19)
    P = np.exp(rng.uniform(np.log(0.005 * 101325.0), np.log(500 * 101325.0),
                           T.size))
    ref = oracle(name).eval_stages(T, P, Y)
    lib = hip_lib(name)
    Td, Pd, Yd = to_dev(T, P, Y)
    C = lib.eval_conc(Td, Pd, Yd)['C']
    fwd, rev = lib.eval_rxn_rates(Td, Pd, C)
    fe = np.abs(to_host(fwd) - ref['fwd']) / np.abs(ref['fwd'])
    re_ = np.abs(to_host(rev) - ref['rev']) / np.maximum(np.abs(ref['rev']),
                                                         1e-300)
    S = metrics.gross_scale(specs, reacs, ref['fwd'], ref['rev'], ref['pm'])
    W = to_host(lib.eval_wdot(Td, Pd, Yd))
    we = metrics.rel_err(W, ref['wdot'], S, TAU).max()
    print('fwd', fe.max(), 'rev', re_.max(), 'wdot', we)
    assert fe.max() <= 1e-12 and re_.max() <= 1e-12 and we <= 1e-12


def test_full_size_properties():
    """At the headline mixed-precision size (53 sp, 2^22 states), where the
    oracle would take minutes: the fused FP64 kernel against the chained
    stage kernels (a separate code path, pinned against the oracle above)
    to 1e-12 conditioned, and element and mass conservation of the fused
    rates to 1e-12 of their gross scale."""
    import torch
    name = 'mech53'
    specs, reacs, _ = mechanism(name)
    n = 1 << 22
    T, P, Y = _states(name, n, 20)
    lib = hip_lib(name)
    Td, Pd, Yd = to_dev(T, P, Y)
    del T, P, Y
    W = lib.eval_wdot(Td, Pd, Yd)
    Wst = torch.empty_like(W)
    S = torch.empty_like(W)
    chunk = 1 << 20
    for a in range(0, n, chunk):
        b = a + chunk
        Tc, Pc = Td[a:b].contiguous(), Pd[a:b].contiguous()
        C = lib.eval_conc(Tc, Pc, Yd[:, a:b].contiguous())['C']
        fwd, rev = lib.eval_rxn_rates(Tc, Pc, C)
        pm = lib.get_rxn_pres_mod(Tc, Pc, C)
        Wst[:, a:b] = lib.eval_spec_rates(fwd, rev, pm)
        S[:, a:b] = metrics.gross_scale(specs, reacs, fwd, rev, pm)
        del C, fwd, rev, pm
    e = metrics.rel_err(W, Wst, S, TAU).max().item()
    mw = torch.tensor([sp.mw for sp in specs], dtype=torch.float64,
                      device='cuda')
    elems = sorted({el.upper() for sp in specs for el, _c in sp.elem})
    A = torch.zeros((len(elems) + 1, len(specs)), dtype=torch.float64,
                    device='cuda')
    for k, sp in enumerate(specs):
        for el, c in sp.elem:
            A[elems.index(el.upper()), k] += c
    A[-1] = mw
    resid = (A @ W).abs() / (A.abs() @ S).clamp_min(1e-300)
    print('fused vs stages', e, 'balance residual', resid.max().item())
    assert e <= 1e-12
    assert resid.max().item() <= 1e-12


def test_wdot_empty_and_bad_args():
    import ctypes
    import torch
    lib = hip_lib('mech13')
    e = torch.empty(0, dtype=torch.float64, device='cuda')
    Y = torch.empty((lib.nsp, 0), dtype=torch.float64, device='cuda')
    W = lib.eval_wdot(e, e, Y)
    assert W.shape == (lib.nsp, 0)
    raw = lib.lib
    null = ctypes.c_void_p(0)
    # unknown mode, n < 0, ld < n, NULL pointer with work: all rejected
    T = torch.ones(4, dtype=torch.float64, device='cuda')
    Y4 = torch.ones((lib.nsp, 4), dtype=torch.float64, device='cuda')
    p = [ctypes.c_void_p(t.data_ptr()) for t in (T, T, Y4, Y4)]
    assert raw.crs_eval_wdot(*p, 4, 4, 7, null) == 1
    assert b'mode' in raw.crs_last_error()
    assert raw.crs_eval_wdot(*p, -1, 4, 0, null) == 1
    assert raw.crs_eval_wdot(*p, 4, 3, 0, null) == 1
    assert raw.crs_eval_wdot(p[0], null, p[2], p[3], 4, 4, 0, null) == 1
    assert b'P' in raw.crs_last_error()
    torch.cuda.synchronize()
