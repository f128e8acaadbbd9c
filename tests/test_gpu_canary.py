"""Guard bands around every output of the C ABI (GPU).

Each entry point runs on arrays cut out of larger buffers: guard rows
before and after every [rows][ld] array, padding columns n..ld-1 in every
row (ld > n, n not a multiple of the block size) and guard words around
every [n] vector, all holding a NaN canary bit pattern.  After the call the
canary must be intact everywhere outside the entries the entry point
documents as written, and every written entry must be finite (a canary
left there means an entry was never written).  Inputs are guarded the same
way and the unread last row of Y holds the canary too: a kernel that read
it, or read past its rows, would turn the canary's NaN into its results.
Mechanisms: mech13 (one wave per SIMD) and mech53 (two waves, the 512-
thread blocks of the fused kernels).
"""

import ctypes

import numpy as np
import pytest

from create_rate_subs_amd import states

from _helpers import hip_lib, mechanism

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')

CANARY = 0x7FF4DEAD5EEDF00D       # a quiet NaN
G = 2                             # guard rows of a [rows][ld] array
GV = 64                           # guard words of an [n] vector
N = 1037                          # not a multiple of any block size
LD = N + 29


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _s():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Guarded(object):
    """A [rows][ld] FP64 array inside guard rows, or an [n] vector inside
    guard words; ``view`` is the array the entry point sees."""

    def __init__(self, rows, n=N, ld=LD, vector=False):
        self.vector = vector
        if vector:
            self.buf = torch.full((n + 2 * GV,), CANARY, dtype=torch.int64,
                                  device='cuda').view(torch.float64)
            self.view = self.buf[GV:GV + n]
        else:
            self.buf = torch.full((rows + 2 * G, ld), CANARY,
                                  dtype=torch.int64,
                                  device='cuda').view(torch.float64)
            self.view = self.buf[G:G + rows, :n]
        self.rows, self.n = rows, n

    def fill(self, a):
        self.view.copy_(torch.as_tensor(np.ascontiguousarray(a),
                                        dtype=torch.float64))
        return self

    def check(self, what, cols=None, rows=None):
        """Canary intact outside rows x cols of the view (all of it by
        default), written and finite inside."""
        torch.cuda.synchronize()
        bits = self.buf.view(torch.int64).cpu().numpy()
        vals = self.buf.cpu().numpy()
        live = np.zeros(bits.shape, dtype=bool)
        cols = np.arange(self.n) if cols is None else np.asarray(cols)
        if self.vector:
            live[GV + cols] = True
        else:
            rows = np.arange(self.rows) if rows is None else np.asarray(rows)
            live[np.ix_(G + rows, cols)] = True
        outside = bits[~live]
        assert (outside == CANARY).all(), \
            '{}: {} entries written outside the output'.format(
                what, int((outside != CANARY).sum()))
        inside_bits, inside = bits[live], vals[live]
        assert not (inside_bits == CANARY).any(), \
            '{}: {} output entries never written'.format(
                what, int((inside_bits == CANARY).sum()))
        assert np.isfinite(inside).all(), \
            '{}: {} non-finite outputs'.format(
                what, int((~np.isfinite(inside)).sum()))


def _inputs(name, seed=71):
    specs, _r, _f = mechanism(name)
    nsp = len(specs)
    T, P, Y = states.random_states(nsp, N, seed)
    Tg = Guarded(1, vector=True).fill(T)
    Pg = Guarded(1, vector=True).fill(P)
    Yg = Guarded(nsp)
    Yg.view[:nsp - 1].copy_(torch.as_tensor(Y[:nsp - 1]))
    return nsp, Tg, Pg, Yg


@pytest.mark.parametrize('name', ['mech13', 'mech53'])
def test_stage_entry_points_stay_inside(name):
    lib = hip_lib(name)
    nsp, T, P, Y = _inputs(name)
    yN, mw, rho = (Guarded(1, vector=True) for _ in range(3))
    C = Guarded(nsp)
    assert lib.lib.crs_eval_conc(_p(T.view), _p(P.view), _p(Y.view),
                                 _p(yN.view), _p(mw.view), _p(rho.view),
                                 _p(C.view), N, LD, _s()) == 0
    for g, what in ((yN, 'y_N'), (mw, 'mw_avg'), (rho, 'rho'), (C, 'conc')):
        g.check('eval_conc ' + what)

    fwd, rev = Guarded(lib.nr), Guarded(max(lib.nrev, 1))
    assert lib.lib.crs_eval_rxn_rates(_p(T.view), _p(P.view), _p(C.view),
                                      _p(fwd.view), _p(rev.view), N, LD,
                                      _s()) == 0
    fwd.check('eval_rxn_rates fwd')
    rev.check('eval_rxn_rates rev', rows=np.arange(lib.nrev))

    pm = Guarded(max(lib.npd, 1))
    assert lib.lib.crs_get_rxn_pres_mod(_p(T.view), _p(P.view), _p(C.view),
                                        _p(pm.view), N, LD, _s()) == 0
    pm.check('get_rxn_pres_mod', rows=np.arange(lib.npd))

    # species rates with the bath gas's rate in a separate vector
    sp, dyN = Guarded(nsp - 1), Guarded(1, vector=True)
    assert lib.lib.crs_eval_spec_rates(_p(fwd.view), _p(rev.view),
                                       _p(pm.view), _p(sp.view),
                                       _p(dyN.view), N, LD, _s()) == 0
    sp.check('eval_spec_rates sp_rates')
    dyN.check('eval_spec_rates dy_N')


@pytest.mark.parametrize('name', ['mech13', 'mech53'])
@pytest.mark.parametrize('mode', [0, 1])
def test_fused_entry_points_stay_inside(name, mode):
    lib = hip_lib(name)
    nsp, T, P, Y = _inputs(name)
    W = Guarded(nsp)
    assert lib.lib.crs_eval_wdot(_p(T.view), _p(P.view), _p(Y.view),
                                 _p(W.view), N, LD, mode, _s()) == 0
    W.check('eval_wdot mode {}'.format(mode))

    # subset: only the listed columns are written (out of order, ragged)
    rng = np.random.default_rng(5)
    cols = np.sort(rng.choice(N, size=300, replace=False))
    idx = torch.as_tensor(rng.permutation(cols).astype(np.int32),
                          device='cuda')
    Ws = Guarded(nsp)
    assert lib.lib.crs_eval_wdot_subset(_p(T.view), _p(P.view), _p(Y.view),
                                        _p(Ws.view), _p(idx), idx.numel(), LD,
                                        mode, _s()) == 0
    Ws.check('eval_wdot_subset', cols=cols)
    torch.testing.assert_close(Ws.view[:, cols], W.view[:, cols], rtol=0,
                               atol=0)

    # device count below the list length: only the first k are written
    k = 117
    cnt = torch.tensor([k], dtype=torch.int32, device='cuda')
    Wd = Guarded(nsp)
    assert lib.lib.crs_eval_wdot_subset_dev(
        _p(T.view), _p(P.view), _p(Y.view), _p(Wd.view), _p(idx), _p(cnt),
        idx.numel(), LD, mode, _s()) == 0
    Wd.check('eval_wdot_subset_dev', cols=np.sort(idx[:k].cpu().numpy()))


@pytest.mark.parametrize('name', ['mech13', 'mech53'])
@pytest.mark.parametrize('extra_row', [False, True])
def test_dydt_and_rk4_stay_inside(name, extra_row):
    lib = hip_lib(name)
    nsp, T, P, Y = _inputs(name)
    rows = nsp + (1 if extra_row else 0)     # an integrator's NN-th slot
    y = Guarded(rows)
    y.view[0].copy_(T.view)
    y.view[1:nsp].copy_(Y.view[:nsp - 1])
    dy = Guarded(rows)
    assert lib.lib.crs_dydt(_p(P.view), _p(y.view), _p(dy.view), N, LD, 0,
                            _s()) == 0
    dy.check('dydt', rows=np.arange(nsp))

    # RK4 from near-equilibrium states (random compositions of these
    # synthetic mechanisms are too stiff for any explicit step), against
    # the same call on plain contiguous arrays
    specs, _r, _f = mechanism(name)
    Te, Pe, Ye = states.equilibrium_states(specs, N, 73, perturb=1e-3)
    Pg = Guarded(1, vector=True).fill(Pe)
    y.view[0].copy_(torch.as_tensor(Te))
    y.view[1:nsp].copy_(torch.as_tensor(Ye[:nsp - 1]))
    plain = torch.empty((rows, N), dtype=torch.float64, device='cuda')
    plain[:nsp] = y.view[:nsp]
    work = Guarded(3 * nsp)
    T0 = Guarded(1, vector=True).fill(Te)
    tig = Guarded(1, vector=True)
    tig.view.fill_(float('nan'))
    before = y.buf.clone()
    h = 1.0e-39
    assert lib.lib.crs_rk4(_p(Pg.view), _p(y.view), _p(work.view), N, LD,
                           ctypes.c_double(0.0), ctypes.c_double(h), 2,
                           _p(T0.view), _p(tig.view), ctypes.c_double(400.0),
                           _s()) == 0
    lib.rk4(torch.as_tensor(Pe, device='cuda'), plain[:nsp], 0.0, h, 2)
    torch.cuda.synchronize()
    # y: only its NSP live rows and n columns may change
    outside = torch.ones_like(y.buf, dtype=torch.bool)
    outside[G:G + nsp, :N] = False
    assert torch.equal(y.buf.view(torch.int64)[outside],
                       before.view(torch.int64)[outside])
    y.check('rk4 y', rows=np.arange(nsp))
    assert torch.equal(y.view[:nsp].contiguous().view(torch.int64),
                       plain[:nsp].view(torch.int64))
    work.check('rk4 work')
    # t_ign: NaN (not ignited) stays NaN at this step size; no guard word
    # touched
    bits = tig.buf.view(torch.int64).cpu().numpy()
    assert (bits[:GV] == CANARY).all() and (bits[GV + N:] == CANARY).all()


def test_surrogate_inference_stays_inside():
    from create_rate_subs_amd.surrogate import Surrogate
    name = 'mech13'
    lib = hip_lib(name)
    specs, reacs, _f = mechanism(name)
    nsp, T, P, Y = _inputs(name)
    Wx = Guarded(nsp)
    assert lib.lib.crs_eval_wdot(_p(T.view), _p(P.view), _p(Y.view),
                                 _p(Wx.view), N, LD, 0, _s()) == 0
    sur = Surrogate(lib, specs, reacs, seed=3)
    sur.fit_normalisation(T.view, P.view, Y.view, Wx.view)
    W = Guarded(nsp)
    sur.infer(T.view, P.view, Y.view, W.view)
    W.check('crs_mlp_infer')
