"""crs_mlp_eval_gated_noalloc against crs_mlp_eval_gated (GPU).

The no-allocation entry point keeps its fallback list and count in a static
device array of the surrogate library instead of caller workspace.  Through
the C ABI (ctypes), two calls in a row, the second with fewer states: the
rates it delivers must equal crs_mlp_eval_gated's bit for bit (the same
fused inference, the same gate and the same exact subset kernel), states
beyond the static list's size are refused with an error, and an empty call
does nothing.
"""

import ctypes

import pytest
import torch

from create_rate_subs_amd.runtime import MODES
from create_rate_subs_amd.surrogate import Surrogate, _p, _s, load_mlp
from create_rate_subs_amd import states

from _helpers import hip_lib, mechanism

pytestmark = pytest.mark.gpu

MAX_STATES = 1 << 22


def _bind(lib):
    fn = lib.lib.crs_mlp_eval_gated_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                   ctypes.c_int64, ctypes.c_void_p, ctypes.c_int,
                   ctypes.c_void_p]
    return fn


def _states(name, n, seed):
    specs, _r, _f = mechanism(name)
    T, P, Y = states.random_states(len(specs), n, seed)
    return [torch.from_numpy(a).cuda() for a in (T, P, Y)]


@pytest.mark.parametrize('name,mode', [('mech53', 'fp64'),
                                       ('mech13', 'mixed')])
def test_noalloc_matches_workspace_entry(name, mode):
    specs, reacs, _ = mechanism(name)
    lib = hip_lib(name)
    net = Surrogate(lib, specs, reacs, hidden=256, seed=29)
    Td, Pd, Yd = _states(name, 1 << 14, 91)
    net.train(Td, Pd, Yd, lib.eval_wdot(Td, Pd, Yd), epochs=1, batch=4096)
    mlp = load_mlp()
    noalloc = _bind(mlp)
    desc = net.net_desc()
    exact = net.exact_fn()
    for n, seed in ((5000, 92), (1234, 93)):      # second call smaller
        T, P, Y = _states(name, n, seed)
        T[7] = 9000.0                              # always sent back
        Wsur = torch.empty_like(Y)
        net.predict(T, P, Y, Wsur)
        net.gate.rmax = float(net.residuals(Wsur).median().item())
        desc = net.net_desc()
        ref = torch.full_like(Y, float('nan'))
        cnt = net.evaluate(T, P, Y, ref, mode=mode, read_count=False)
        out = torch.full_like(Y, float('nan'))
        rc = noalloc(ctypes.byref(desc), _p(T), _p(P), _p(Y), _p(out), n,
                     Y.stride(0), exact, MODES[mode], _s())
        assert rc == 0, mlp.lib.crs_mlp_last_error()
        torch.cuda.synchronize()
        m = int(cnt.item())
        print(name, mode, n, 'fallback', m)
        assert 0 < m < n
        assert torch.equal(out, ref)
        assert torch.equal(out[:, 7], lib.eval_wdot(T, P, Y, mode)[:, 7])
    # beyond the static list: refused before any work is enqueued
    rc = noalloc(ctypes.byref(desc), _p(T), _p(P), _p(Y), _p(out),
                 MAX_STATES + 1, MAX_STATES + 1, exact, MODES[mode], _s())
    assert rc != 0
    assert b'exceeds' in mlp.lib.crs_mlp_last_error()
    # nothing to do
    assert noalloc(ctypes.byref(desc), None, None, None, None, 0, 0, exact,
                   MODES[mode], _s()) == 0
