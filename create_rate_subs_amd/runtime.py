"""ctypes binding of a mechanism library, on torch device tensors.

``load(path)`` opens one ``libcrs_<mech>.so`` (see include/crs_hip.h) and
returns a :class:`MechLib` whose methods take and return torch FP64 tensors
on the current GPU and enqueue on torch's current stream.  There is no
fallback: a missing library, a missing GPU or a failing call raises.

torch is imported before the library is opened so that the library binds
to the HIP runtime torch already loaded (same ``libamdhip64.so.7`` soname)
and shares its device memory and streams.
"""

import ctypes
import os

import torch

__all__ = ['CrsError', 'MechLib', 'load', 'MODES', 'EXPORTED_SYMBOLS']

MODES = {'fp64': 0, 'mixed': 1}

# Every entry point include/crs_hip.h declares.
EXPORTED_SYMBOLS = ('crs_get_mech_info', 'crs_last_error', 'crs_eval_conc',
                    'crs_eval_rxn_rates', 'crs_get_rxn_pres_mod',
                    'crs_eval_spec_rates', 'crs_eval_wdot',
                    'crs_eval_wdot_subset', 'crs_eval_wdot_subset_dev',
                    'crs_dydt', 'crs_rk4')


class CrsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__('[crs error {}] {}'.format(code, msg))
        self.code = code


class _Info(ctypes.Structure):
    _fields_ = [('nsp', ctypes.c_int32), ('nr', ctypes.c_int32),
                ('nrev', ctypes.c_int32), ('npd', ctypes.c_int32),
                ('modes', ctypes.c_int32), ('name', ctypes.c_char_p)]


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_SIGS = {
    'crs_get_mech_info': [ctypes.POINTER(_Info)],
    'crs_eval_conc': [_P] * 7 + [_I64, _I64, _P],
    'crs_eval_rxn_rates': [_P] * 5 + [_I64, _I64, _P],
    'crs_get_rxn_pres_mod': [_P] * 4 + [_I64, _I64, _P],
    'crs_eval_spec_rates': [_P] * 5 + [_I64, _I64, _P],
    'crs_eval_wdot': [_P] * 4 + [_I64, _I64, ctypes.c_int, _P],
    'crs_eval_wdot_subset': [_P] * 5 + [_I64, _I64, ctypes.c_int, _P],
    'crs_eval_wdot_subset_dev': [_P] * 6 + [_I64, _I64, ctypes.c_int, _P],
    'crs_dydt': [_P] * 3 + [_I64, _I64, ctypes.c_int, _P],
    'crs_rk4': [_P] * 3 + [_I64, _I64, ctypes.c_double, ctypes.c_double,
                           ctypes.c_int, _P, _P, ctypes.c_double, _P],
}


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class MechLib(object):
    """One mechanism library.  Sizes: ``nsp``, ``nr``, ``nrev``, ``npd``."""

    def __init__(self, path):
        if not os.path.exists(path):
            raise FileNotFoundError('mechanism library not built: ' + path)
        self.path = os.path.abspath(path)
        self.lib = ctypes.CDLL(self.path)
        for name in EXPORTED_SYMBOLS:
            fn = getattr(self.lib, name)
            fn.restype = ctypes.c_int
            if name in _SIGS:
                fn.argtypes = _SIGS[name]
        self.lib.crs_last_error.restype = ctypes.c_char_p
        self.lib.crs_last_error.argtypes = []
        info = _Info()
        self._check(self.lib.crs_get_mech_info(ctypes.byref(info)))
        self.nsp, self.nr = info.nsp, info.nr
        self.nrev, self.npd = info.nrev, info.npd
        self.modes = info.modes
        self.name = info.name.decode()

    # -- helpers -----------------------------------------------------------
    def _check(self, rc):
        if rc != 0:
            raise CrsError(rc, self.lib.crs_last_error().decode())

    @staticmethod
    def _vec(x, n, what):
        if not (x.is_cuda and x.dtype == torch.float64 and x.dim() == 1 and
                x.is_contiguous() and x.numel() == n):
            raise ValueError('{} must be a contiguous cuda float64 vector '
                             'of length {}'.format(what, n))

    @staticmethod
    def _rows(x, rows, n, what):
        if not (x.is_cuda and x.dtype == torch.float64 and x.dim() == 2 and
                x.shape[0] == rows and x.shape[1] == n and x.stride(1) == 1):
            raise ValueError('{} must be a cuda float64 [{}][{}] array with '
                             'unit column stride'.format(what, rows, n))
        return x.stride(0) if rows > 1 else max(x.stride(0), n)

    def _new(self, rows, n, ld=None):
        ld = ld or n
        buf = torch.empty((max(rows, 1), ld), dtype=torch.float64,
                          device='cuda')
        return buf[:rows, :n]

    # -- the four batched reference functions -----------------------------
    def eval_conc(self, T, P, Y):
        n = T.numel()
        self._vec(T, n, 'T')
        self._vec(P, n, 'P')
        ld = self._rows(Y, self.nsp, n, 'Y')
        y_N = torch.empty(n, dtype=torch.float64, device='cuda')
        mw = torch.empty_like(y_N)
        rho = torch.empty_like(y_N)
        C = self._new(self.nsp, n, ld)
        self._check(self.lib.crs_eval_conc(_ptr(T), _ptr(P), _ptr(Y),
                                           _ptr(y_N), _ptr(mw), _ptr(rho),
                                           _ptr(C), n, ld, _stream()))
        return dict(y_N=y_N, mw=mw, rho=rho, C=C)

    def eval_rxn_rates(self, T, P, C):
        n = T.numel()
        self._vec(T, n, 'T')
        self._vec(P, n, 'P')
        ld = self._rows(C, self.nsp, n, 'C')
        fwd = self._new(self.nr, n, ld)
        rev = self._new(self.nrev, n, ld)
        self._check(self.lib.crs_eval_rxn_rates(
            _ptr(T), _ptr(P), _ptr(C), _ptr(fwd), _ptr(rev), n, ld,
            _stream()))
        return fwd, rev

    def get_rxn_pres_mod(self, T, P, C):
        n = T.numel()
        self._vec(T, n, 'T')
        self._vec(P, n, 'P')
        ld = self._rows(C, self.nsp, n, 'C')
        pm = self._new(self.npd, n, ld)
        self._check(self.lib.crs_get_rxn_pres_mod(
            _ptr(T), _ptr(P), _ptr(C), _ptr(pm), n, ld, _stream()))
        return pm

    def eval_spec_rates(self, fwd, rev, pm):
        n = fwd.shape[1]
        ld = self._rows(fwd, self.nr, n, 'fwd')
        if self.nrev:
            if self._rows(rev, self.nrev, n, 'rev') != ld:
                raise ValueError('rev and fwd must share the row stride')
        if self.npd:
            if self._rows(pm, self.npd, n, 'pres_mod') != ld:
                raise ValueError('pres_mod and fwd must share the row stride')
        W = self._new(self.nsp, n, ld)
        self._check(self.lib.crs_eval_spec_rates(
            _ptr(fwd), _ptr(rev) if self.nrev else None,
            _ptr(pm) if self.npd else None, _ptr(W),
            ctypes.c_void_p(W.data_ptr() + 8 * (self.nsp - 1) * ld), n, ld,
            _stream()))
        return W

    # -- fused path --------------------------------------------------------
    def eval_wdot(self, T, P, Y, mode='fp64', out=None):
        n = T.numel()
        self._vec(T, n, 'T')
        self._vec(P, n, 'P')
        ld = self._rows(Y, self.nsp, n, 'Y')
        W = out if out is not None else self._new(self.nsp, n, ld)
        if self._rows(W, self.nsp, n, 'wdot') != ld:
            raise ValueError('wdot and Y must share the row stride')
        self._check(self.lib.crs_eval_wdot(_ptr(T), _ptr(P), _ptr(Y), _ptr(W),
                                           n, ld, MODES[mode], _stream()))
        return W

    def dydt(self, P, y, out=None, mode='fp64'):
        """Constant-pressure RHS (the reference's dydt, CONP): y [>= NSP][ld]
        with row 0 = T and rows 1..NSP-1 the mass fractions; returns dy of
        the same shape (rows beyond NSP-1 untouched)."""
        n = P.numel()
        self._vec(P, n, 'P')
        rows = y.shape[0] if y.dim() == 2 else -1
        if rows not in (self.nsp, self.nsp + 1):
            raise ValueError('y must have NSP or NSP+1 rows')
        ld = self._rows(y, rows, n, 'y')
        dy = out if out is not None else self._new(rows, n, ld)
        if self._rows(dy, dy.shape[0], n, 'dy') != ld or \
                dy.shape[0] not in (self.nsp, self.nsp + 1):
            raise ValueError('dy must have NSP or NSP+1 rows and the row '
                             'stride of y')
        self._check(self.lib.crs_dydt(_ptr(P), _ptr(y), _ptr(dy), n, ld,
                                      MODES[mode], _stream()))
        return dy

    def rk4(self, P, y, t0, h, nsteps, work=None, T0=None, t_ign=None,
            dT_ign=400.0):
        """nsteps RK4 steps of the CONP system on y (updated in place);
        returns y.  ``work``: [3 NSP][ld] scratch (allocated if None).
        With T0 and t_ign: t_ign (initialised to NaN) gets the start time
        of the first step after which T >= T0 + dT_ign."""
        n = P.numel()
        self._vec(P, n, 'P')
        rows = y.shape[0] if y.dim() == 2 else -1
        if rows not in (self.nsp, self.nsp + 1):
            raise ValueError('y must have NSP or NSP+1 rows')
        ld = self._rows(y, rows, n, 'y')
        if work is None:
            work = self._new(3 * self.nsp, n, ld)
        elif self._rows(work, 3 * self.nsp, n, 'work') != ld:
            raise ValueError('work must share the row stride of y')
        if (T0 is None) != (t_ign is None):
            raise ValueError('T0 and t_ign go together')
        if T0 is not None:
            self._vec(T0, n, 'T0')
            self._vec(t_ign, n, 't_ign')
        self._check(self.lib.crs_rk4(_ptr(P), _ptr(y), _ptr(work), n, ld,
                                     float(t0), float(h), int(nsteps),
                                     _ptr(T0), _ptr(t_ign), float(dT_ign),
                                     _stream()))
        return y

    def eval_wdot_subset_dev(self, T, P, Y, W, idx, count, mode='fp64'):
        """States idx[t], t < min(count[0], idx.numel()); the count is a
        device int32 tensor that is never read on the host."""
        n = T.numel()
        ld = self._rows(Y, self.nsp, n, 'Y')
        if self._rows(W, self.nsp, n, 'wdot') != ld:
            raise ValueError('wdot and Y must share the row stride')
        for t, what in ((idx, 'idx'), (count, 'count')):
            if not (t.is_cuda and t.dtype == torch.int32 and t.dim() == 1
                    and t.is_contiguous()):
                raise ValueError(what + ' must be a contiguous cuda int32 '
                                 'vector')
        self._check(self.lib.crs_eval_wdot_subset_dev(
            _ptr(T), _ptr(P), _ptr(Y), _ptr(W), _ptr(idx), _ptr(count),
            idx.numel(), ld, MODES[mode], _stream()))
        return W

    def eval_wdot_subset(self, T, P, Y, W, idx, mode='fp64'):
        n = T.numel()
        ld = self._rows(Y, self.nsp, n, 'Y')
        if self._rows(W, self.nsp, n, 'wdot') != ld:
            raise ValueError('wdot and Y must share the row stride')
        if not (idx.is_cuda and idx.dtype == torch.int32 and idx.dim() == 1
                and idx.is_contiguous()):
            raise ValueError('idx must be a contiguous cuda int32 vector')
        self._check(self.lib.crs_eval_wdot_subset(
            _ptr(T), _ptr(P), _ptr(Y), _ptr(W), _ptr(idx), idx.numel(), ld,
            MODES[mode], _stream()))
        return W


_LOADED = {}


def load(path):
    """Open (once per process) the mechanism library at ``path``."""
    path = os.path.abspath(path)
    if path not in _LOADED:
        _LOADED[path] = MechLib(path)
    return _LOADED[path]
