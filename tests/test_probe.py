"""The ceiling probes of csrc/fma_probe.hip (create_rate_subs_amd.tune
ceiling / bandwidth): argument checks on the CPU, results on the GPU.

The FMA chains are checked against the same recurrence in exact rational
arithmetic rounded once per step (what one fused multiply-add does), the
copy bit for bit, and the row pattern against a float64 host evaluation.
"""

import ctypes
import os
from fractions import Fraction

import numpy as np
import pytest
# before the probe library is opened: it then binds to the HIP runtime
# torch loaded (same libamdhip64 soname) instead of loading a second one
import torch

from create_rate_subs_amd import build

VP = ctypes.c_void_p


def _lib():
    if not os.path.exists(build.PROBE_LIB):
        build.build_probe()
    lib = ctypes.CDLL(build.PROBE_LIB)
    lib.crs_probe_fma.restype = ctypes.c_int
    lib.crs_probe_fma.argtypes = [ctypes.c_int] * 4 + [VP, VP]
    lib.crs_probe_copy.restype = ctypes.c_int
    lib.crs_probe_copy.argtypes = [VP, VP, ctypes.c_int64, VP]
    lib.crs_probe_rows.restype = ctypes.c_int
    lib.crs_probe_rows.argtypes = [VP, VP, VP, ctypes.c_int64,
                                   ctypes.c_int64, ctypes.c_int, VP]
    return lib


def test_probe_arguments_refused_before_launch():
    lib = _lib()
    buf = (ctypes.c_double * 64)()
    p = ctypes.cast(buf, VP)
    assert lib.crs_probe_fma(3, 1, 1, 1, p, None) == -1       # k
    assert lib.crs_probe_fma(1, 5, 1, 1, p, None) == -1       # waves
    assert lib.crs_probe_fma(1, 1, 0, 1, p, None) == -1       # blocks
    assert lib.crs_probe_fma(1, 1, 1, 1, None, None) == -1
    assert lib.crs_probe_copy(p, p, 24, None) == -1           # not 16 B
    assert lib.crs_probe_copy(None, p, 16, None) == -1
    assert lib.crs_probe_rows(p, p, p, 6, 8, 4, None) == -1   # n % 4
    assert lib.crs_probe_rows(p, p, p, 8, 6, 1, None) == -1   # ld < n
    assert lib.crs_probe_rows(p, p, p, 4, 6, 1, None) == -1   # ld % 4
    assert lib.crs_probe_rows(p, p, p, 8, 8, 3, None) == -1   # per lane


def _fma_chain(a, b, c, steps):
    fa, fb, fc = Fraction(a), Fraction(b), Fraction(c)
    for _ in range(steps):
        fa = Fraction(float(fa * fb + fc))
    return float(fa)


@pytest.mark.gpu
@pytest.mark.parametrize('k', [1, 8])
def test_fma_chains_match_exact_rounding(k):
    lib = _lib()
    blocks, waves, n = 2, 1, 3
    out = torch.empty(blocks * 256 * waves, dtype=torch.float64,
                      device='cuda')
    s = torch.cuda.current_stream().cuda_stream
    assert lib.crs_probe_fma(k, waves, blocks, n, out.data_ptr(), s) == 0
    got = out.cpu().numpy()
    for tid in (0, 1, 63, 255):
        want = 0.0
        for j in range(k):
            want += _fma_chain(float(tid + j), 0.999, 1e-3, n * 16)
        assert got[tid] == want, (tid, got[tid], want)
        assert got[256 + tid] == want


@pytest.mark.gpu
@pytest.mark.parametrize('per_lane', [1, 2, 4])
def test_copy_and_row_pattern(per_lane):
    lib = _lib()
    s = torch.cuda.current_stream().cuda_stream
    g = torch.Generator(device='cpu').manual_seed(7)
    rows, n, ld = 53, 1024, 1028
    Y = torch.rand((rows, ld), generator=g, dtype=torch.float64).cuda()
    W = torch.rand((rows, ld), generator=g, dtype=torch.float64).cuda()
    dy = torch.full((rows, ld), 5.0, dtype=torch.float64, device='cuda')
    assert lib.crs_probe_rows(Y.data_ptr(), W.data_ptr(), dy.data_ptr(), n,
                              ld, per_lane, s) == 0
    torch.cuda.synchronize()
    Yh, Wh, dh = Y.cpu().numpy(), W.cpu().numpy(), dy.cpu().numpy()
    sc = (Yh[:, :n] * (1.0 + np.arange(rows)[:, None] * 1e-3)).sum(0)
    np.testing.assert_allclose(dh[:, :n], Wh[:, :n] * sc, rtol=1e-13)
    assert (dh[:, n:] == 5.0).all()
    dst = torch.zeros_like(Y)
    assert lib.crs_probe_copy(Y.data_ptr(), dst.data_ptr(), Y.numel() * 8,
                              s) == 0
    torch.cuda.synchronize()
    assert torch.equal(dst, Y)
