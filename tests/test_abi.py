"""The C ABI: header declarations, exported symbols, host-only queries (CPU)."""

import ctypes
import os
import re

import pytest

from create_rate_subs_amd import build, runtime, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'crs_hip.h')


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r'^\S.*?\b(crs_\w+)\s*\(', text, re.M)))


def test_header_declares_exactly_the_runtime_symbols():
    assert _declared() == sorted(runtime.EXPORTED_SYMBOLS)


def _libs():
    out = []
    for name in workloads.LIBRARIES:
        p = build.lib_path(name)
        if os.path.exists(p):
            out.append((name, p))
    return out


def test_libraries_export_every_symbol():
    libs = _libs()
    if not any(n == 'mech13' for n, _ in libs):
        build.build_mechanism('mech13')
        libs = _libs()
    for name, path in libs:
        lib = ctypes.CDLL(path)
        for sym in _declared():
            assert hasattr(lib, sym), (name, sym)


def test_mech_info_is_host_only():
    path = build.lib_path('mech13')
    if not os.path.exists(path):
        build.build_mechanism('mech13')
    lib = ctypes.CDLL(path)

    class Info(ctypes.Structure):
        _fields_ = [('nsp', ctypes.c_int32), ('nr', ctypes.c_int32),
                    ('nrev', ctypes.c_int32), ('npd', ctypes.c_int32),
                    ('modes', ctypes.c_int32), ('name', ctypes.c_char_p)]
    info = Info()
    assert lib.crs_get_mech_info(ctypes.byref(info)) == 0
    assert (info.nsp, info.nr, info.nrev, info.npd) == (13, 27, 23, 4)
    assert info.name == b'mech13'
    lib.crs_last_error.restype = ctypes.c_char_p
    assert lib.crs_get_mech_info(None) == 1
    assert b'NULL' in lib.crs_last_error()
