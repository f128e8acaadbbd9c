"""Host AddressSanitizer / UndefinedBehaviorSanitizer runs (CPU only).

The CPU oracle and the C ABI shims of a mechanism library and of the
surrogate library, built with -fsanitize=address,undefined by
tests/sanitize/run_sanitizers.py and run without a GPU; plus two planted
mistakes in the oracle driver that the checks must catch.
"""

import importlib.util
import os
import shutil

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
_spec = importlib.util.spec_from_file_location(
    'run_sanitizers', os.path.join(HERE, 'sanitize', 'run_sanitizers.py'))
san = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(san)

needs_gcc = pytest.mark.skipif(shutil.which('gcc') is None,
                               reason='gcc not found')
needs_hipcc = pytest.mark.skipif(not os.path.exists(san.pkg_build.hipcc()),
                                 reason='hipcc not found')


@needs_gcc
@pytest.mark.parametrize('mech', ['mech13', 'mech19fx', 'mech21pc'])
def test_oracle_under_sanitizers(mech):
    rc, out = san.run(san.build_oracle_san(mech), 37, 2)
    assert rc == 0, out
    assert out.rstrip().endswith('ok'), out


@needs_gcc
def test_oracle_single_state_single_thread():
    rc, out = san.run(san.build_oracle_san('mech13'), 1, 1)
    assert rc == 0 and out.rstrip().endswith('ok'), out


@needs_gcc
def test_planted_padding_write_is_caught():
    rc, out = san.run(san.build_oracle_san('mech13', plant='PAD_WRITE'), 37,
                      2)
    assert rc == 1, out
    assert 'padding entries written' in out


@needs_gcc
def test_planted_overflow_is_caught():
    rc, out = san.run(san.build_oracle_san('mech13', plant='OOB_WRITE'), 37,
                      2)
    assert rc == 23, out              # ASAN_OPTIONS exitcode
    assert 'heap-buffer-overflow' in out


@needs_hipcc
@pytest.mark.parametrize('mech', ['mech13', 'mech19fx'])
def test_c_abi_shims_under_sanitizers(mech):
    rc, out = san.run(san.build_abi_san(mech))
    assert rc == 0, out
    assert out.rstrip().endswith('ok'), out
