"""The generator's command line (CPU only).

``python -m create_rate_subs_amd -l hip -i mech -t therm [-ls SP] -o dir``
takes the reference's flags (create_rate_subs.py:561-593) and must write
exactly the source write_hip emits for the same mechanism, with the -ls
species moved last (the reference parses -ls and drops it at :593; here it
is honoured, and this is the documented divergence).
"""

import os
import re
import subprocess
import sys

import pytest

from create_rate_subs_amd import hip_emitter, workloads
from create_rate_subs_amd.create_rate_subs import load_mechanism

from _helpers import ROOT


def _run(*args):
    return subprocess.run([sys.executable, '-m', 'create_rate_subs_amd'] +
                          list(args), cwd=ROOT, capture_output=True,
                          text=True, timeout=300)


@pytest.mark.parametrize('last', [None, 'CO'])
def test_cli_emits_write_hip_source(tmp_path, last):
    mpath, tpath = workloads.mechanism_files('mech13')
    out = tmp_path / 'out'
    args = ['-l', 'hip', '-i', mpath, '-t', tpath, '-o', str(out)]
    if last:
        args += ['-ls', last]
    res = _run(*args)
    assert res.returncode == 0, res.stderr
    got = (out / 'crs_mech13.hip').read_text()
    _e, specs, reacs, fwd = load_mechanism(mpath, tpath, last_spec=last)
    ref_dir = tmp_path / 'ref'
    src = hip_emitter.write_hip(str(ref_dir), specs, reacs, name='mech13',
                                spec_order=fwd)
    assert got == open(src).read()
    # the species order the library reports ends with the requested species
    names = [sp.name for sp in specs]
    assert names[-1] == (last or 'N2')
    if last:
        assert 'N2' in names[:-1]


def test_cli_rejects_unknown_language(tmp_path):
    mpath, tpath = workloads.mechanism_files('mech13')
    res = _run('-l', 'fortran', '-i', mpath, '-t', tpath, '-o',
               str(tmp_path))
    assert res.returncode != 0
    assert re.search(r'invalid choice', res.stderr)


def test_cli_requires_input():
    res = _run('-l', 'hip')
    assert res.returncode != 0
    assert '-i/--input' in res.stderr
