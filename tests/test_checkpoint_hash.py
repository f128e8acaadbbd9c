"""The mechanism hash a surrogate checkpoint carries (CPU only).

It must be stable for the same tables, and change with the species order,
a single constant's last bit, a third-body efficiency or a NASA
coefficient, so a network trained on one mechanism never loads silently
into another with the same species count.
"""

import copy

import numpy as np

from create_rate_subs_amd.surrogate import mechanism_hash

from _helpers import mechanism


def test_hash_is_stable_and_distinguishes_mechanisms():
    s13, r13, _ = mechanism('mech13')
    s13co, r13co, _ = mechanism('mech13co')
    s53, r53, _ = mechanism('mech53')
    h = mechanism_hash(s13, r13)
    assert h == mechanism_hash(copy.deepcopy(s13), copy.deepcopy(r13))
    assert len(s13) == len(s13co)
    assert h != mechanism_hash(s13co, r13co)
    assert h != mechanism_hash(s53, r53)


def test_hash_sees_single_constants():
    specs, reacs, _ = mechanism('mech53')
    h = mechanism_hash(specs, reacs)

    r = copy.deepcopy(reacs)
    r[7].A = np.nextafter(r[7].A, np.inf)
    assert mechanism_hash(specs, r) != h

    r = copy.deepcopy(reacs)
    j = next(i for i, x in enumerate(r) if x.thd_body_eff)
    sp, eff = r[j].thd_body_eff[0]
    r[j].thd_body_eff[0] = (sp, eff + 0.01)
    assert mechanism_hash(specs, r) != h

    s = copy.deepcopy(specs)
    s[3].lo[5] += 1e-9
    assert mechanism_hash(s, reacs) != h

    r = copy.deepcopy(reacs)
    j = next(i for i, x in enumerate(r) if x.troe)
    r[j].troe_par[0] *= 1.001
    assert mechanism_hash(specs, r) != h
