"""The seeded workloads: which mechanism, how many states, which seeds.

Mechanism seed 1000+k and state seed 2000+k belong to workload k.  The
53-species mechanism is shared by the FP64, mixed-precision and surrogate
workloads (it is generated once, from seed 1001).

==========  =====================================================  =========
workload    what                                                   states
==========  =====================================================  =========
w0_13       13 sp / 27 rxn, FP64, CPU-runnable                     4096
w1_53       53 sp / 325 rxn, FP64 drop-in                          2^20
w2_53       53 sp / 325 rxn, mixed precision (+ near-eq slice)     2^22 (+2^18)
w3_111      111 sp / 784 rxn, FP64 and mixed                       2^20
w4_53       53 sp / 325 rxn, learned surrogate                     2^22
==========  =====================================================  =========
"""

import os
from dataclasses import dataclass

from . import synth

__all__ = ['MECHANISMS', 'REORDERED', 'LIBRARIES', 'last_species',
           'WORKLOADS', 'Workload', 'mechanism_files',
           'BUILD_ROOT']

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD_ROOT = os.path.join(HERE, '_build')

MECHANISMS = synth.MECHS

# Libraries built from a synthetic mechanism with another species moved to
# the end (the generator's -ls option): name -> (mechanism, last species).
# The synthetic bath gas is inert; a reactive last species exercises the
# implied-species terms (dy_N) of the path and of dydt.
REORDERED = {'mech13co': ('mech13', 'CO')}
LIBRARIES = list(MECHANISMS) + list(REORDERED)


def last_species(name):
    """The species a library's mechanism is reordered to end with, or None."""
    return REORDERED[name][1] if name in REORDERED else None


@dataclass(frozen=True)
class Workload:
    name: str
    mech: str
    n_states: int
    state_seed: int
    near_eq_states: int = 0


WORKLOADS = {
    'w0_13': Workload('w0_13', 'mech13', 4096, 2000),
    'w1_53': Workload('w1_53', 'mech53', 1 << 20, 2001),
    'w2_53': Workload('w2_53', 'mech53', 1 << 22, 2002, near_eq_states=1 << 18),
    'w3_111': Workload('w3_111', 'mech111', 1 << 20, 2003),
    'w4_53': Workload('w4_53', 'mech53', 1 << 22, 2004),
}


def mechanism_files(name, directory=None):
    """Write (if needed) and return the Chemkin/thermo paths of a mechanism."""
    cfg = MECHANISMS[REORDERED[name][0] if name in REORDERED else name]
    directory = directory or os.path.join(BUILD_ROOT, 'mechs')
    mpath = os.path.join(directory, cfg.name + '.dat')
    tpath = os.path.join(directory, cfg.name + '_therm.dat')
    mech, thermo = synth.generate_mechanism(cfg)
    os.makedirs(directory, exist_ok=True)
    for path, text in ((mpath, mech), (tpath, thermo)):
        old = None
        if os.path.exists(path):
            with open(path) as fh:
                old = fh.read()
        if old != text:
            with open(path, 'w') as fh:
                fh.write(text)
    return mpath, tpath
