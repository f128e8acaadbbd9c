"""Write tests/golden/wdot_kat.json: known-answer wdot vectors.

For three synthetic mechanisms that between them hold every rate form of
the path (mech13: elementary, three-body, Troe, Lindemann; mech21pc: PLOG
and Chebyshev beside them; mech19fx: 3-parameter Troe, 5-parameter SRI, a
named collider, chemically activated reactions, split REV reactions and
non-integer orders), eight seeded states each, wdot evaluated in mpmath at
50 digits from the closed forms (oracle/numpy_ref.mp_state, which restates
the reference's definitions independently of both the C oracle and the
kernels) and rounded to double.  S, the gross scale of the conditioned
error metric, comes from the NumPy evaluator.

    python tests/golden/make_wdot_kat.py

The reference ships no golden vectors and cannot be run here (Python 2),
so these are vectors of the mathematics, not of the reference's output.
"""

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from create_rate_subs_amd import states, workloads  # noqa: E402
from create_rate_subs_amd.create_rate_subs import load_mechanism  # noqa: E402
from oracle import numpy_ref  # noqa: E402

MECHS = ('mech13', 'mech21pc', 'mech19fx')
COUNT = 8
SEED = 41
OUT = os.path.join(HERE, 'wdot_kat.json')


def make(name):
    mpath, tpath = workloads.mechanism_files(name)
    _e, specs, reacs, _f = load_mechanism(
        mpath, tpath, last_spec=workloads.last_species(name))
    T, P, Y = states.random_states(len(specs), COUNT, SEED)
    ref = numpy_ref.evaluate(specs, reacs, T, P, Y)
    S = numpy_ref.gross_scale(specs, reacs, ref['fwd'], ref['rev'],
                              ref['pm'])
    wdot = np.array([[float(v) for v in numpy_ref.mp_state(
        specs, reacs, T[i], P[i], Y[:, i])] for i in range(COUNT)]).T
    return dict(nsp=len(specs), states=COUNT, seed=SEED, T=T.tolist(),
                P=P.tolist(), Y=Y.tolist(), wdot=wdot.tolist(),
                S=np.asarray(S).tolist())


def main():
    data = dict(dps=50, mechanisms={name: make(name) for name in MECHS})
    with open(OUT, 'w') as fh:
        json.dump(data, fh, indent=0)
    print('wrote', OUT)


if __name__ == '__main__':
    main()
