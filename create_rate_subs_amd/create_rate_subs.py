"""Driver and command-line interface (create_rate_subs.py:430-593).

``create_rate_subs(lang, mech_name, therm_name=None, last_spec=None)``
parses a Chemkin mechanism, moves the bath gas to the last species index
and writes the rate subroutines for the chosen language into ``./out/``.
This package emits one language, ``hip``: a mechanism-specific HIP source
plus the C-ABI entry points declared in ``include/crs_hip.h``, built into a
shared library that Python loads with ctypes (``runtime.load``).

Same flags as the reference (``-l/-i/-t/-ls``, create_rate_subs.py:565-589)
plus ``-o`` (output directory, default ``./out/``) and ``--build``
(compile the emitted library with hipcc).  Unlike the reference, the
``-ls`` value is passed through (the reference parses it and then drops it,
create_rate_subs.py:593).
"""

import argparse
import os
import sys

from . import chem, mech_ct, utils
from . import mech_interpret as mech

__all__ = ['load_mechanism', 'create_rate_subs', 'main', 'langs']

langs = ['hip']


def _default_last_species(specs):
    """Index of the first of N2, Ar, He matched by name *and* exact MW.

    create_rate_subs.py:491-507.
    """
    wt = chem.get_elem_wt()
    for name, mw in (('N2', wt['n'] * 2.), ('Ar', wt['ar']), ('He', wt['he'])):
        for i, sp in enumerate(specs):
            if sp.name.lower() == name.lower() and sp.mw == mw:
                return i
    return None


def load_mechanism(mech_name, therm_name=None, last_spec=None, log=None):
    """Parse and reorder a mechanism as create_rate_subs.py:466-526 does.

    Returns ``(elems, specs, reacs, fwd_spec_mapping)`` with the bath gas
    (``last_spec`` if given and found, else the first of N2/Ar/He, else the
    mechanism's last species) moved to the last index and every reaction's
    species lists turned into indices.
    """
    say = log if log is not None else (lambda msg: None)
    # Cantera formats (create_rate_subs.py:466-471); YAML is read without
    # cantera, .cti/.xml need cantera itself
    if mech_name.endswith(('.cti', '.xml', '.yaml', '.yml')):
        elems, specs, reacs = mech_ct.read_mech_ct(mech_name)
    else:
        elems, specs, reacs = mech.read_mech(mech_name, therm_name)

    idx = None
    if last_spec is not None:
        idx = next((i for i, sp in enumerate(specs)
                    if sp.name.lower() == last_spec.lower().strip()), None)
        if idx is None:
            say('Warning: User specified last species {} not found in '
                'mechanism.  Attempting to find a default species.'
                .format(last_spec))
    if idx is None:
        idx = _default_last_species(specs)
        if idx is not None:
            say('Default last species {} found.'.format(specs[idx].name))
    if idx is None:
        say('Warning: Neither a user specified or default last species '
            'could be found. Proceeding using the last species in the base '
            'mechanism: {}'.format(specs[-1].name))
        idx = len(specs) - 1

    fwd, _back = utils.get_species_mappings(len(specs), idx)
    specs[:] = [specs[i] for i in fwd]
    utils.reassign_species_lists(reacs, specs)
    return elems, specs, reacs, fwd


def create_rate_subs(lang, mech_name, therm_name=None, last_spec=None,
                     out_dir='./out/', build=False, name=None):
    """Emit (and optionally build) the rate library for a mechanism.

    Returns 0 like the reference (create_rate_subs.py:557).  With
    ``build=True`` the shared library path is printed.
    """
    lang = lang.lower()
    if lang not in langs:
        raise ValueError('language needs to be one of: ' + ', '.join(langs))
    utils.create_dir(out_dir)
    _elems, specs, reacs, fwd = load_mechanism(mech_name, therm_name,
                                               last_spec, log=print)
    from . import hip_emitter
    if name is None:
        name = os.path.splitext(os.path.basename(mech_name))[0]
    def write(opts, d):
        return hip_emitter.write_hip(d, specs, reacs, name=name,
                                     spec_order=fwd, opts=opts)
    write({}, out_dir)
    if build:
        from . import build as _build
        print(_build.build_with_groups(write, out_dir))
    return 0


def main(argv=None):
    parser = argparse.ArgumentParser(
        description='Generates source code for species and reaction rates.')
    parser.add_argument('-l', '--lang', type=str, choices=langs,
                        required=True,
                        help='Programming language for output source files.')
    parser.add_argument('-i', '--input', type=str, required=True,
                        help='Input mechanism filename (e.g., mech.dat).')
    parser.add_argument('-t', '--thermo', type=str, default=None,
                        help='Thermodynamic database filename (e.g., '
                             'therm.dat), or nothing if in mechanism.')
    parser.add_argument('-ls', '--last_species', required=False, type=str,
                        default=None,
                        help='The name of the species to set as the last in '
                             'the mechanism. If not specified, defaults to '
                             'the first of N2, AR, and HE in the mechanism.')
    parser.add_argument('-o', '--out', type=str, default='./out/',
                        help='Output directory (default ./out/).')
    parser.add_argument('--build', action='store_true',
                        help='Compile the emitted HIP library for gfx950.')
    args = parser.parse_args(argv)
    return create_rate_subs(args.lang, args.input, args.thermo,
                            args.last_species, out_dir=args.out,
                            build=args.build)


if __name__ == '__main__':
    sys.exit(main())
