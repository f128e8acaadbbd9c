"""Helpers shared by the parser, the emitter and the runtime.

Restates utils.py:40-79 and :199-228 of the reference (species reordering,
net stoichiometric coefficients, integer test) and the number formatting the
reference's generator relies on.  The generator is Python 2, so every place
where it prints a float with ``str()`` or an empty ``'{}'`` format spec it
gets 12 significant digits, not Python 3's shortest round-trip repr.  The
functions below reproduce those spellings so that every constant this
package derives has exactly the value the reference would have emitted.
"""

import errno
import os

__all__ = ['py2_str', 'fmt_e16', 'fmt_e8', 'fmt_g6', 'fmt_w3', 'emitted',
           'get_nu', 'is_integer', 'get_species_mappings',
           'reassign_species_lists', 'create_dir']


def _add_dot_0(text):
    """Append '.0' when a %g rendering looks like an integer (Py_DTSF_ADD_DOT_0)."""
    if any(c in text for c in '.eEnN'):
        return text
    return text + '.0'


def py2_str(x):
    """``str(float)`` / ``'{}'.format(float)`` as Python 2.7 spells it.

    Python 2.7 renders floats with ``%.12g`` and appends '.0' to integral
    results.  Used by the reference for the Arrhenius exponent b
    (rate_subs.py:78,93), the literal A (:68), third-body efficiency offsets
    (:909-911) and falloff thresholds (:593).
    """
    if isinstance(x, bool) or isinstance(x, int):
        return str(x)
    return _add_dot_0('%.12g' % x)


def fmt_e16(x):
    """``'{:.16e}'`` — 17 significant digits, round-trip exact."""
    return '{:.16e}'.format(x)


def fmt_e8(x):
    """``'{:.8e}'`` — 9 significant digits (RU, Troe parameters)."""
    return '{:.8e}'.format(x)


def fmt_g6(x):
    """``'{:.6}'`` as used for SRI parameters (rate_subs.py:1020-1036).

    Both Python 2.7 and 3 treat an omitted presentation type with a
    precision as ``g`` plus a forced '.0' on integral results.
    """
    return _add_dot_0('%.6g' % x)


def fmt_w3(x):
    """``'{:3}'`` for a fractional stoichiometric coefficient (rate_subs.py:1193).

    In Python 2.7 an omitted type with no precision is ``%.12g`` with '.0'
    added, right-aligned in a field of width 3.
    """
    return '{:>3}'.format(py2_str(float(x)))


def emitted(text):
    """Numeric value of an emitted literal: what a C compiler reads back."""
    return float(text)


def get_nu(isp, rxn):
    """Net stoichiometric coefficient of species ``isp`` in ``rxn``.

    utils.py:63-79: products minus reactants, 0 for non-participants or
    species whose net change is zero.
    """
    in_prod = isp in rxn.prod
    in_reac = isp in rxn.reac
    if in_prod and in_reac:
        nu = (rxn.prod_nu[rxn.prod.index(isp)] -
              rxn.reac_nu[rxn.reac.index(isp)])
        return 0 if nu == 0 else nu
    if in_prod:
        return rxn.prod_nu[rxn.prod.index(isp)]
    if in_reac:
        return -rxn.reac_nu[rxn.reac.index(isp)]
    return 0


def is_integer(val):
    """True if ``val`` holds an integral value, int or float (utils.py:215-228)."""
    if isinstance(val, bool):
        return True
    if isinstance(val, int):
        return True
    try:
        return float(val).is_integer()
    except (TypeError, ValueError, OverflowError):
        return False


def get_species_mappings(num_specs, last_species):
    """Species permutation that moves ``last_species`` to the end.

    utils.py:40-60.  Returns ``(fwd, back)``: ``fwd[i]`` is the original
    index of the species placed at new position ``i``; ``back[j]`` is the new
    position of original species ``j``.
    """
    fwd = [i for i in range(num_specs) if i != last_species] + [last_species]
    back = [0] * num_specs
    for new, old in enumerate(fwd):
        back[old] = new
    return fwd, back


def reassign_species_lists(reacs, specs):
    """Replace species names in every reaction by their index in ``specs``.

    utils.py:199-212; third-body efficiency names and a named falloff
    collider are mapped too.
    """
    index = {sp.name: i for i, sp in enumerate(specs)}
    for rxn in reacs:
        rxn.reac = [index[s] for s in rxn.reac]
        rxn.prod = [index[s] for s in rxn.prod]
        rxn.thd_body_eff = [(index[name], eff)
                            for name, eff in rxn.thd_body_eff]
        if rxn.pdep_sp != '':
            rxn.pdep_sp = index[rxn.pdep_sp]


def create_dir(path):
    """mkdir -p (utils.py:125-144)."""
    try:
        os.makedirs(path)
    except OSError as exc:
        if exc.errno != errno.EEXIST:
            raise
