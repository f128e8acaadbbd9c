"""Chemkin mechanism and NASA-7 thermo parsing.

A Python 3 restatement of the reference's reader (mech_interpret.py:52-712
``read_mech`` and :715-863 ``read_thermo``) with the same unit conventions,
auxiliary keywords and reaction bookkeeping, so that a mechanism parsed here
produces the same constants the reference would emit.

Knowing divergences (both only turn a hang or crash of the reference into a
clean result or error):

* ``read_thermo`` stops at end of file; the reference spins forever there
  when no END line follows (mech_interpret.py:762-772).
* a species line that gives only two temperatures with no common-range line
  raises ``ValueError`` instead of a ``NameError`` (mech_interpret.py:753-758).
"""

import copy
import math
import re

import numpy as np

from . import chem

__all__ = ['read_mech', 'read_thermo', 'act_energy_fact', 'pre_units',
           'act_energy_units', 'MechanismError']

# mech_interpret.py:31-33
pre_units = ['moles', 'molecules']
act_energy_units = ['kelvins', 'evolts', 'cal/mole', 'joules/kmole',
                    'kcal/mole', 'joules/mole', 'kjoules/mole']

# Activation-energy factor to an activation temperature [K]
# (mech_interpret.py:39-46).
act_energy_fact = {
    'kelvins': 1.0,
    'evolts': 11595.,
    'cal/mole': 4.184 / chem.RU_JOUL,
    'kcal/mole': 4184. / chem.RU_JOUL,
    'joules/mole': 1. / chem.RU_JOUL,
    'kjoules/mole': 1000.0 / chem.RU_JOUL,
    'joules/kmole': 1. / (chem.RU_JOUL * 1000.),
}

_BLANK = re.compile(r'^\s*$')
_COMMENT = re.compile(r'^\s*!')


class MechanismError(ValueError):
    """Malformed or unsupported mechanism input."""


def _skippable(line):
    return bool(_BLANK.search(line) or _COMMENT.search(line))


def _split_side(side):
    """Split one side of a reaction equation into species terms.

    Handles the '(+M)' / '(+SP)' falloff marker and species whose names
    contain '+' or parentheses, as mech_interpret.py:241-336 does.  Returns
    ``(terms, pdep, pdep_sp, thd)`` where ``terms`` is a list of
    ``(name, nu)`` before merging and ``thd`` is set for a bare 'M' term or
    a '(+M)' marker.
    """
    pdep = False
    thd = False
    pdep_sp = ''
    rest = side
    offset = 0
    while '(' in rest:
        i1 = rest.find('(')
        i2 = rest.find(')')
        inner = rest[i1 + 1:i2].strip()
        if inner == '+':
            # '(+)' inside a species name
            offset += i2 + 1
            rest = rest[i2 + 1:]
        elif inner[:1] == '+':
            pdep = True
            pdep_sp = rest[i1 + 1:i2].replace('+', ' ').strip()
            if pdep_sp.lower() == 'm':
                thd = True
                pdep_sp = ''
            side = side[:offset + i1] + side[offset + i2 + 1:]
            break
        else:
            offset += i2 + 1
            rest = rest[i2 + 1:]

    pieces = side.split('+')
    # consecutive '+' means a species name ended in '+'
    while '' in pieces:
        k = pieces.index('')
        pieces[k - 1] = pieces[k - 1] + '+'
        del pieces[k]
    # re-join names like 'A(+)B' that the split tore apart
    k = 0
    while k < len(pieces) - 1:
        if pieces[k][-1:] == '(' and pieces[k + 1][:1] == ')':
            pieces[k] = pieces[k] + '+' + pieces[k + 1]
            del pieces[k + 1]
        k += 1

    terms = []
    for piece in pieces:
        piece = piece.strip()
        if piece[:1].isdigit():
            i = 0
            while i < len(piece) and not piece[i].isalpha():
                i += 1
            coef = piece[:i]
            nu = float(coef) if '.' in coef else int(coef)
            piece = piece[i:].strip()
        else:
            nu = 1
        terms.append((piece, nu))
    return terms, pdep, pdep_sp, thd


def _merge(terms, is_m):
    names, nus = [], []
    thd = False
    for name, nu in terms:
        if is_m(name):
            thd = True
            continue
        if name in names:
            nus[names.index(name)] += nu
        else:
            names.append(name)
            nus.append(nu)
    return names, nus, thd


def _aux_numbers(line):
    return line.replace('/', ' ').replace(',', ' ').split()


def read_mech(mech_filename, therm_filename=None, elem_wt=None):
    """Parse a Chemkin mechanism; return ``(elems, specs, reacs)``.

    Follows mech_interpret.py:52-712.  Pre-exponential factors are converted
    from cgs/mole to SI/kmol and activation energies to activation
    temperatures [K].  Reactions with explicit REV parameters are split into
    two irreversible reactions (mech_interpret.py:673-693).  ``elem_wt`` is
    the atomic-weight table (a fresh copy of the built-in one by default);
    ELEM-section weight declarations update it.
    """
    if elem_wt is None:
        elem_wt = chem.get_elem_wt()
    elems, specs, reacs = [], [], []
    key = ''
    units_E = 'cal/mole'
    units_A = 'moles'
    thd = pdep = False
    cheb_flag = False

    with open(mech_filename, 'r') as fh:
        lines = fh.readlines()

    for raw in lines:
        line = raw
        if _skippable(line):
            continue
        bang = line.find('!')
        if bang > 0:
            line = line[:bang]

        head4 = line[0:4].lower()
        if head4 in ('elem', 'spec'):
            key = head4
            toks = line.split()
            if len(toks) > 1:
                line = line[line.index(toks[1]):]
            else:
                continue
        elif head4 == 'reac':
            key = 'reac'
            units_E = 'cal/mole'
            units_A = 'moles'
            for unit in line.split()[1:]:
                u = unit.lower()
                if u in pre_units:
                    units_A = u
                elif u in act_energy_units:
                    units_E = u
                else:
                    raise MechanismError(
                        'unsupported units on REACTION line: ' + unit)
            if units_A == 'molecules':
                raise NotImplementedError('Molecules units not supported.')
            continue
        elif head4 == 'ther':
            read_thermo(mech_filename, elems, specs, elem_wt)
            continue
        elif line[0:3].lower() == 'end':
            key = ''
            continue

        line = line.strip()

        if key == 'elem':
            last = ''
            for tok in line.replace('/', ' ').split():
                if tok.isalpha():
                    if tok[0:3] == 'end':
                        continue
                    if tok not in elems:
                        elems.append(tok)
                    last = tok
                else:
                    elem_wt[last.lower()] = float(tok)

        elif key == 'spec':
            for tok in line.split():
                if tok[0:3] == 'end':
                    continue
                if not any(sp.name == tok for sp in specs):
                    specs.append(chem.SpecInfo(tok))

        elif key == 'reac':
            if '=' in line:
                cheb_flag = False
                toks = line.split()
                A, b, E = (float(t) for t in toks[-3:])
                eqn = line[:line.index(toks[-3])].strip()
                if '<=>' in eqn:
                    i = eqn.index('<=>')
                    rev, lhs, rhs = True, eqn[:i], eqn[i + 3:]
                elif '=>' in eqn:
                    i = eqn.index('=>')
                    rev, lhs, rhs = False, eqn[:i], eqn[i + 2:]
                else:
                    i = eqn.index('=')
                    rev, lhs, rhs = True, eqn[:i], eqn[i + 1:]

                # each side is scanned for its own '(+M)' marker; the
                # product side's collider name wins when both give one
                lterms, lp, lsp, lthd = _split_side(lhs.strip())
                rterms, rp, rsp, rthd = _split_side(rhs.strip())
                pdep = lp or rp
                pdep_sp = rsp if rp else lsp
                reac_names, reac_nu, m_l = _merge(
                    lterms, lambda s: s.lower() == 'm')
                prod_names, prod_nu, m_r = _merge(
                    rterms, lambda s: s in ('m', 'M'))
                thd = lthd or rthd or m_l or m_r
                if pdep:
                    # falloff and plain third-body reactions are distinct
                    thd = False

                E *= act_energy_fact[units_E]
                if units_A == 'moles':
                    order = sum(reac_nu)
                    if thd:
                        A /= 1000. ** order
                    else:
                        A /= 1000. ** (order - 1.)

                rxn = chem.ReacInfo(rev, reac_names, reac_nu, prod_names,
                                    prod_nu, A, b, E)
                rxn.thd_body = thd
                rxn.pdep = pdep
                if pdep:
                    rxn.pdep_sp = pdep_sp
                reacs.append(rxn)
            else:
                if not reacs:
                    raise MechanismError('auxiliary line before any '
                                         'reaction: ' + line)
                last = reacs[-1]
                aux = line[0:3].lower()
                if aux == 'dup':
                    last.dup = True
                elif aux == 'rev':
                    p = _aux_numbers(line)
                    par1, par2, par3 = float(p[1]), float(p[2]), float(p[3])
                    par3 *= act_energy_fact[units_E]
                    if units_A == 'moles':
                        order = sum(last.prod_nu)
                        if thd:
                            par1 /= 1000. ** order
                        else:
                            par1 /= 1000. ** (order - 1.)
                    if par1 != 0.0:
                        last.rev_par.extend([par1, par2, par3])
                    else:
                        last.rev = False
                elif aux == 'low':
                    p = _aux_numbers(line)
                    par1, par2, par3 = float(p[1]), float(p[2]), float(p[3])
                    par3 *= act_energy_fact[units_E]
                    if units_A == 'moles':
                        par1 /= 1000. ** sum(last.reac_nu)
                    last.low.extend([par1, par2, par3])
                elif aux == 'hig':
                    p = _aux_numbers(line)
                    par1, par2, par3 = float(p[1]), float(p[2]), float(p[3])
                    par3 *= act_energy_fact[units_E]
                    if units_A == 'moles':
                        par1 /= 1000. ** (sum(last.reac_nu) - 2.)
                    last.high.extend([par1, par2, par3])
                elif aux == 'tro':
                    p = _aux_numbers(line)
                    last.troe = True
                    last.troe_par.extend(float(x) for x in p[1:4])
                    if len(p) > 4:
                        last.troe_par.append(float(p[4]))
                elif aux == 'sri':
                    p = _aux_numbers(line)
                    last.sri = True
                    last.sri_par.extend(float(x) for x in p[1:4])
                    if len(p) > 4:
                        last.sri_par.extend([float(p[4]), float(p[5])])
                elif aux == 'che':
                    last.cheb = True
                    p = line.replace('/', ' ').split()
                    if cheb_flag:
                        last.cheb_par.extend(float(x) for x in p[1:])
                    else:
                        cheb_flag = True
                        last.pdep = False
                        last.cheb_n_temp = int(p[1])
                        last.cheb_n_pres = int(p[2])
                        last.cheb_par = [float(x) for x in p[3:]]
                elif aux == 'pch':
                    p = line.replace('/', ' ').split()
                    last.cheb_plim = [float(p[1]) * chem.PA,
                                      float(p[2]) * chem.PA]
                    if len(p) > 3 and p[3].lower() == 'tcheb':
                        last.cheb_tlim = [float(p[4]), float(p[5])]
                elif aux == 'tch':
                    p = line.replace('/', ' ').split()
                    last.cheb_tlim = [float(p[1]), float(p[2])]
                    if len(p) > 3 and p[3].lower() == 'pcheb':
                        last.cheb_plim = [float(p[4]) * chem.PA,
                                          float(p[5]) * chem.PA]
                elif aux == 'plo':
                    p = line.replace('/', ' ').split()
                    if not last.plog:
                        last.plog = True
                        last.pdep = False
                        last.plog_par = []
                    pars = [float(x) for x in p[1:5]]
                    pars[0] *= 101325.0
                    pars[3] *= act_energy_fact[units_E]
                    if units_A == 'moles':
                        pars[1] /= 1000. ** (sum(last.reac_nu) - 1.)
                    last.plog_par.append(pars)
                else:
                    p = line.replace('/', ' ').split()
                    for i in range(0, len(p), 2):
                        last.thd_body_eff.append([p[i], float(p[i + 1])])

    for idx, rxn in enumerate(reacs):
        if rxn.cheb:
            n, m = rxn.cheb_n_temp, rxn.cheb_n_pres
            if len(rxn.cheb_par) != n * m:
                raise MechanismError('incorrect number of CHEB coefficients '
                                     'in reaction ' + repr(idx))
            if units_A == 'moles':
                order = sum(rxn.reac_nu)
                rxn.cheb_par[0] += math.log10(0.001 ** (order - 1.))
            rxn.cheb_par = np.reshape(rxn.cheb_par, (n, m))

    # split explicit-REV reactions into two irreversible ones
    out = []
    for rxn in reacs:
        if rxn.rev_par:
            back = copy.deepcopy(rxn)
            back.A, back.b, back.E = rxn.rev_par[0:3]
            back.rev = False
            back.rev_par = []
            back.reac, back.reac_nu = rxn.prod[:], rxn.prod_nu[:]
            back.prod, back.prod_nu = rxn.reac[:], rxn.reac_nu[:]
            rxn.rev = False
            rxn.rev_par = []
            out.extend([rxn, back])
        else:
            out.append(rxn)
    reacs[:] = out

    if any(not sp.mw for sp in specs):
        if therm_filename:
            read_thermo(therm_filename, elems, specs, elem_wt)
        else:
            raise MechanismError('no thermo file given and species are '
                                 'missing thermo data')
    missing = [sp.name for sp in specs if not sp.mw]
    if missing:
        raise MechanismError('missing thermo data for ' + ', '.join(missing))
    return elems, specs, reacs


def _fields(text, width):
    return [text[i:i + width] for i in range(0, len(text), width)]


def read_thermo(filename, elems, specs, elem_wt=None):
    """Fill NASA-7 data, element composition and MW of ``specs`` in place.

    Fixed-column reader after mech_interpret.py:715-863: name in columns
    0-17 (cut at the first blank), elements in 5-character fields of
    columns 24-43, temperatures in columns 45-72, then three lines of
    15-character coefficients (upper range a0..a6, then lower range
    a0..a6).  MW is the running sum of count x atomic weight.
    """
    if elem_wt is None:
        elem_wt = chem.get_elem_wt()
    with open(filename, 'r') as fh:
        lines = fh.readlines()
    pos = 0
    nlines = len(lines)
    while pos < nlines:
        line = lines[pos]
        pos += 1
        if _skippable(line):
            continue
        if 'thermo' in line.lower():
            break
    T_ranges = None
    if pos < nlines:
        toks = lines[pos].split()
        if toks and toks[0][0:1].isdigit():
            T_ranges = [float(t) for t in toks]
            pos += 1

    by_name = {}
    for i, sp in enumerate(specs):
        by_name.setdefault(sp.name, i)

    while pos < nlines:
        line = lines[pos]
        pos += 1
        if line[0:3].lower() == 'end':
            break
        if _skippable(line):
            continue
        name = line[0:18].strip()
        if name.find(' ') > 0:
            name = name[:name.find(' ')]
        idx = by_name.get(name)
        if idx is None or specs[idx].mw:
            pos += 3
            continue
        sp = specs[idx]
        for field in _fields(line[24:44], 5):
            e = field[0:2].strip()
            if e == '' or e == '0':
                continue
            count = int(float(field[2:].strip()))
            sp.elem.append([e, count])
            sp.mw += count * elem_wt[e.lower()]
        temps = [float(t) for t in line[45:73].split()]
        if len(temps) >= 3:
            t_mid = temps[2]
        elif T_ranges is not None:
            t_mid = T_ranges[1]
        else:
            raise ValueError('species %s gives no common temperature and '
                             'the file has no range line' % name)
        sp.Trange = [temps[0], t_mid, temps[1]]
        if pos + 3 > nlines:
            raise ValueError('truncated thermo entry for ' + name)
        c2 = _fields(lines[pos][0:75], 15)
        c3 = _fields(lines[pos + 1][0:75], 15)
        c4 = _fields(lines[pos + 2][0:75], 15)
        pos += 3
        sp.hi[0:5] = [float(c) for c in c2[0:5]]
        sp.hi[5], sp.hi[6] = float(c3[0]), float(c3[1])
        sp.lo[0:3] = [float(c) for c in c3[2:5]]
        sp.lo[3:7] = [float(c) for c in c4[0:4]]
        if all(s.mw != 0.0 for s in specs):
            break
    return None
