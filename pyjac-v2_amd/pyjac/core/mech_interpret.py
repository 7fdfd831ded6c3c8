"""Chemkin-format mechanism and thermodynamic-database reader.

Restates the semantics of ``pyjac/core/mech_interpret.py`` (``read_mech``
:106-810, ``read_thermo`` :812-969, ``sort_reactions`` :58-103): unit
conversion of pre-exponential factors to kmol/m^3/s (:497-512), activation
energies to activation temperatures (:39-50), the LOW/HIGH/TROE/SRI/PLOG/CHEB
auxiliary lines, splitting of reactions with explicit reverse parameters into
two irreversible reactions (:759-776) and the Chebyshev log10 unit fix
(:735-746).  The resulting ``SpecInfo``/``ReacInfo`` lists are identical in
content and order to the reference's, which is what makes the downstream
tables (and hence the kernels' results) match.
"""

import copy
import logging
import math
import re

import numpy as np

from . import chem_model as chem
from .enum_types import reaction_type, falloff_form, thd_body_type, \
    reversible_type, reaction_sorting
from .exceptions import MechanismError

_PRE_UNITS = ('moles', 'molecules')
_ACT_ENERGY_FACT = {
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


def sort_reactions(reacs, sort_type, return_order=False):
    """Stable multi-key sort by (reaction type, falloff form, third-body
    form, reversibility), reference mech_interpret.py:58-103."""
    order = np.arange(len(reacs))
    if sort_type is not None and sort_type != reaction_sorting.none:
        enums = (reaction_type, falloff_form, thd_body_type, reversible_type)
        keys = np.array([[int(r.get_type(e)[0]) for e in enums]
                         for r in reacs], dtype=np.float64)
        for col in reversed(range(len(enums))):
            idx = keys[:, col].argsort(kind='mergesort')
            order = order[idx]
            keys = keys[idx]
    if return_order:
        return order
    return [reacs[i] for i in order]


def _read_str_num(string):
    return [float(n) for n in string.split()]


def _split_str(seq, length):
    return [seq[i:i + length] for i in range(0, len(seq), length)]


def _parse_side(side):
    """Split one side of a reaction equation into species / coefficients and
    detect ``+M`` and ``(+M)`` / ``(+species)`` (reference :307-490)."""
    thd = False
    pdep = False
    pdep_sp = ''
    sub = side
    while '(' in sub:
        i1 = sub.find('(')
        i2 = sub.find(')')
        inside = sub[i1 + 1:i2].strip()
        if inside == '+':
            sub = sub[i2 + 1:]
        elif inside[:1] == '+':
            pdep = True
            pdep_sp = sub[i1 + 1:i2].replace('+', ' ').strip()
            if pdep_sp.lower() == 'm':
                thd = True
                pdep_sp = ''
            base = side.find(sub)
            side = side[:i1 + base] + side[i2 + base + 1:]
            break
        else:
            sub = sub[i2 + 1:]

    items = side.split('+')
    while '' in items:
        i = items.index('')
        items[i - 1] = items[i - 1] + '+'
        del items[i]
    for sp in list(items):
        i = items.index(sp)
        if i < len(items) - 1:
            nxt = items[i + 1]
            if sp[-1:] == '(' and nxt[:1] == ')':
                items[i] = sp + '+' + nxt
                del items[i + 1]

    names, nus = [], []
    for sp in items:
        sp = sp.strip()
        if sp[0:1].isdigit():
            for j in range(len(sp)):
                if sp[j:j + 1].isalpha():
                    break
            coef = sp[0:j]
            nu = float(coef) if '.' in coef else int(coef)
            sp = sp[j:].strip()
        else:
            nu = 1
        if sp.lower() == 'm':
            thd = True
            continue
        if sp not in names:
            names.append(sp)
            nus.append(nu)
        else:
            nus[names.index(sp)] += nu
    return names, nus, thd, pdep, pdep_sp


def _aux_numbers(line):
    return line.replace('/', ' ').replace(',', ' ').split()


def read_mech(mech_filename, therm_filename=None, sort_type=None):
    """Read a Chemkin mechanism.

    Returns ``(elems, specs, reacs)`` with species/reactions in file order
    (reactions optionally sorted) and every reaction finalised.
    """
    elems, specs, reacs = [], [], []
    elem_wt = chem.get_elem_wt()
    key = ''
    units_E = 'cal/mole'
    units_A = 'moles'
    cheb_flag = False
    thd = pdep = False

    with open(mech_filename, 'r') as fh:
        lines = fh.readlines()

    for raw in lines:
        if _BLANK.search(raw) or _COMMENT.search(raw):
            continue
        line = raw.strip()
        ind = line.find('!')
        if ind > 0:
            line = line[:ind]
        head = line[0:4].lower()
        if head == 'elem':
            key = 'elem'
            parts = line.split()
            if len(parts) > 1:
                line = line[line.index(parts[1]):]
            else:
                continue
        elif head == 'spec':
            key = 'spec'
            parts = line.split()
            if len(parts) > 1:
                line = line[line.index(parts[1]):]
            else:
                continue
        elif head == 'reac':
            key = 'reac'
            units_E = 'cal/mole'
            units_A = 'moles'
            for unit in line.split()[1:]:
                u = unit.lower()
                if u in _PRE_UNITS:
                    units_A = u
                elif u in _ACT_ENERGY_FACT:
                    units_E = u
                else:
                    raise MechanismError(
                        'Unsupported units {} on REACTION line'.format(unit))
            if units_A == 'molecules':
                raise MechanismError('Molecules units not supported')
            continue
        elif head == 'ther':
            _read_thermo(mech_filename, elems, specs, elem_wt)
            continue
        elif line[0:3].lower() == 'end':
            key = ''
            continue

        if key == 'elem':
            last = ''
            for e in line.replace('/', ' ').split():
                if e.isalpha():
                    if e[0:3] == 'end':
                        continue
                    if e not in elems:
                        elems.append(e)
                    last = e
                else:
                    elem_wt[last.lower()] = float(e)
        elif key == 'spec':
            for s in line.split():
                if s[0:3] == 'end':
                    continue
                if not any(sp.name == s for sp in specs):
                    specs.append(chem.SpecInfo(s))
        elif key == 'reac':
            if '=' in line:
                cheb_flag = False
                parts = line.split()
                n = len(parts)
                A = float(parts[n - 3])
                b = float(parts[n - 2])
                E = float(parts[n - 1])
                eq = line[0:line.index(parts[n - 3])].strip()
                if '<=>' in eq:
                    i = eq.index('<=>')
                    rev, lhs, rhs = True, eq[:i].strip(), eq[i + 3:].strip()
                elif '=>' in eq:
                    i = eq.index('=>')
                    rev, lhs, rhs = False, eq[:i].strip(), eq[i + 2:].strip()
                else:
                    i = eq.index('=')
                    rev, lhs, rhs = True, eq[:i].strip(), eq[i + 1:].strip()
                r_names, r_nu, t1, p1, s1 = _parse_side(lhs)
                p_names, p_nu, t2, p2, s2 = _parse_side(rhs)
                thd = t1 or t2
                pdep = p1 or p2
                pdep_sp = s1 if p1 else (s2 if p2 else '')
                if p2 and not p1:
                    pdep_sp = s2
                elif p1 and p2:
                    pdep_sp = s2
                if pdep:
                    thd = False
                E *= _ACT_ENERGY_FACT[units_E]
                if units_A == 'moles':
                    order = sum(r_nu)
                    if thd:
                        A /= 1000. ** order
                    else:
                        A /= 1000. ** (order - 1.)
                rxn = chem.ReacInfo(rev, r_names, r_nu, p_names, p_nu, A, b, E)
                rxn.thd_body = thd
                rxn.pdep = pdep
                if pdep:
                    rxn.pdep_sp = pdep_sp
                reacs.append(rxn)
            else:
                _read_aux(line, reacs[-1], units_A, units_E, thd, pdep,
                          cheb_flag)
                if line[0:3].lower() == 'che':
                    cheb_flag = True

    for idx, rxn in enumerate(reacs):
        if rxn.cheb:
            nT, nP = rxn.cheb_n_temp, rxn.cheb_n_pres
            if len(rxn.cheb_par) != nT * nP:
                raise MechanismError(
                    'Incorrect number of CHEB coefficients in reaction '
                    '{}'.format(idx))
            if units_A == 'moles':
                rxn.cheb_par[0] += math.log10(
                    0.001 ** (sum(rxn.reac_nu) - 1.))
            rxn.cheb_par = np.reshape(rxn.cheb_par, (nT, nP))

    names = set(sp.name for sp in specs)
    for idx, rxn in enumerate(reacs):
        for s in set(list(rxn.reac) + list(rxn.prod)):
            if s not in names:
                raise MechanismError(
                    'Reaction {} contains unknown species {}'.format(idx, s))

    # explicit reverse parameters -> two irreversible reactions
    for rxn in list(reacs):
        if rxn.rev_par:
            new = copy.deepcopy(rxn)
            idx = reacs.index(rxn)
            reacs[idx].rev = False
            reacs[idx].rev_par = []
            new.A, new.b, new.E = new.rev_par[0], new.rev_par[1], \
                new.rev_par[2]
            new.rev = False
            new.rev_par = []
            new.prod = list(rxn.reac)
            new.prod_nu = list(rxn.reac_nu)
            new.reac = list(rxn.prod)
            new.reac_nu = list(rxn.prod_nu)
            reacs.insert(idx + 1, new)

    if any(not sp.mw for sp in specs):
        if not therm_filename:
            raise MechanismError('No thermo data for some species and no '
                                 'thermo file given')
        _read_thermo(therm_filename, elems, specs, elem_wt)
    missing = [sp.name for sp in specs if not sp.mw]
    if missing:
        raise MechanismError('Missing thermo data for ' + ', '.join(missing))

    for rxn in reacs:
        rxn.finalize(len(specs))
    for sp in specs:
        sp.finalize()
    if sort_type:
        reacs = sort_reactions(reacs, sort_type)
    return elems, specs, reacs


def _read_aux(line, rxn, units_A, units_E, thd, pdep, cheb_flag):
    """One auxiliary line (DUP/REV/LOW/HIGH/TROE/SRI/CHEB/PCHEB/TCHEB/PLOG or
    third-body efficiencies), reference mech_interpret.py:520-706."""
    aux = line[0:3].lower()
    if aux == 'dup':
        rxn.dup = True
    elif aux == 'rev':
        p = _aux_numbers(line)
        a1, a2, a3 = float(p[1]), float(p[2]), float(p[3])
        a3 *= _ACT_ENERGY_FACT[units_E]
        if units_A == 'moles':
            order = sum(rxn.prod_nu)
            if thd:
                a1 /= 1000. ** order
            else:
                a1 /= 1000. ** (order - 1.)
        if a1 != 0.0:
            rxn.rev_par.extend([a1, a2, a3])
        else:
            rxn.rev = False
    elif aux == 'low':
        p = _aux_numbers(line)
        a1, a2, a3 = float(p[1]), float(p[2]), float(p[3])
        a3 *= _ACT_ENERGY_FACT[units_E]
        if units_A == 'moles':
            a1 /= 1000. ** sum(rxn.reac_nu)
        rxn.low.extend([a1, a2, a3])
    elif aux == 'hig':
        p = _aux_numbers(line)
        a1, a2, a3 = float(p[1]), float(p[2]), float(p[3])
        a3 *= _ACT_ENERGY_FACT[units_E]
        if units_A == 'moles':
            a1 /= 1000. ** (sum(rxn.reac_nu) - 2.)
        rxn.high.extend([a1, a2, a3])
    elif aux == 'tro':
        p = _aux_numbers(line)
        rxn.troe = True
        a1, a2, a3 = float(p[1]), float(p[2]), float(p[3])
        if a2 == 0:
            a2 = 1e-30
        if a3 == 0:
            a3 = 1e-30
        rxn.troe_par.extend([a1, a2, a3])
        if len(p) > 4:
            rxn.troe_par.append(float(p[4]))
    elif aux == 'sri':
        p = _aux_numbers(line)
        rxn.sri = True
        rxn.sri_par.extend([float(p[1]), float(p[2]), float(p[3])])
        if len(p) > 4:
            rxn.sri_par.extend([float(p[4]), float(p[5])])
    elif aux == 'che':
        rxn.cheb = True
        p = line.replace('/', ' ').split()
        if cheb_flag:
            rxn.cheb_par.extend(float(x) for x in p[1:])
        else:
            rxn.pdep = False
            rxn.cheb_n_temp = int(p[1])
            rxn.cheb_n_pres = int(p[2])
            rxn.cheb_par = [float(x) for x in p[3:]]
    elif aux == 'pch':
        p = line.replace('/', ' ').split()
        rxn.cheb_plim = [float(p[1]) * chem.PA, float(p[2]) * chem.PA]
        if len(p) > 3 and p[3].lower() == 'tcheb':
            rxn.cheb_tlim = [float(p[4]), float(p[5])]
    elif aux == 'tch':
        p = line.replace('/', ' ').split()
        rxn.cheb_tlim = [float(p[1]), float(p[2])]
        if len(p) > 3 and p[3].lower() == 'pcheb':
            rxn.cheb_plim = [float(p[4]) * chem.PA, float(p[5]) * chem.PA]
    elif aux == 'plo':
        p = line.replace('/', ' ').split()
        if not rxn.plog:
            rxn.plog = True
            rxn.pdep = False
            rxn.plog_par = []
        pars = [float(x) for x in p[1:5]]
        pars[0] *= 101325.0
        pars[3] *= _ACT_ENERGY_FACT[units_E]
        if units_A == 'moles':
            pars[1] /= 1000. ** (sum(rxn.reac_nu) - 1.)
        rxn.plog_par.append(pars)
    else:
        p = line.replace('/', ' ').split()
        for i in range(0, len(p), 2):
            rxn.thd_body_eff.append([p[i], float(p[i + 1])])


def _read_thermo(filename, elems, specs, elem_wt):
    """NASA-7 fixed-column thermo reader (reference :812-969)."""
    name_map = {e.lower(): e for e in elem_wt}
    with open(filename, 'r') as fh:
        lines = fh.readlines()
    pos = 0
    while pos < len(lines):
        line = lines[pos]
        pos += 1
        if _BLANK.search(line) or _COMMENT.search(line):
            continue
        if 'thermo' in line.lower():
            break
    T_ranges = None
    if pos < len(lines):
        parts = lines[pos].split()
        if parts and parts[0][0:1].isdigit():
            T_ranges = _read_str_num(lines[pos])
            pos += 1
    by_name = {sp.name: sp for sp in specs}
    while pos < len(lines):
        line = lines[pos]
        pos += 1
        if line[0:3].lower() == 'end':
            break
        if _BLANK.search(line) or _COMMENT.search(line):
            continue
        name = line[0:18].strip()
        if name.find(' ') > 0:
            name = name[0:name.find(' ')]
        sp = by_name.get(name)
        if sp is None or sp.mw:
            pos += 3
            continue
        for e_str in _split_str(line[24:44], 5):
            e = e_str[0:2].strip()
            if e == '' or e == '0':
                continue
            num = int(float(e_str[2:].strip()))
            sp.elem.append([name_map[e.lower()], num])
            sp.mw += num * elem_wt[e.lower()]
        T_spec = _read_str_num(line[45:74])
        T_com = T_spec[2] if len(T_spec) == 3 else T_ranges[1]
        sp.Trange = [T_spec[0], T_com, T_spec[1]]
        c2 = _split_str(lines[pos][0:75], 15)
        c3 = _split_str(lines[pos + 1][0:75], 15)
        c4 = _split_str(lines[pos + 2][0:75], 15)
        pos += 3
        sp.hi[0:5] = [float(c) for c in c2[0:5]]
        sp.hi[5] = float(c3[0])
        sp.hi[6] = float(c3[1])
        sp.lo[0:3] = [float(c) for c in c3[2:5]]
        sp.lo[3:7] = [float(c) for c in c4[0:4]]
        if not any(s.mw == 0.0 for s in specs):
            break


def find_last_species(specs, last_spec=None, return_map=False):
    """Move the bath gas to the end of the species list: the user's choice
    if found, else the first of N2, Ar, He matched by name (case-insensitive)
    and exact molecular weight, else the current last species (reference
    create_jacobian.py:5131-5209)."""
    log = logging.getLogger(__name__)
    idx = None
    if last_spec is not None:
        idx = next((i for i, sp in enumerate(specs)
                    if sp.name.lower() == last_spec.lower().strip()), None)
        if idx is None:
            log.warning('last species %s not found', last_spec)
    if idx is None:
        wt = chem.get_elem_wt()
        for name, mw in (('N2', wt['n'] * 2.), ('Ar', wt['ar']),
                         ('He', wt['he'])):
            idx = next((i for i, sp in enumerate(specs)
                        if sp.name.lower() == name.lower() and sp.mw == mw),
                       None)
            if idx is not None:
                break
    if idx is None:
        idx = len(specs) - 1
    if return_map:
        gmap = np.arange(len(specs), dtype=np.int32)
        gmap[idx:-1] = gmap[idx + 1:]
        gmap[-1] = idx
        return gmap
    specs.append(specs.pop(idx))
    return specs


def reassign_species_lists(reacs, specs):
    """Replace species names in reactions by indices, sorting reactants,
    products and third-body efficiencies by species index (reference
    utils.py:613-645)."""
    smap = {sp.name: i for i, sp in enumerate(specs)}
    for rxn in reacs:
        pairs = sorted(zip(rxn.reac, rxn.reac_nu), key=lambda x: smap[x[0]])
        rxn.reac = tuple(smap[s] for s, _ in pairs)
        rxn.reac_nu = tuple(nu for _, nu in pairs)
        pairs = sorted(zip(rxn.prod, rxn.prod_nu), key=lambda x: smap[x[0]])
        rxn.prod = tuple(smap[s] for s, _ in pairs)
        rxn.prod_nu = tuple(nu for _, nu in pairs)
        rxn.thd_body_eff = sorted([(smap[t[0]], t[1])
                                   for t in rxn.thd_body_eff],
                                  key=lambda x: x[0])
        rxn.pdep_sp = smap[rxn.pdep_sp] if rxn.pdep_sp != '' else None


def load_mechanism(mech_filename, therm_filename=None, last_spec=None,
                   sort_type=None):
    """Parse, move the bath gas last and index the species lists; the same
    front-end sequence as reference create_jacobian.py:5434-5470: Cantera
    ``.cti`` and ``.yaml`` files go through their readers, anything else is
    Chemkin."""
    name = str(mech_filename)
    if name.endswith('.xml'):
        raise MechanismError('Unsupported mechanism format (Chemkin, Cantera CTI or YAML): '
                             + name)
    if name.endswith('.cti'):
        from .cti import read_mech_ct
        elems, specs, reacs = read_mech_ct(name, sort_type)
    elif name.endswith(('.yaml', '.yml')):
        from .cantera_yaml import read_mech_yaml
        elems, specs, reacs = read_mech_yaml(name, sort_type)
    else:
        elems, specs, reacs = read_mech(mech_filename, therm_filename, sort_type)
    if not specs:
        raise MechanismError('No species found in ' + mech_filename)
    if not reacs:
        raise MechanismError('No reactions found in ' + mech_filename)
    specs = find_last_species(specs, last_spec=last_spec)
    reassign_species_lists(reacs, specs)
    return elems, specs, reacs
