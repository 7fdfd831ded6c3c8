"""Chemkin-II mechanism and NASA-7 thermodynamic-data reader.

Produces the ``SpecInfo``/``ReacInfo`` records (core/chem_model.py) that the
table builders consume, with the values the reference's reader assigns
(pyjac/core/mech_interpret.py: ``read_mech`` :106-810, ``read_thermo``
:812-969): pre-exponential factors in kmol/m^3/s units, activation energies
as activation temperatures, LOW/HIGH/TROE/SRI/PLOG/CHEB parameters,
reactions with explicit reverse parameters split into two irreversible
reactions, the Chebyshev unit shift, and the bath gas moved last
(core/create_jacobian.py:5131-5209).

The reader is organised around the file's structure rather than line by
line: ``_scan`` classifies every line by section (ELEMENTS, SPECIES, THERMO,
REACTIONS) and hands the THERMO block to the fixed-column record reader;
reaction equations are split by ``_Equation`` with one scanner per side;
auxiliary lines are read as ``KEY`` or ``KEY / values /`` items and
dispatched through a table of keyword handlers.
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

# activation-energy units of the REACTIONS line -> factor to an activation
# temperature [K] (the gas constant in J/(mol K), chem_model.RU_JOUL)
ENERGY_UNITS = {
    'kelvins': 1.0,
    'evolts': 11595.,
    'cal/mole': 4.184 / chem.RU_JOUL,
    'kcal/mole': 4184. / chem.RU_JOUL,
    'joules/mole': 1. / chem.RU_JOUL,
    'kjoules/mole': 1000.0 / chem.RU_JOUL,
    'joules/kmole': 1. / (chem.RU_JOUL * 1000.),
}
# the name the Cantera YAML front end imports
_ACT_ENERGY_FACT = ENERGY_UNITS

_QUANTITY_UNITS = ('moles', 'molecules')

# auxiliary keywords of the Chemkin format this reader does not implement:
# refused with a message rather than read as third-body efficiencies
_UNSUPPORTED_AUX = ('FORD', 'RORD', 'UNITS', 'LT', 'RLT', 'JAN', 'FIT1', 'HV', 'TDEP',
                    'EXCI', 'MOME', 'XSMI', 'COLLEFF', 'MWON', 'MWOFF', 'USRPROG')

_SECTIONS = (('ELEM', 'elements'), ('SPEC', 'species'), ('THER', 'thermo'),
             ('REAC', 'reactions'), ('TRAN', 'transport'))

# KEY or KEY / values / on an auxiliary line
_ITEM = re.compile(r'([^\s/]+)\s*(?:/([^/]*)/)?')
# the three Arrhenius numbers closing a reaction line
_HEADER = re.compile(r'^(?P<eq>.*?\S)\s+(?P<A>\S+)\s+(?P<b>\S+)\s+(?P<E>\S+)\s*$')
# falloff marker (+M) or (+species) on one side of an equation
_FALLOFF = re.compile(r'\(\s*\+\s*([^()\s]+)\s*\)')
# a leading stoichiometric coefficient
_COEF = re.compile(r'^(\d+\.\d*|\.\d+|\d+)\s*(?=[A-Za-z(\[])')


def _number(tok):
    """A Chemkin number (Fortran D exponents accepted)."""
    try:
        return float(tok)
    except ValueError:
        try:
            return float(tok.replace('D', 'E').replace('d', 'e'))
        except ValueError:
            raise MechanismError('not a number: {!r}'.format(tok))


def _numbers(text):
    return [_number(t) for t in (text or '').replace(',', ' ').split()]


def _section_keyword(token):
    """Section a line's first token opens, or None."""
    t = token.upper()
    for prefix, name in _SECTIONS:
        if t.startswith(prefix):
            return name
    return None


# ------------------------------------------------------------------ thermo
class _ThermoRecord(object):
    """One NASA-7 record of the fixed-column format (four lines), read when
    a species of the mechanism takes it: composition fields in columns
    25-44, temperature range in 46-74, fifteen-column coefficients (high
    set, then low set) on lines 2-4."""

    def __init__(self, lines, T_common):
        self.lines = lines
        self.T_common = T_common
        self.name = lines[0][0:18].split()[0]

    def apply(self, sp, elem_wt):
        head, body = self.lines[0], self.lines[1:4]
        wt_key = {e.lower(): e for e in elem_wt}
        for fld in _fields(head[24:44], 5):
            sym = fld[0:2].strip()
            if sym in ('', '0'):
                continue
            key = wt_key.get(sym.lower())
            if key is None:
                raise MechanismError('species {}: unknown element {}'.format(sp.name, sym))
            count = int(float(fld[2:].strip()))
            sp.elem.append([key, count])
            sp.mw += count * elem_wt[key]
        temps = _numbers(head[45:74])
        if len(temps) == 3:
            T_mid = temps[2]
        elif self.T_common is not None:
            T_mid = self.T_common[1]
        else:
            raise MechanismError('thermo record {} has no common temperature'.format(sp.name))
        sp.Trange = [temps[0], T_mid, temps[1]]
        coef = [_number(c) for ln in body for c in _fields(ln[0:75], 15)[:5] if c.strip()]
        if len(coef) < 14:
            raise MechanismError('thermo record {} has {} coefficients'.format(sp.name, len(coef)))
        sp.hi[:] = coef[0:7]
        sp.lo[:] = coef[7:14]


def _fields(text, width):
    return [text[i:i + width] for i in range(0, len(text), width)]


def _thermo_records(lines, pos):
    """Records of a THERMO block starting at lines[pos] (after the THERMO
    keyword line); returns ({name: first record}, index after the block)."""
    def skip(p):
        while p < len(lines) and (not lines[p].strip() or lines[p].lstrip().startswith('!')):
            p += 1
        return p

    records = {}
    pos = skip(pos)
    T_common = None
    if pos < len(lines) and lines[pos].split()[0][:1].isdigit():
        T_common = _numbers(lines[pos])   # default T_lo T_mid T_hi
        pos += 1
    while True:
        pos = skip(pos)
        if pos >= len(lines) or lines[pos].split()[0].upper().startswith('END'):
            return records, pos + 1
        if pos + 3 >= len(lines):
            raise MechanismError('truncated thermo record: ' + lines[pos].strip())
        rec = _ThermoRecord(lines[pos:pos + 4], T_common)
        pos += 4
        records.setdefault(rec.name, rec)


def _read_thermo_file(path):
    """{name: record} of a thermo database (a THERMO block, with or without
    its keyword line)."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    for i, ln in enumerate(lines):
        tok = ln.split()
        if tok and not ln.lstrip().startswith('!') and tok[0].upper().startswith('THER'):
            return _thermo_records(lines, i + 1)[0]
    return _thermo_records(lines, 0)[0]


# -------------------------------------------------------------- equations
class _Side(object):
    """Species and coefficients of one side of an equation, with its
    third-body ('+ M') and falloff ('(+M)', '(+species)') markers."""

    def __init__(self, text):
        self.falloff = None       # None, '' for (+M), or a species name
        marks = _FALLOFF.findall(text)
        if len(marks) > 1:
            raise MechanismError('several falloff markers in ' + text)
        if marks:
            self.falloff = '' if marks[0].upper() == 'M' else marks[0]
            text = _FALLOFF.sub(' ', text, count=1)
        self.third_body = False
        self.names, self.nus = [], []
        for term in self._terms(text):
            m = _COEF.match(term)
            if m is None:
                nu, name = 1, term
            else:
                nu = float(m.group(1)) if '.' in m.group(1) else int(m.group(1))
                name = term[m.end():].strip()
            if name.upper() == 'M':
                self.third_body = True
            elif name in self.names:
                self.nus[self.names.index(name)] += nu
            else:
                self.names.append(name)
                self.nus.append(nu)

    @staticmethod
    def _terms(text):
        """Split at the '+' signs that separate terms: a '+' belongs to a
        species name (ions such as H3O+) when no term follows it."""
        terms, cur, i, s = [], '', 0, text.strip()
        while i < len(s):
            ch = s[i]
            if ch == '+' and cur.strip():
                rest = s[i + 1:].lstrip()
                if rest and rest[0] != '+':
                    terms.append(cur.strip())
                    cur = ''
                    i += 1
                    continue
            cur += ch
            i += 1
        if cur.strip():
            terms.append(cur.strip())
        return terms


class _Equation(object):
    """A reaction equation: reactant and product sides and the arrow."""

    def __init__(self, text):
        for arrow, rev in (('<=>', True), ('=>', False), ('=', True)):
            if arrow in text:
                left, right = text.split(arrow, 1)
                break
        else:
            raise MechanismError('no reaction arrow in ' + text)
        self.reversible = rev
        self.lhs, self.rhs = _Side(left), _Side(right)
        self.falloff = self.lhs.falloff is not None or self.rhs.falloff is not None
        # the product side's marker names the falloff collider when both have one
        self.collider = self.rhs.falloff if self.rhs.falloff is not None else \
            (self.lhs.falloff or '')
        self.third_body = (self.lhs.third_body or self.rhs.third_body) and not self.falloff


# ------------------------------------------------------------------ units
class _Units(object):
    """Units of one REACTIONS section and the conversions to the SI kmol
    units of the kernels."""

    def __init__(self, tokens):
        self.energy = 'cal/mole'
        self.quantity = 'moles'
        for tok in tokens:
            u = tok.lower()
            if u in _QUANTITY_UNITS:
                self.quantity = u
            elif u in ENERGY_UNITS:
                self.energy = u
            else:
                raise MechanismError('Unsupported units {} on REACTIONS line'.format(tok))
        if self.quantity == 'molecules':
            raise MechanismError('Molecules units not supported')

    def temperature(self, E):
        return E * ENERGY_UNITS[self.energy]

    def rate(self, A, order):
        """A of a rate constant of the given total order, cm/mol -> m/kmol."""
        return A / 1000. ** (order - 1.) if self.quantity == 'moles' else A


# ---------------------------------------------------- auxiliary keywords
def _aux_rev(rxn, vals, units, ctx):
    A, b, E = vals[0:3]
    order = sum(rxn.prod_nu) + (1 if ctx['thd'] else 0)
    A = units.rate(A, order)
    if A != 0.0:
        rxn.rev_par.extend([A, b, units.temperature(E)])
    else:
        rxn.rev = False


def _aux_low(rxn, vals, units, ctx):
    rxn.low.extend([units.rate(vals[0], sum(rxn.reac_nu) + 1), vals[1],
                    units.temperature(vals[2])])


def _aux_high(rxn, vals, units, ctx):
    rxn.high.extend([units.rate(vals[0], sum(rxn.reac_nu) - 1), vals[1],
                     units.temperature(vals[2])])


def _aux_troe(rxn, vals, units, ctx):
    rxn.troe = True
    a, T3, T1 = vals[0:3]
    rxn.troe_par.extend([a, T3 if T3 != 0 else 1e-30, T1 if T1 != 0 else 1e-30])
    if len(vals) > 3:
        rxn.troe_par.append(vals[3])


def _aux_sri(rxn, vals, units, ctx):
    rxn.sri = True
    rxn.sri_par.extend(vals[0:5] if len(vals) > 3 else vals[0:3])


def _aux_cheb(rxn, vals, units, ctx):
    if not rxn.cheb:
        # the first CHEB item of a reaction carries the basis sizes
        rxn.cheb = True
        rxn.pdep = False
        rxn.cheb_n_temp, rxn.cheb_n_pres = int(vals[0]), int(vals[1])
        rxn.cheb_par = list(vals[2:])
    else:
        rxn.cheb_par.extend(vals)


def _aux_pcheb(rxn, vals, units, ctx):
    rxn.cheb_plim = [vals[0] * chem.PA, vals[1] * chem.PA]


def _aux_tcheb(rxn, vals, units, ctx):
    rxn.cheb_tlim = [vals[0], vals[1]]


def _aux_plog(rxn, vals, units, ctx):
    if not rxn.plog:
        rxn.plog = True
        rxn.pdep = False
        rxn.plog_par = []
    P, A, b, E = vals[0:4]
    rxn.plog_par.append([P * chem.PA, units.rate(A, sum(rxn.reac_nu)), b,
                         units.temperature(E)])


_AUX = {'REV': _aux_rev, 'LOW': _aux_low, 'HIGH': _aux_high, 'TROE': _aux_troe,
        'SRI': _aux_sri, 'CHEB': _aux_cheb, 'PCHEB': _aux_pcheb, 'TCHEB': _aux_tcheb,
        'PLOG': _aux_plog}


def _auxiliary(line, rxn, units, ctx, species):
    for m in _ITEM.finditer(line):
        key, val = m.group(1), m.group(2)
        up = key.upper()
        if up.startswith('DUP'):
            rxn.dup = True
        elif up in _AUX:
            _AUX[up](rxn, _numbers(val), units, ctx)
        elif up in _UNSUPPORTED_AUX:
            raise MechanismError('auxiliary keyword {} is not supported: {}'.format(key, line))
        elif val is not None:
            rxn.thd_body_eff.append([key, _number(val.strip())])
        else:
            raise MechanismError('cannot read auxiliary item {!r} of {}'.format(key, line))


# ------------------------------------------------------------------- file
class _Mechanism(object):
    def __init__(self):
        self.elems, self.specs, self.reacs = [], [], []
        self.elem_wt = chem.get_elem_wt()
        self.thermo = {}
        self.units = _Units([])
        self.ctx = None   # state of the reaction the auxiliary lines belong to

    def add_elements(self, text):
        for m in _ITEM.finditer(text):
            name, wt = m.group(1), m.group(2)
            if name not in self.elems:
                self.elems.append(name)
            if wt is not None:
                self.elem_wt[name.lower()] = _number(wt.strip())

    def add_species(self, text):
        known = {sp.name for sp in self.specs}
        for name in text.split():
            if name not in known:
                self.specs.append(chem.SpecInfo(name))
                known.add(name)

    def add_reaction(self, m):
        eq = _Equation(m.group('eq'))
        A, b, E = (_number(m.group(k)) for k in ('A', 'b', 'E'))
        order = sum(eq.lhs.nus) + (1 if eq.third_body else 0)
        rxn = chem.ReacInfo(eq.reversible, eq.lhs.names, eq.lhs.nus, eq.rhs.names,
                            eq.rhs.nus, self.units.rate(A, order), b,
                            self.units.temperature(E))
        rxn.thd_body = eq.third_body
        rxn.pdep = eq.falloff
        if eq.falloff:
            rxn.pdep_sp = eq.collider
        self.reacs.append(rxn)
        self.ctx = {'thd': eq.third_body}


def _scan(path, mech):
    """Classify the file's lines by section and feed them to mech."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    section, pos = None, 0
    while pos < len(lines):
        text = lines[pos].split('!', 1)[0].strip()
        pos += 1
        if not text:
            continue
        first = text.split()[0]
        if first.upper() == 'END':
            section = None
            continue
        opened = _section_keyword(first)
        if opened is not None:
            section = opened
            rest = text[len(first):].strip()
            if opened == 'thermo':
                records, pos = _thermo_records(lines, pos)
                for name, rec in records.items():
                    mech.thermo.setdefault(name, rec)
                section = None
                continue
            if opened == 'reactions':
                mech.units = _Units(rest.split())
                continue
            text = rest
            if not text:
                continue
        # a trailing END closes the section on the same line
        toks = text.split()
        closes = bool(toks) and toks[-1].upper() == 'END'
        if closes:
            text = text[:text.rfind(toks[-1])].strip()
        if section == 'elements':
            mech.add_elements(text)
        elif section == 'species':
            mech.add_species(text)
        elif section == 'reactions':
            m = _HEADER.match(text) if '=' in text else None
            if m is not None:
                mech.add_reaction(m)
            elif mech.reacs:
                _auxiliary(text, mech.reacs[-1], mech.units, mech.ctx, mech.specs)
            else:
                raise MechanismError('auxiliary line before any reaction: ' + text)
        if closes:
            section = None


def _split_explicit_reverse(reacs):
    """A reaction with REV parameters becomes two irreversible reactions,
    the reverse one right after the forward one."""
    out = []
    for rxn in reacs:
        out.append(rxn)
        if not rxn.rev_par:
            continue
        back = copy.deepcopy(rxn)
        back.A, back.b, back.E = rxn.rev_par[0:3]
        back.reac, back.prod = list(rxn.prod), list(rxn.reac)
        back.reac_nu, back.prod_nu = list(rxn.prod_nu), list(rxn.reac_nu)
        for r in (rxn, back):
            r.rev = False
            r.rev_par = []
        out.append(back)
    return out


def read_mech(mech_filename, therm_filename=None, sort_type=None):
    """Read a Chemkin mechanism: ``(elems, specs, reacs)`` with species and
    reactions in file order (reactions optionally sorted), every record
    finalised."""
    mech = _Mechanism()
    _scan(mech_filename, mech)
    # Chebyshev tables: size check, unit shift of the leading coefficient
    # (log10 of the kmol conversion of the rate), [n_T][n_P] layout
    for idx, rxn in enumerate(mech.reacs):
        if not rxn.cheb:
            continue
        nT, nP = rxn.cheb_n_temp, rxn.cheb_n_pres
        if len(rxn.cheb_par) != nT * nP:
            raise MechanismError('Incorrect number of CHEB coefficients in reaction '
                                 '{}'.format(idx))
        if mech.units.quantity == 'moles':
            rxn.cheb_par[0] += math.log10(0.001 ** (sum(rxn.reac_nu) - 1.))
        rxn.cheb_par = np.reshape(rxn.cheb_par, (nT, nP))
    names = {sp.name for sp in mech.specs}
    for idx, rxn in enumerate(mech.reacs):
        unknown = sorted((set(rxn.reac) | set(rxn.prod)) - names)
        if unknown:
            raise MechanismError('Reaction {} contains unknown species {}'.format(
                idx, ', '.join(unknown)))
        bad = [e for e, _ in rxn.thd_body_eff if e not in names]
        if bad:
            raise MechanismError('Reaction {}: efficiency of unknown species {}'.format(
                idx, ', '.join(bad)))
    reacs = _split_explicit_reverse(mech.reacs)
    # thermodynamic data: the mechanism's THERMO block first, the database
    # file for the species it leaves out
    lacking = [sp for sp in mech.specs if sp.name not in mech.thermo]
    db = _read_thermo_file(therm_filename) if (lacking and therm_filename) else {}
    for sp in mech.specs:
        rec = mech.thermo.get(sp.name) or db.get(sp.name)
        if rec is not None:
            rec.apply(sp, mech.elem_wt)
    missing = [sp.name for sp in mech.specs if not sp.mw]
    if missing:
        if lacking and not therm_filename:
            raise MechanismError('No thermo data for some species and no thermo file given')
        raise MechanismError('Missing thermo data for ' + ', '.join(missing))
    for rxn in reacs:
        rxn.finalize(len(mech.specs))
    for sp in mech.specs:
        sp.finalize()
    if sort_type:
        reacs = sort_reactions(reacs, sort_type)
    return mech.elems, mech.specs, reacs


# -------------------------------------------------------- post-processing
def sort_reactions(reacs, sort_type, return_order=False):
    """Stable sort by (reaction type, falloff form, third-body form,
    reversibility), reference mech_interpret.py:58-103."""
    order = list(range(len(reacs)))
    if sort_type is not None and sort_type != reaction_sorting.none:
        enums = (reaction_type, falloff_form, thd_body_type, reversible_type)
        key = [tuple(int(r.get_type(e)[0]) for e in enums) for r in reacs]
        order.sort(key=lambda i: key[i])
    if return_order:
        return np.asarray(order, dtype=np.int64)
    return [reacs[i] for i in order]


def _bath_gas_index(specs, last_spec=None):
    """Index of the bath gas: the named species, else N2, Ar or He (name,
    case-insensitive, and exact molecular weight), else the last one."""
    if last_spec is not None:
        want = last_spec.strip().lower()
        for i, sp in enumerate(specs):
            if sp.name.lower() == want:
                return i
        logging.getLogger(__name__).warning('last species %s not found', last_spec)
    wt = chem.get_elem_wt()
    for name, mw in (('n2', 2. * wt['n']), ('ar', wt['ar']), ('he', wt['he'])):
        for i, sp in enumerate(specs):
            if sp.name.lower() == name and sp.mw == mw:
                return i
    return len(specs) - 1


def find_last_species(specs, last_spec=None, return_map=False):
    """Move the bath gas to the end of the species list (reference
    create_jacobian.py:5131-5209).  With return_map, the permutation instead
    (new position -> old index)."""
    idx = _bath_gas_index(specs, last_spec)
    perm = [i for i in range(len(specs)) if i != idx] + [idx]
    if return_map:
        return np.asarray(perm, dtype=np.int32)
    specs[:] = [specs[i] for i in perm]
    return specs


def reassign_species_lists(reacs, specs):
    """Species names in reactions -> indices, with reactants, products and
    third-body efficiencies ordered by species index (reference
    utils.py:613-645)."""
    index = {sp.name: i for i, sp in enumerate(specs)}

    def ordered(names, nus):
        pairs = sorted(((index[s], nu) for s, nu in zip(names, nus)), key=lambda p: p[0])
        return tuple(p[0] for p in pairs), tuple(p[1] for p in pairs)

    for rxn in reacs:
        rxn.reac, rxn.reac_nu = ordered(rxn.reac, rxn.reac_nu)
        rxn.prod, rxn.prod_nu = ordered(rxn.prod, rxn.prod_nu)
        rxn.thd_body_eff = sorted(((index[s], e) for s, e in rxn.thd_body_eff),
                                  key=lambda p: p[0])
        rxn.pdep_sp = index[rxn.pdep_sp] if rxn.pdep_sp != '' else None


def load_mechanism(mech_filename, therm_filename=None, last_spec=None, sort_type=None):
    """Read a mechanism (Cantera ``.cti`` / ``.yaml`` through their readers,
    anything else as Chemkin), move the bath gas last and index the species
    lists: the front-end sequence of reference create_jacobian.py:5434-5470."""
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
        raise MechanismError('No species found in ' + name)
    if not reacs:
        raise MechanismError('No reactions found in ' + name)
    specs = find_last_species(specs, last_spec=last_spec)
    reassign_species_lists(reacs, specs)
    return elems, specs, reacs
