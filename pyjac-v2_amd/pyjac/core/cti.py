"""Cantera CTI mechanisms without Cantera.

The reference reads ``.cti`` inputs through Cantera (``read_mech_ct``,
core/mech_interpret.py:972-1241, dispatched at core/create_jacobian.py:5434).
Cantera is not a dependency here: a CTI file is parsed into its syntax tree
-- never executed -- and its directives are evaluated from literals only
(numbers, strings, lists, tuples, simple arithmetic, ``OneAtm`` and the
constructor calls ``NASA``, ``Troe``, ``SRI``, ``Arrhenius``, ``state``,
``gas_transport`` ..., which become plain records).  The mechanism is then
written as Chemkin text in Chemkin's units (cm, mol, the file's energy unit)
and read by the Chemkin reader, so both inputs produce their tables through
the same code.

Supported: ``units``; the first ``ideal_gas`` phase (its elements and
species); ``species`` with two ``NASA`` ranges; ``reaction``,
``three_body_reaction``, ``falloff_reaction`` (Lindemann, ``Troe``, ``SRI``),
``chemically_activated_reaction``, ``pdep_arrhenius`` (PLOG) and
``chebyshev_reaction``; ``efficiencies``, ``default_efficiency`` and
``options='duplicate'``.  As in the reference's Cantera path, elementary
reactions with a zero pre-exponential factor are dropped and a single
collider of efficiency one with a zero default efficiency becomes the
falloff's named bath species.

Molecular weights follow from the element table of the Chemkin reader; the
reference's Cantera path takes Cantera's own weights, which can differ in the
last digits.
"""

import ast
import os
import re
import tempfile

import numpy as np

from .exceptions import MechanismError

ONE_ATM = 101325.0
AVOGADRO = 6.02214076e23

# CTI units -> factors to Chemkin's cm and mol, and the Chemkin energy keyword
_LENGTH = {'cm': 1.0, 'm': 100.0, 'mm': 0.1}
_QUANTITY = {'mol': 1.0, 'kmol': 1000.0, 'molec': 1.0 / AVOGADRO}
_ENERGY = {'cal/mol': 'CAL/MOLE', 'kcal/mol': 'KCAL/MOLE', 'J/mol': 'JOULES/MOLE',
           'kJ/mol': 'KJOULES/MOLE', 'J/kmol': 'JOULES/KMOLE', 'K': 'KELVINS',
           'eV': 'EVOLTS'}
# CTI pressure units -> atm (Chemkin's PLOG and Chebyshev pressure unit)
_PRESSURE = {'Pa': 1.0 / ONE_ATM, 'atm': 1.0, 'bar': 1e5 / ONE_ATM,
             'dyn/cm2': 0.1 / ONE_ATM}

_CONSTANTS = {'OneAtm': ONE_ATM, 'True': True, 'False': False, 'None': None}


class Call(object):
    """A constructor call of the file, as a record."""

    def __init__(self, name, args, kwargs):
        self.name, self.args, self.kwargs = name, args, kwargs

    def arg(self, i, key, default=None):
        """Argument ``key``, given by keyword or at position ``i``."""
        if key in self.kwargs:
            return self.kwargs[key]
        return self.args[i] if len(self.args) > i else default

    def __repr__(self):
        return 'Call({}, {}, {})'.format(self.name, self.args, self.kwargs)


def _value(node):
    """Literal value of an expression node; constructor calls become Call
    records; anything else is refused."""
    if isinstance(node, ast.Constant):
        return node.value
    if isinstance(node, (ast.List, ast.Tuple)):
        return [_value(e) for e in node.elts]
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
        v = _value(node.operand)
        return -v if isinstance(node.op, ast.USub) else v
    if isinstance(node, ast.BinOp) and isinstance(node.op, (ast.Mult, ast.Div, ast.Add,
                                                            ast.Sub, ast.Pow)):
        a, b = _value(node.left), _value(node.right)
        if not all(isinstance(x, (int, float)) for x in (a, b)):
            raise MechanismError('CTI: arithmetic on non-numbers at line {}'.format(node.lineno))
        if isinstance(node.op, ast.Mult):
            return a * b
        if isinstance(node.op, ast.Div):
            return a / b
        if isinstance(node.op, ast.Add):
            return a + b
        if isinstance(node.op, ast.Sub):
            return a - b
        return a ** b
    if isinstance(node, ast.Name) and node.id in _CONSTANTS:
        return _CONSTANTS[node.id]
    if isinstance(node, ast.Call) and isinstance(node.func, ast.Name):
        return Call(node.func.id, [_value(a) for a in node.args],
                    {k.arg: _value(k.value) for k in node.keywords})
    raise MechanismError('CTI: unsupported expression at line {}'.format(
        getattr(node, 'lineno', '?')))


def parse_cti(text):
    """Top-level directives of a CTI file, in file order (Call records)."""
    try:
        tree = ast.parse(text)
    except SyntaxError as e:
        raise MechanismError('CTI: syntax error: {}'.format(e))
    out = []
    for stmt in tree.body:
        if isinstance(stmt, ast.Expr) and isinstance(stmt.value, ast.Call):
            out.append(_value(stmt.value))
        elif isinstance(stmt, ast.Expr) and isinstance(stmt.value, ast.Constant):
            continue  # a docstring
        else:
            raise MechanismError('CTI: unsupported statement at line {}'.format(stmt.lineno))
    return out


class _Units(object):
    """The file's ``units()`` directive (Cantera's defaults are SI)."""

    def __init__(self, length='m', quantity='kmol', act_energy='J/kmol', time='s',
                 pressure='Pa', **_):
        for v, table in ((length, _LENGTH), (quantity, _QUANTITY), (act_energy, _ENERGY),
                         (pressure, _PRESSURE)):
            if v not in table:
                raise MechanismError('CTI: unsupported unit {}'.format(v))
        if time != 's':
            raise MechanismError('CTI: unsupported time unit {}'.format(time))
        self.energy = act_energy
        self.pressure = pressure
        # (length^3 / quantity) of the file -> cm^3 / mol
        self.conc = _LENGTH[length] ** 3 / _QUANTITY[quantity]

    def atm(self, p):
        """A pressure, given bare in the file's unit or as (value, unit)."""
        if isinstance(p, list):
            if p[1] not in _PRESSURE:
                raise MechanismError('CTI: unsupported pressure unit {}'.format(p[1]))
            return float(p[0]) * _PRESSURE[p[1]]
        return float(p) * _PRESSURE[self.pressure]


def _arrhenius(k):
    """(A, b, E) of a rate given as [A, b, E] or Arrhenius(...)."""
    if isinstance(k, Call):
        if k.name != 'Arrhenius':
            raise MechanismError('CTI: unsupported rate {}'.format(k.name))
        vals = (k.arg(0, 'A'), k.arg(1, 'b', 0.0), k.arg(2, 'E', 0.0))
    else:
        vals = tuple(k[:3]) + (0.0,) * (3 - len(k[:3]))
    if any(isinstance(v, list) for v in vals):
        raise MechanismError('CTI: rate parameters with explicit units are not supported')
    return tuple(float(v) for v in vals)


def _fmt_rate(units, k, conc_power):
    """'A b E' in Chemkin units for a rate whose A carries the given power of
    (cm^3/mol)."""
    A, b, E = _arrhenius(k)
    return '{!r} {!r} {!r}'.format(A * units.conc ** conc_power, b, E)


def _split_equation(eq):
    for arrow in ('<=>', '=>', '='):
        if arrow in eq:
            lhs, rhs = eq.split(arrow, 1)
            return lhs, arrow, rhs
    raise MechanismError('CTI: no arrow in reaction {}'.format(eq))


_PDEP = re.compile(r'\(\s*\+\s*([^)\s]+)\s*\)')


def _order(side):
    """Sum of the stoichiometric coefficients of one side, without colliders."""
    side = _PDEP.sub(' ', side)
    total = 0.0
    for term in side.split('+'):
        t = term.split()
        if not t or t == ['M']:
            continue
        total += float(t[0]) if len(t) > 1 else 1.0
    return total


def _chemkin_equation(eq, bath=None):
    """'2 OH (+ M) <=> H2O2 (+ M)' -> '2OH(+M)<=>H2O2(+M)'; with ``bath``, the
    collider M of a falloff becomes that species."""
    eq = _PDEP.sub(lambda m: '(+{})'.format(bath if bath and m.group(1) == 'M'
                                            else m.group(1)), eq)
    return ''.join(eq.split())


def _thermo_entry(name, atoms, nasa):
    """Chemkin NASA-7 entry (four fixed-width lines) of one species."""
    if len(atoms) > 4:
        raise MechanismError('CTI: species {} has more than four elements'.format(name))
    ranges = sorted(nasa, key=lambda c: float(c.arg(0, 'Trange')[0]))
    (Tlo, Tmid), (Tmid2, Thi) = ([float(t) for t in c.arg(0, 'Trange')] for c in ranges)
    if abs(Tmid - Tmid2) > 1e-9 * Tmid:
        raise MechanismError('CTI: species {}: NASA ranges do not meet'.format(name))
    clo, chi = ([float(v) for v in c.arg(1, 'coeffs')] for c in ranges)
    if len(clo) != 7 or len(chi) != 7:
        raise MechanismError('CTI: species {}: NASA ranges need seven coefficients'.format(name))
    elems = ''.join('{:<2}{:>3d}'.format(e, n) for e, n in atoms)
    num = lambda vals: ''.join('{:15.8E}'.format(v) for v in vals)
    return '\n'.join((
        '{:<18}{:<6}{:<20}G{:10.3f}{:10.3f}{:8.2f}'.format(name, 'CTI', elems, Tlo, Thi,
                                                          Tmid).ljust(79) + '1',
        num(chi[:5]).ljust(79) + '2',
        num([chi[5], chi[6], clo[0], clo[1], clo[2]]).ljust(79) + '3',
        num(clo[3:7]).ljust(79) + '4'))


def _atoms(spec):
    out = []
    for tok in str(spec).split():
        e, n = tok.split(':')
        count = float(n)
        if count != int(count):
            raise MechanismError('CTI: fractional atom count in {}'.format(spec))
        out.append((e.upper(), int(count)))
    return out


def _efficiency_line(call, species):
    """Chemkin efficiencies of a collider reaction; a default efficiency other
    than one is written out for every unlisted species (mech_interpret.py:
    1066-1083 does the same on Cantera's reaction)."""
    eff = [tok.split(':') for tok in str(call.kwargs.get('efficiencies') or '').split()]
    default = float(call.kwargs.get('default_efficiency', 1.0))
    listed = set(s for s, _ in eff)
    pairs = [(s, float(v)) for s, v in eff]
    if default != 1.0:
        pairs += [(s, default) for s in species if s not in listed]
    return ' '.join('{}/{!r}/'.format(s, v) for s, v in pairs)


def _falloff_lines(fo):
    if fo is None or (isinstance(fo, Call) and fo.name == 'Lindemann'):
        return []
    if isinstance(fo, Call) and fo.name == 'Troe':
        pars = [fo.arg(0, 'A'), fo.arg(1, 'T3'), fo.arg(2, 'T1'), fo.arg(3, 'T2')]
        pars = pars[:3] if pars[3] is None else pars
        return ['TROE / ' + ' '.join(repr(float(p)) for p in pars) + ' /']
    if isinstance(fo, Call) and fo.name == 'SRI':
        pars = [fo.arg(0, 'A'), fo.arg(1, 'B'), fo.arg(2, 'C')]
        D, E = fo.arg(3, 'D'), fo.arg(4, 'E')
        if D is not None or E is not None:
            pars += [1.0 if D is None else D, 0.0 if E is None else E]
        return ['SRI / ' + ' '.join(repr(float(p)) for p in pars) + ' /']
    raise MechanismError('CTI: unsupported falloff form {}'.format(fo))


def _reaction_lines(c, units, species):
    """Chemkin lines of one reaction directive (empty when it is dropped)."""
    eq = c.arg(0, 'equation')
    lhs, _, _ = _split_equation(eq)
    n = _order(lhs)
    lines = []
    if c.name == 'reaction':
        k = c.arg(1, 'kf')
        if _arrhenius(k)[0] == 0.0:
            return []
        lines.append('{} {}'.format(_chemkin_equation(eq), _fmt_rate(units, k, n - 1)))
    elif c.name == 'three_body_reaction':
        # the collider counts in the rate's order
        lines.append('{} {}'.format(_chemkin_equation(eq), _fmt_rate(units, c.arg(1, 'kf'), n)))
        eff = _efficiency_line(c, species)
        if eff:
            lines.append(eff)
    elif c.name in ('falloff_reaction', 'chemically_activated_reaction'):
        # a single named collider: Cantera's zero default efficiency with one
        # efficiency of one, the "(+SP)" form of a Chemkin equation
        eff_toks = str(c.kwargs.get('efficiencies') or '').split()
        bath = None
        if float(c.kwargs.get('default_efficiency', 1.0)) == 0.0 and len(eff_toks) == 1 \
                and float(eff_toks[0].split(':')[1]) == 1.0:
            bath = eff_toks[0].split(':')[0]
        if c.name == 'falloff_reaction':
            # k_inf on the reaction line, k_0 (one more collider) on LOW
            main, aux, key = c.arg(1, 'kf'), c.arg(2, 'kf0'), 'LOW'
            main_pow, aux_pow = n - 1, n
        else:
            # k_0 on the reaction line, k_inf (one collider fewer) on HIGH
            main, aux, key = c.arg(1, 'kLow'), c.arg(2, 'kHigh'), 'HIGH'
            main_pow, aux_pow = n - 1, n - 2
        lines.append('{} {}'.format(_chemkin_equation(eq, bath), _fmt_rate(units, main, main_pow)))
        lines.append('{} / {} /'.format(key, _fmt_rate(units, aux, aux_pow)))
        lines += _falloff_lines(c.kwargs.get('falloff'))
        if bath is None:
            eff = _efficiency_line(c, species)
            if eff:
                lines.append(eff)
    elif c.name == 'pdep_arrhenius':
        rates = c.args[1:]
        if not rates:
            raise MechanismError('CTI: pdep_arrhenius without rates: {}'.format(eq))
        # the reaction line is a placeholder: PLOG replaces it
        lines.append('{} {}'.format(_chemkin_equation(eq), _fmt_rate(units, rates[0][1:], n - 1)))
        for r in rates:
            lines.append('PLOG / {!r} {} /'.format(units.atm(r[0]),
                                                      _fmt_rate(units, r[1:], n - 1)))
    elif c.name == 'chebyshev_reaction':
        coef = np.array(c.kwargs['coeffs'], dtype=np.float64)
        if coef.ndim != 2:
            raise MechanismError('CTI: Chebyshev coefficients must be a matrix: {}'.format(eq))
        # log10 k in the file's concentration units -> cm^3/mol: the (0, 0)
        # basis function is one everywhere
        coef[0, 0] += (n - 1) * np.log10(units.conc)
        lines.append('{} 1.0 0.0 0.0'.format(_chemkin_equation(eq)))
        lines.append('TCHEB / {!r} {!r} /'.format(float(c.kwargs.get('Tmin', 300.0)),
                                                       float(c.kwargs.get('Tmax', 2500.0))))
        lines.append('PCHEB / {!r} {!r} /'.format(units.atm(c.kwargs['Pmin']),
                                                       units.atm(c.kwargs['Pmax'])))
        flat = ['{:d} {:d}'.format(*coef.shape)] + [repr(float(v)) for v in coef.reshape(-1)]
        for i in range(0, len(flat), 4):
            lines.append('CHEB / ' + ' '.join(flat[i:i + 4]) + ' /')
    else:
        raise MechanismError('CTI: unsupported reaction type {}'.format(c.name))
    opts = c.kwargs.get('options', [])
    opts = [opts] if isinstance(opts, str) else list(opts)
    if 'duplicate' in opts:
        lines.append('DUPLICATE')
    return lines


_REACTIONS = ('reaction', 'three_body_reaction', 'falloff_reaction',
              'chemically_activated_reaction', 'pdep_arrhenius', 'chebyshev_reaction')


def _convert(text):
    """(Chemkin text, {species: (lo, hi) NASA coefficients}) of the first
    phase of a CTI file."""
    return convert_directives(parse_cti(text))


def convert_directives(calls):
    """(Chemkin text, {species: (lo, hi) NASA coefficients}) of a
    mechanism given as CTI directives (Call records, in file order)."""
    units = _Units()
    phase = None
    species = {}
    reactions = []
    for c in calls:
        if c.name == 'units':
            units = _Units(**c.kwargs)
        elif c.name == 'ideal_gas':
            phase = phase or c
        elif c.name == 'species':
            name = c.arg(0, 'name')
            thermo = c.arg(2, 'thermo')
            nasa = thermo if isinstance(thermo, list) else [thermo]
            if len(nasa) != 2 or any(not isinstance(t, Call) or t.name != 'NASA' for t in nasa):
                raise MechanismError('CTI: species {} needs two NASA ranges'.format(name))
            species[name] = (_atoms(c.arg(1, 'atoms')), nasa)
        elif c.name in _REACTIONS:
            reactions.append(c)
        elif c.name.endswith('reaction'):
            raise MechanismError('CTI: unsupported reaction type {}'.format(c.name))
    if phase is None:
        raise MechanismError('CTI: no ideal_gas phase')
    if str(phase.kwargs.get('reactions', 'all')).strip() not in ('all', ''):
        raise MechanismError('CTI: only reactions="all" is supported')
    elements = str(phase.arg(1, 'elements')).split()
    names = str(phase.arg(2, 'species')).split()
    missing = [s for s in names if s not in species]
    if missing:
        raise MechanismError('CTI: no species entry for ' + ', '.join(missing))
    out = ['ELEMENTS', ' '.join(e.upper() for e in elements), 'END', 'SPECIES']
    out += [' '.join(names[i:i + 6]) for i in range(0, len(names), 6)]
    out += ['END', 'THERMO', '   300.000  1000.000  5000.000']
    out += [_thermo_entry(s, *species[s]) for s in names]
    out += ['END', 'REACTIONS MOLES {}'.format(_ENERGY[units.energy])]
    for c in reactions:
        out += _reaction_lines(c, units, names)
    out.append('END')
    coeffs = {}
    for s in names:
        lo, hi = sorted(species[s][1], key=lambda c: float(c.arg(0, 'Trange')[0]))
        coeffs[s] = ([float(v) for v in lo.arg(1, 'coeffs')],
                     [float(v) for v in hi.arg(1, 'coeffs')])
    return '\n'.join(out) + '\n', coeffs


def cti_to_chemkin(text):
    """Chemkin text (thermo included) of the first phase of a CTI file."""
    return _convert(text)[0]


def read_mech_ct(filename, sort_type=None):
    """``(elems, specs, reacs)`` of a CTI file, as the Chemkin reader returns
    them for the same mechanism (mech_interpret.py:972 returns the same
    triple from Cantera).  The NASA coefficients are set from the file's
    values themselves, not their fixed-width Chemkin rendering."""
    with open(filename) as f:
        return read_converted(*_convert(f.read()), sort_type=sort_type)


def read_converted(text, coeffs, sort_type=None):
    """``(elems, specs, reacs)`` from converted Chemkin text and the exact
    NASA coefficients."""
    from .mech_interpret import read_mech
    with tempfile.TemporaryDirectory(prefix='pyjac_cti_') as d:
        path = os.path.join(d, 'mech.inp')
        with open(path, 'w') as f:
            f.write(text)
        elems, specs, reacs = read_mech(path, None, sort_type)
    for sp in specs:
        lo, hi = coeffs[sp.name]
        sp.lo[:] = lo
        sp.hi[:] = hi
    return elems, specs, reacs


def write_chemkin(path, directory=None):
    """Write the Chemkin form of the CTI file ``path`` into ``directory``
    (default: a new temporary folder) and return its path."""
    with open(path) as f:
        text = cti_to_chemkin(f.read())
    directory = directory or tempfile.mkdtemp(prefix='pyjac_cti_')
    out = os.path.join(directory, os.path.splitext(os.path.basename(path))[0] + '.inp')
    with open(out, 'w') as f:
        f.write(text)
    return out
