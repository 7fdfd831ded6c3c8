"""Cantera YAML mechanisms without Cantera.

The reference reads Cantera inputs through Cantera (``read_mech_ct``,
core/mech_interpret.py:972-1241).  A Cantera YAML file is loaded with
``yaml.safe_load`` (plain data, nothing executed) and its first phase,
species and reactions are rewritten as the equivalent CTI directives, which
the CTI reader's conversion then turns into Chemkin text for the Chemkin
reader -- so YAML, CTI and Chemkin inputs of one mechanism give the same
tables.

Supported: ``units`` (length, quantity, activation-energy, pressure);
``ideal-gas`` phases with explicit element and species lists; ``NASA7``
thermo with two temperature ranges; reactions of type elementary (no
``type``), ``three-body``, ``falloff`` (Lindemann, ``Troe``, ``SRI``),
``chemically-activated``, ``pressure-dependent-Arrhenius`` and
``Chebyshev``; ``efficiencies``, ``default-efficiency`` and ``duplicate``.
Activation energies and pressures may carry their own units
(``"6260 cal/mol"``, ``"0.04 atm"``); pre-exponential factors must be in the
file's units.
"""

import re

from .cti import Call, _PRESSURE, convert_directives, read_converted
from .exceptions import MechanismError
from .mech_interpret import _ACT_ENERGY_FACT

# Cantera's energy unit names -> the Chemkin reader's keys (activation
# temperature factors)
_E_KEY = {'cal/mol': 'cal/mole', 'kcal/mol': 'kcal/mole', 'J/mol': 'joules/mole',
          'kJ/mol': 'kjoules/mole', 'J/kmol': 'joules/kmole', 'K': 'kelvins', 'eV': 'evolts'}

_QUANTITY = re.compile(r'^\s*([-+0-9.eE]+)\s*(\S.*)?$')


def _split(v):
    """(number, unit or None) of a YAML value, plain or "value unit"."""
    if isinstance(v, (int, float)):
        return float(v), None
    m = _QUANTITY.match(str(v))
    if not m:
        raise MechanismError('YAML: cannot read quantity {!r}'.format(v))
    return float(m.group(1)), (m.group(2) or '').strip() or None


class _Units(object):
    def __init__(self, d):
        d = d or {}
        self.length = d.get('length', 'm')
        self.quantity = d.get('quantity', 'kmol')
        self.energy = d.get('activation-energy', 'J/kmol')
        self.pressure = d.get('pressure', 'Pa')
        if self.energy not in _E_KEY:
            raise MechanismError('YAML: unsupported energy unit {}'.format(self.energy))
        if self.pressure not in _PRESSURE:
            raise MechanismError('YAML: unsupported pressure unit {}'.format(self.pressure))
        if d.get('time', 's') != 's':
            raise MechanismError('YAML: unsupported time unit {}'.format(d['time']))

    def cti(self):
        return Call('units', [], {'length': self.length, 'quantity': self.quantity,
                                  'act_energy': self.energy, 'pressure': self.pressure})

    def energy_value(self, v):
        """An activation energy in the file's default unit."""
        x, unit = _split(v)
        if unit is None or unit == self.energy:
            return x
        if unit not in _E_KEY:
            raise MechanismError('YAML: unsupported energy unit {}'.format(unit))
        return x * _ACT_ENERGY_FACT[_E_KEY[unit]] / _ACT_ENERGY_FACT[_E_KEY[self.energy]]

    def pressure_value(self, v):
        """A pressure as the CTI reader takes it: bare (file's unit) or
        [value, unit]."""
        x, unit = _split(v)
        if unit is None:
            return x
        if unit not in _PRESSURE:
            raise MechanismError('YAML: unsupported pressure unit {}'.format(unit))
        return [x, unit]


def _rate(units, d, what):
    if not isinstance(d, dict):
        raise MechanismError('YAML: {} must be a mapping of A, b, Ea'.format(what))
    A, unit = _split(d.get('A', 0.0))
    if unit is not None:
        raise MechanismError('YAML: pre-exponential factors with explicit units are not '
                             'supported ({})'.format(d['A']))
    return [A, float(d.get('b', 0.0)), units.energy_value(d.get('Ea', 0.0))]


def _efficiency_kwargs(r):
    kw = {}
    eff = r.get('efficiencies')
    if eff:
        kw['efficiencies'] = ' '.join('{}:{!r}'.format(k, float(v)) for k, v in eff.items())
    if 'default-efficiency' in r:
        kw['default_efficiency'] = float(r['default-efficiency'])
    return kw


def _falloff(r):
    if 'Troe' in r:
        t = r['Troe']
        return Call('Troe', [], {k: float(t[k]) for k in ('A', 'T3', 'T1', 'T2') if k in t})
    if 'SRI' in r:
        t = r['SRI']
        return Call('SRI', [], {k: float(t[k]) for k in ('A', 'B', 'C', 'D', 'E') if k in t})
    return None


def _reaction(units, r):
    """The CTI directive of one YAML reaction."""
    eq = r.get('equation')
    if not eq:
        raise MechanismError('YAML: reaction without an equation')
    for key in ('orders', 'negative-orders', 'nonreactant-orders'):
        if key in r:
            raise MechanismError('YAML: reaction orders are not supported ({})'.format(eq))
    kind = r.get('type', 'elementary')
    kw = {'options': ['duplicate']} if r.get('duplicate') else {}
    if kind in ('elementary', 'reaction'):
        return Call('reaction', [eq, _rate(units, r['rate-constant'], eq)], kw)
    if kind == 'three-body':
        kw.update(_efficiency_kwargs(r))
        return Call('three_body_reaction', [eq, _rate(units, r['rate-constant'], eq)], kw)
    if kind in ('falloff', 'chemically-activated'):
        kw.update(_efficiency_kwargs(r))
        fo = _falloff(r)
        if fo is not None:
            kw['falloff'] = fo
        low = _rate(units, r['low-P-rate-constant'], eq)
        high = _rate(units, r['high-P-rate-constant'], eq)
        if kind == 'falloff':
            kw.update(kf=high, kf0=low)
            return Call('falloff_reaction', [eq], kw)
        kw.update(kLow=low, kHigh=high)
        return Call('chemically_activated_reaction', [eq], kw)
    if kind == 'pressure-dependent-Arrhenius':
        rates = [[units.pressure_value(k['P'])] + _rate(units, k, eq)
                 for k in r['rate-constants']]
        return Call('pdep_arrhenius', [eq] + rates, kw)
    if kind == 'Chebyshev':
        Tr, Pr = r['temperature-range'], r['pressure-range']
        kw.update(Tmin=float(Tr[0]), Tmax=float(Tr[1]), Pmin=units.pressure_value(Pr[0]),
                  Pmax=units.pressure_value(Pr[1]), coeffs=[[float(x) for x in row]
                                                            for row in r['data']])
        return Call('chebyshev_reaction', [eq], kw)
    raise MechanismError('YAML: unsupported reaction type {} ({})'.format(kind, eq))


def _species(d):
    name = d['name']
    th = d.get('thermo') or {}
    if th.get('model') != 'NASA7':
        raise MechanismError('YAML: species {}: only NASA7 thermo is supported'.format(name))
    Tr, data = th['temperature-ranges'], th['data']
    if len(Tr) != 3 or len(data) != 2:
        raise MechanismError('YAML: species {}: two NASA7 ranges are needed'.format(name))
    nasa = [Call('NASA', [[float(Tr[0]), float(Tr[1])], [float(x) for x in data[0]]], {}),
            Call('NASA', [[float(Tr[1]), float(Tr[2])], [float(x) for x in data[1]]], {})]
    atoms = ' '.join('{}:{}'.format(e, n) for e, n in d['composition'].items())
    return Call('species', [], {'name': name, 'atoms': atoms, 'thermo': nasa})


def yaml_directives(doc):
    """CTI directives (Call records) equivalent to a loaded Cantera YAML
    document."""
    if not isinstance(doc, dict):
        raise MechanismError('YAML: not a Cantera mechanism')
    units = _Units(doc.get('units'))
    phases = doc.get('phases') or []
    if not phases:
        raise MechanismError('YAML: no phase')
    ph = phases[0]
    if ph.get('thermo', 'ideal-gas') != 'ideal-gas':
        raise MechanismError('YAML: only ideal-gas phases are supported')
    species = doc.get('species') or []
    names = ph.get('species', [s['name'] for s in species])
    if not isinstance(names, list) or not all(isinstance(n, str) for n in names):
        raise MechanismError('YAML: the phase must list its species by name')
    if ph.get('reactions', 'all') not in ('all', 'declared-species'):
        raise MechanismError('YAML: the phase must take the file\'s reactions')
    calls = [units.cti(),
             Call('ideal_gas', [], {'elements': ' '.join(ph['elements']),
                                    'species': ' '.join(names), 'reactions': 'all'})]
    calls += [_species(s) for s in species]
    calls += [_reaction(units, r) for r in doc.get('reactions') or []]
    return calls


def _loader():
    """``yaml.SafeLoader`` with YAML 1.2 booleans: only true/false are
    booleans, so species such as NO, ON or Y stay names (Cantera writes its
    files for a YAML 1.2 reader)."""
    import yaml

    class Loader(yaml.SafeLoader):
        pass
    Loader.yaml_implicit_resolvers = {
        k: [(tag, rx) for tag, rx in v if tag != 'tag:yaml.org,2002:bool']
        for k, v in yaml.SafeLoader.yaml_implicit_resolvers.items()}
    Loader.add_implicit_resolver('tag:yaml.org,2002:bool',
                                 re.compile(r'^(?:true|True|TRUE|false|False|FALSE)$'),
                                 list('tTfF'))
    return Loader


def load_yaml(text):
    """The document of a Cantera YAML text (plain data only)."""
    import yaml
    return yaml.load(text, Loader=_loader())


def read_mech_yaml(filename, sort_type=None):
    """``(elems, specs, reacs)`` of a Cantera YAML file (the triple of
    mech_interpret.py:972)."""
    with open(filename) as f:
        doc = load_yaml(f.read())
    return read_converted(*convert_directives(yaml_directives(doc)), sort_type=sort_type)
