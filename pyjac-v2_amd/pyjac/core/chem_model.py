"""Species / reaction records and physical constants.

Mirrors the data model of ``pyjac/core/chem_model.py`` (constants :20-28,
``ReacInfo`` :82, ``ReacInfo.finalize`` :275, ``SpecInfo`` :443) so that the
table builders see exactly the same reaction classification as the reference.
"""

import numpy as np

from .enum_types import (reaction_type, thd_body_type, falloff_form,
                         reversible_type)

#: universal gas constant, J/(kmol K)  (reference chem_model.py:20)
RU = 8314.4621
RU_JOUL = 8.3144621
RUC = RU / 4.18400
#: one standard atmosphere, Pa  (reference chem_model.py:28)
PA = 101325.0

# Atomic weights, kg/kmol.  The molecular weights built from these must be
# bit-identical to the reference's (the bath-gas search compares N2/AR/HE
# weights with ``==``, reference create_jacobian.py:5183-5190), so the values
# are the reference's table (chem_model.py:33-77), which is standard data.
_ELEM_WT = (
    ('h', 1.00794), ('he', 4.00260), ('li', 6.93900), ('be', 9.01220),
    ('b', 10.81100), ('c', 12.0110), ('n', 14.00674), ('o', 15.99940),
    ('f', 18.99840), ('ne', 20.18300), ('na', 22.98980), ('mg', 24.31200),
    ('al', 26.98150), ('si', 28.08600), ('p', 30.97380), ('s', 32.06400),
    ('cl', 35.45300), ('ar', 39.94800), ('k', 39.10200), ('ca', 40.08000),
    ('sc', 44.95600), ('ti', 47.90000), ('v', 50.94200), ('cr', 51.99600),
    ('mn', 54.93800), ('fe', 55.84700), ('co', 58.93320), ('ni', 58.71000),
    ('cu', 63.54000), ('zn', 65.37000), ('ga', 69.72000), ('ge', 72.59000),
    ('as', 74.92160), ('se', 78.96000), ('br', 79.90090), ('kr', 83.80000),
    ('rb', 85.47000), ('sr', 87.62000), ('y', 88.90500), ('zr', 91.22000),
    ('nb', 92.90600), ('mo', 95.94000), ('tc', 99.00000), ('ru', 101.07000),
    ('rh', 102.90500), ('pd', 106.40000), ('ag', 107.87000),
    ('cd', 112.40000), ('in', 114.82000), ('sn', 118.69000),
    ('sb', 121.75000), ('te', 127.60000), ('i', 126.90440),
    ('xe', 131.30000), ('cs', 132.90500), ('ba', 137.34000),
    ('la', 138.91000), ('ce', 140.12000), ('pr', 140.90700),
    ('nd', 144.24000), ('pm', 145.00000), ('sm', 150.35000),
    ('eu', 151.96000), ('gd', 157.25000), ('tb', 158.92400),
    ('dy', 162.50000), ('ho', 164.93000), ('er', 167.26000),
    ('tm', 168.93400), ('yb', 173.04000), ('lu', 174.99700),
    ('hf', 178.49000), ('ta', 180.94800), ('w', 183.85000),
    ('re', 186.20000), ('os', 190.20000), ('ir', 192.20000),
    ('pt', 195.09000), ('au', 196.96700), ('hg', 200.59000),
    ('tl', 204.37000), ('pb', 207.19000), ('bi', 208.98000),
    ('po', 210.00000), ('at', 210.00000), ('rn', 222.00000),
    ('fr', 223.00000), ('ra', 226.00000), ('ac', 227.00000),
    ('th', 232.03800), ('pa', 231.00000), ('u', 238.03000),
    ('np', 237.00000), ('pu', 242.00000), ('am', 243.00000),
    ('cm', 247.00000), ('bk', 249.00000), ('cf', 251.00000),
    ('es', 254.00000), ('fm', 253.00000), ('d', 2.01410),
    ('e', 5.48578e-4))


def get_elem_wt():
    """Fresh dict of lower-case element name -> atomic weight [kg/kmol]."""
    return dict(_ELEM_WT)


class SpecInfo(object):
    """A species: name, elemental composition, molecular weight and NASA-7
    coefficients (``lo``/``hi``) with ``Trange = [T_low, T_mid, T_high]``
    (reference chem_model.py:443-520)."""

    def __init__(self, name):
        self.name = name
        self.elem = []
        self.mw = 0.0
        self.hi = np.zeros(7)
        self.lo = np.zeros(7)
        self.Trange = [300.0, 1000.0, 5000.0]

    def finalize(self):
        self.elem = [x for x in self.elem if x[1] != 0]

    def __repr__(self):
        return 'SpecInfo({})'.format(self.name)


def _is_integer(val):
    try:
        return float(val).is_integer()
    except (TypeError, ValueError):
        return False


class ReacInfo(object):
    """A reaction (reference chem_model.py:82-441).

    ``A`` is in SI kmol units, ``E`` is stored as an activation temperature
    [K].  After :meth:`finalize` the ``type`` list holds one enum of each of
    :class:`reversible_type`, :class:`reaction_type`, :class:`thd_body_type`
    and :class:`falloff_form`, which is what the table builders match on.
    """

    def __init__(self, rev, reactants, reac_nu, products, prod_nu, A, b, E):
        self.reac = reactants
        self.reac_nu = reac_nu
        self.prod = products
        self.prod_nu = prod_nu
        self.A = A
        self.b = b
        self.E = E
        self.rev = rev
        self.rev_par = []
        self.dup = False
        self.thd_body = False
        self.thd_body_eff = []
        self.pdep = False
        self.pdep_sp = ''
        self.low = []
        self.high = []
        self.troe = False
        self.troe_par = []
        self.sri = False
        self.sri_par = []
        self.cheb = False
        self.cheb_n_temp = 0
        self.cheb_n_pres = 0
        self.cheb_plim = [0.001 * PA, 100. * PA]
        self.cheb_tlim = [300., 2500.]
        self.cheb_par = None
        self.plog = False
        self.plog_par = None
        self.type = []
        self.rxn_str = ''

    def finalize(self, num_species):
        """Classify the reaction (reference chem_model.py:275-372)."""
        if self.rev:
            self.type.append(reversible_type.explicit if self.rev_par
                             else reversible_type.non_explicit)
        else:
            self.type.append(reversible_type.non_reversible)

        if self.thd_body:
            self.type.append(reaction_type.thd)
        elif self.pdep and self.low:
            self.type.append(reaction_type.fall)
        elif self.pdep and self.high:
            self.type.append(reaction_type.chem)
        elif self.plog:
            self.type.append(reaction_type.plog)
        elif self.cheb:
            self.type.append(reaction_type.cheb)
        else:
            self.type.append(reaction_type.elementary)

        if self.match((reaction_type.fall, reaction_type.chem,
                       reaction_type.thd)):
            if self.pdep_sp:
                self.type.append(thd_body_type.species)
            elif not self.thd_body_eff or (
                    len(self.thd_body_eff) == num_species and
                    all(t[1] == 1.0 for t in self.thd_body_eff)):
                self.type.append(thd_body_type.unity)
            else:
                self.type.append(thd_body_type.mix)
        else:
            self.type.append(thd_body_type.none)

        if self.match((reaction_type.fall, reaction_type.chem)):
            if self.troe:
                self.type.append(falloff_form.troe)
            elif self.sri:
                self.type.append(falloff_form.sri)
            else:
                self.type.append(falloff_form.lind)
        else:
            self.type.append(falloff_form.none)

        if self.plog or self.cheb:
            self.A = 0
            self.b = 0
            self.E = 0

        self.rxn_str = self._make_string()

    def _side(self, species, nus):
        parts = []
        for s, nu in zip(species, nus):
            nu_str = '{}'.format(int(nu) if _is_integer(nu) else nu)
            parts.append('{} {}'.format(nu_str, s) if nu != 1 else str(s))
        out = ' + '.join(parts)
        if self.match((reaction_type.fall, reaction_type.chem)):
            out += ' (+{})'.format(self.pdep_sp) if self.pdep_sp else ' (+M)'
        elif self.match((reaction_type.thd,)):
            out += ' + M'
        return out

    def _make_string(self):
        prefix = {reaction_type.fall: 'FalloffReaction: ',
                  reaction_type.chem: 'ChemicallyActivatedReaction: ',
                  reaction_type.thd: 'ThreeBodyReaction: ',
                  reaction_type.plog: 'PLOGReaction: ',
                  reaction_type.cheb: 'ChebyshevReaction: ',
                  reaction_type.elementary: 'ElementaryReaction: '}
        rtype = self.get_type(reaction_type)[0]
        arrow = ' <=> ' if self.rev else ' => '
        return (prefix[rtype] + self._side(self.reac, self.reac_nu) + arrow +
                self._side(self.prod, self.prod_nu))

    def __str__(self):
        return self.rxn_str

    __repr__ = __str__

    def get_type(self, enum_class):
        return [x for x in self.type if isinstance(x, enum_class)]

    def match(self, reac_types):
        """True if, for every enum class present in ``reac_types``, this
        reaction carries one of the listed members (reference
        chem_model.py:392-441)."""
        if not reac_types:
            return True
        if not isinstance(reac_types, (tuple, list)):
            reac_types = (reac_types,)
        by_class = {}
        for t in reac_types:
            by_class.setdefault(type(t), []).append(t)
        return all(any(x in self.type for x in members)
                   for members in by_class.values())
