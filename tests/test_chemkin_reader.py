"""The Chemkin reader on small hand-written mechanisms: equation syntax
(coefficients with and without a space, ions ending in '+', third bodies,
falloff colliders), auxiliary items, units, embedded THERMO blocks and the
errors for what it does not read.  The bundled mechanisms are pinned bit for
bit by tests/test_tables.py."""

import pytest

from pyjac.core.chem_model import RU_JOUL
from pyjac.core.exceptions import MechanismError
from pyjac.core.mech_interpret import read_mech

THERMO = """THERMO ALL
   300.000  1000.000  5000.000
{body}END
"""

# one record per species, NASA-7 fixed columns (copied coefficients of H2
# for every species: the reader does not judge the chemistry)
REC = ("{name:<18}      {elems:<20}G   300.000  5000.000 1000.000    1\n"
       " 3.33727920E+00-4.94024731E-05 4.99456778E-07-1.79566394E-10 2.00255376E-14    2\n"
       "-9.50158922E+02-3.20502331E+00 2.34433112E+00 7.98052075E-03-1.94781510E-05    3\n"
       " 2.01572094E-08-7.37611761E-12-9.17935173E+02 6.83010238E-01                   4\n")


def _mech(tmp_path, species, reactions, elements='H O E N', units=''):
    recs = ''.join(REC.format(name=s, elems='H   2') for s in species)
    text = 'ELEMENTS\n{}\nEND\nSPECIES\n{}\nEND\n'.format(elements, ' '.join(species))
    text += THERMO.format(body=recs)
    text += 'REACTIONS {}\n{}\nEND\n'.format(units, reactions)
    p = tmp_path / 'mech.inp'
    p.write_text(text)
    return str(p)


def test_equation_syntax(tmp_path):
    sp = ['H', 'H2', 'O', 'OH', 'H3O+', 'E', 'N2']
    rx = '\n'.join([
        'H+H+M<=>H2+M   1.0E18 -1.0 0.0',
        'H2/2.5/ N2/ .7/',
        '2 H + H2 => 2H2   9.0E16 -0.6 0.0',
        'H3O+ + E=>H2+OH   1.0E13 0.0 1000.0',
        'O+H(+N2)=OH(+N2)   1.0D13 0.0 0.0',
        'LOW / 1.0D20 -1.0 0.0 /',
        'TROE/ 0.5 0.0 0.0 /',
        'DUPLICATE'])
    _, specs, reacs = read_mech(_mech(tmp_path, sp, rx))
    assert [s.name for s in specs] == sp
    r0, r1, r2, r3 = reacs
    assert r0.thd_body and r0.reac == ['H'] and r0.reac_nu == [2]
    assert r0.thd_body_eff == [['H2', 2.5], ['N2', 0.7]]
    # a third-body reaction: A from cm^6/mol^2/s to m^6/kmol^2/s
    assert r0.A == 1.0e18 / 1000. ** 2
    assert r1.reac == ['H', 'H2'] and r1.reac_nu == [2, 1] and not r1.rev
    assert r1.prod == ['H2'] and r1.prod_nu == [2]
    assert r2.reac == ['H3O+', 'E'] and r2.prod == ['H2', 'OH']
    assert r2.E == 1000.0 * (4.184 / RU_JOUL)
    assert r3.pdep and not r3.thd_body and r3.pdep_sp == 'N2' and r3.dup
    assert r3.low == [1.0e20 / 1000. ** 2, -1.0, 0.0]
    assert r3.troe_par == [0.5, 1e-30, 1e-30]


def test_units_and_explicit_reverse(tmp_path):
    sp = ['H', 'H2', 'N2']
    rx = 'H+H=>H2  2.0E12 0.5 10.0\nH+H<=>H2  3.0E12 0.0 2.0\nREV / 4.0E5 0.1 3.0 /'
    _, _, reacs = read_mech(_mech(tmp_path, sp, rx, units='MOLES KJOULES/MOLE'))
    assert len(reacs) == 3
    assert reacs[0].E == 10.0 * (1000.0 / RU_JOUL)
    fwd, back = reacs[1], reacs[2]
    assert not fwd.rev and not back.rev
    assert back.reac == ['H2'] and back.prod == ['H'] and back.prod_nu == [2]
    assert back.A == 4.0e5 and back.b == 0.1 and back.E == 3.0 * (1000.0 / RU_JOUL)


@pytest.mark.parametrize('rx,match', [
    ('H+H<=>H2  1.0 0 0\nFORD / H 1.5 /', 'not supported'),
    ('H+H+M<=>H2+M  1.0 0 0\nXX/2.0/', 'unknown species'),
    ('H+HE<=>H2  1.0 0 0', 'unknown species'),
    ('LOW / 1 2 3 /', 'before any reaction'),
])
def test_errors(tmp_path, rx, match):
    with pytest.raises(MechanismError, match=match):
        read_mech(_mech(tmp_path, ['H', 'H2', 'N2'], rx))


def test_units_errors(tmp_path):
    with pytest.raises(MechanismError, match='Molecules'):
        read_mech(_mech(tmp_path, ['H', 'H2'], 'H+H<=>H2 1 0 0', units='MOLECULES'))
    with pytest.raises(MechanismError, match='Unsupported units'):
        read_mech(_mech(tmp_path, ['H', 'H2'], 'H+H<=>H2 1 0 0', units='FURLONGS'))


def test_missing_thermo(tmp_path):
    p = tmp_path / 'm.inp'
    p.write_text('ELEM H END\nSPEC H H2 END\nREACTIONS\nH+H<=>H2 1 0 0\nEND\n')
    with pytest.raises(MechanismError, match='no thermo file'):
        read_mech(str(p))
