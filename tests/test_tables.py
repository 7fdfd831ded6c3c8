"""Mechanism front-end vs the golden tables of the reference's generated
sources (examples/opencl/{h2o2,ch4}, extracted by
tests/golden/extract_reference_tables.py): the parser, ``assign_rates`` with
hybrid specialisation and ``determine_jac_inds`` must reproduce every table,
the species order (bath gas last) and the reaction strings exactly."""

import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


@pytest.mark.parametrize('name,fixture', [('h2o2', 'h2o2_tables.json'),
                                          ('gri30', 'gri30_tables.json')])
def test_tables_match_reference(mechs, name, fixture):
    mech = mechs(name)
    with open(os.path.join(GOLDEN, fixture)) as fh:
        ref = json.load(fh)
    assert mech.Ns == ref['sizes']['NS']
    assert mech.Nr == ref['sizes']['NR']
    assert mech.NN == ref['sizes']['NN']
    assert mech.species_names == ref['strings']['species_names']
    assert mech.reaction_strings == ref['strings']['rxn_strings']
    assert len(ref['arrays']) >= 30
    for key, val in ref['arrays'].items():
        assert key in mech.named, key
        got = np.asarray(mech.named[key])
        exp = np.asarray(val['values'], dtype=np.int64 if val['dtype'] == 'int' else np.float64)
        assert got.shape == exp.shape, key
        # bit-exact, floats included
        assert np.array_equal(got, exp), key


def test_gri_sparsity_counts(mechs):
    mech = mechs('gri30')
    assert mech.nnz == 1984
    ccs = mech.info['jac_inds']['ccs']
    assert ccs['col_ptr'][-1] == 1984 and len(ccs['col_ptr']) == 55


def test_all_reaction_types_parse(mechs):
    """pyjac/tests/test.inp covers every reaction type."""
    from pyjac.core.enum_types import reaction_type, falloff_form
    mech = mechs('test')
    info = mech.info
    assert mech.Ns == 40 and mech.species_names[-1] == 'N2'
    assert info['plog']['num'] == 5 and info['cheb']['num'] == 2
    assert info['fall']['sri']['num'] == 3 and info['fall']['lind']['num'] == 2
    assert set(info['fall']['ftype'].tolist()) == {0, 1}
    # explicit REV parameters split into two irreversible reactions
    rs = mech.reaction_strings
    assert 'ElementaryReaction: c2h3oh + hocho => ch3cho + hocho' in rs
    assert 'ElementaryReaction: ch3cho + hocho => c2h3oh + hocho' in rs


def test_unit_conversions_closed_form(mechs):
    """Pre-exponential units: cm^3/mol -> m^3/kmol, activation energy
    cal/mol -> K (reference mech_interpret.py:39-50, 497-512)."""
    from pyjac.core.chem_model import RU_JOUL
    mech = mechs('h2o2')
    r = mech.reacs
    # O + H2 <=> H + OH  3.87e4 2.7 6260 (bimolecular: / 1000)
    rx = [x for x in r if str(x) == 'ElementaryReaction: O + H2 <=> H + OH'][0]
    assert rx.A == pytest.approx(3.87e4 / 1000.0, rel=1e-15)
    assert rx.E == pytest.approx(6260.0 * 4.184 / RU_JOUL, rel=1e-15)
    # 2O + M <=> O2 + M 1.2e17 (third body: / 1000^2)
    tb = [x for x in r if str(x) == 'ThreeBodyReaction: 2 O + M <=> O2 + M'][0]
    assert tb.A == pytest.approx(1.2e17 / 1e6, rel=1e-15)
