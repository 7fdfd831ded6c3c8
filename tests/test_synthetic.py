"""Seeded synthetic mechanisms (S150, S500): deterministic, parsed by the
same front-end, with the requested reaction mix, and the checker's
analytical Jacobian equal to its AD Jacobian on them."""

import hashlib

import numpy as np
import pytest

import oracle
from pyjac.core.states import sample_states
from pyjac.core.synthetic import WORKLOADS, synthetic_mechanism


def test_generator_is_deterministic():
    a = synthetic_mechanism(**dict(WORKLOADS['s150'], n_reactions=50))
    b = synthetic_mechanism(**dict(WORKLOADS['s150'], n_reactions=50))
    assert hashlib.sha1(a.encode()).hexdigest() == hashlib.sha1(b.encode()).hexdigest()


@pytest.mark.parametrize('name', ['s150', 's500'])
def test_reaction_mix(mechs, name):
    mech = mechs(name)
    cfg = WORKLOADS[name]
    assert mech.Ns == cfg['n_species'] and mech.Nr == cfg['n_reactions']
    assert mech.species_names[-1] == 'AR'
    kinds = {}
    for r in mech.reacs:
        for t in r.type:
            kinds[str(t).split('.')[-1]] = kinds.get(str(t).split('.')[-1], 0) + 1
    nr = cfg['n_reactions']
    for key, frac in (('thd', cfg['thd']), ('troe', cfg['troe']), ('sri', cfg['sri']),
                      ('plog', cfg['plog']), ('cheb', cfg['cheb'])):
        got = kinds.get(key, 0) / nr
        assert abs(got - frac) <= 0.03, (key, got, frac)
    # the inert bath gas never takes part, so no Jacobian row is made full by it
    assert mech.info['jac_inds']['has_ns'] if 'has_ns' in mech.info['jac_inds'] else True


@pytest.mark.parametrize('name,conp,n', [('s150', True, 4), ('s150', False, 4), ('s500', True, 1)])
def test_analytical_equals_ad(mechs, name, conp, n):
    mech = mechs(name)
    param, phi = sample_states(mech, n, seed=31, T_range=(900.0, 2400.0), alpha=0.5, conp=conp)
    an = oracle.jacobian(mech, conp, param, phi)
    ad = oracle.jacobian_ad(mech, conp, param, phi)
    fro = np.linalg.norm((an - ad).reshape(n, -1), axis=1) / np.linalg.norm(ad.reshape(n, -1), axis=1)
    assert fro.max() < 1e-11, fro
