"""Validation harness on the GPU: the PaSR generator with its reaction
sub-steps on the HIP integrator against the checker's integrator, and the
error norms of the HIP kernels on PaSR states of GRI-Mech 3.0 against the
checker's automatically differentiated Jacobian and species rates."""

import os

import numpy as np
import pytest

import oracle
from pyjac.core.mech_tables import MechTables
from pyjac.functional_tester import pasr
from pyjac.functional_tester.validate import validate_jacobian, validate_species_rates

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MECH = os.path.join(ROOT, 'pyjac-v2_amd', 'pyjac', 'data', 'mechanisms')
AIR = {'O2': 1.0, 'N2': 3.76}


def oracle_reactor(mech, rtol, atol):
    def integrate(param, phi, dt):
        y, st = oracle.integrate_ros2(mech, True, param, phi, dt, rtol=rtol, atol=atol)
        assert st[3] == 0, st
        return y
    return pasr.phi_reactor(mech, integrate)


def test_pasr_gpu_reactor_matches_checker():
    h2 = MechTables.from_file(os.path.join(MECH, 'h2o2.cti'))
    ref = np.load(os.path.join(ROOT, 'tests', 'golden', 'h2_pasr_states.npz'))
    kw = dict(complete_products=['H2O', 'N2'], num_part=10, num_res=0.1, seed=5,
              pilot=(ref['T'][0], ref['Y'][0]))
    gpu = pasr.run_simulation(h2, 'premixed', 300.0, 1.0, 1.0, {'H2': 1.0}, AIR,
                              react=pasr.gpu_reactor(h2, rtol=1e-6, atol=1e-14), **kw)
    cpu = pasr.run_simulation(h2, 'premixed', 300.0, 1.0, 1.0, {'H2': 1.0}, AIR,
                              react=oracle_reactor(h2, 1e-6, 1e-14), **kw)
    assert gpu.shape == cpu.shape == (11, 10, 13)
    np.testing.assert_allclose(gpu[..., 1], cpu[..., 1], rtol=1e-6)
    np.testing.assert_allclose(gpu[..., 3:], cpu[..., 3:], rtol=0, atol=1e-8)


@pytest.fixture(scope='module')
def gri_pasr():
    gri = MechTables.from_file(*[os.path.join(MECH, f) for f in ('grimech30.inp', 'thermo30.dat')])
    data = pasr.run_simulation(gri, 'premixed', 600.0, 1.0, 1.0, {'CH4': 1.0}, AIR,
                               complete_products=['CO2', 'H2O', 'N2'], num_part=16, num_res=0.1,
                               seed=7, react=pasr.gpu_reactor(gri, rtol=1e-6, atol=1e-14))
    return gri, data


def test_gri_pasr_states(gri_pasr):
    gri, data = gri_pasr
    assert data.shape == (11, 16, gri.Ns + 3)
    T = data[..., 1]
    # premixed stoichiometric methane/air from 600 K: the pilot at the
    # adiabatic flame temperature (about 2370 K with dissociation), fresh
    # gas at 600 K
    assert 2300.0 < T[0].min() < 2450.0 and T.min() >= 600.0 - 1e-6
    assert T[0].max() - T[0].min() < 1e-9
    np.testing.assert_allclose(data[..., 3:].sum(axis=-1), 1.0, rtol=1e-12)


@pytest.mark.parametrize('conp', [True, False])
def test_validation_of_the_jacobian_on_pasr_states(gri_pasr, conp):
    """The reference's Jacobian norms of the HIP kernel (F order, sparse)
    against the checker's analytical Jacobian; a perturbation planted in the
    answers shows in the norms."""
    gri, data = gri_pasr
    param, phi = pasr.particle_states(gri, data, conp=conp)

    def answer(p, f):
        return oracle.to_layout(gri, oracle.jacobian(gri, conp, p, f), 'F', 1)
    err = validate_jacobian(gri, param, phi, answer, conp=conp, order='F', jac_format=1,
                            chunk=64)
    print({k: v for k, v in err.items() if not k.endswith('amax')})
    assert err['jac_lapack'] < 1e-10
    assert err['jac_thresholded_15'] < 1e-6
    # entries where the checker's answer is exactly zero: structural zeros
    # are zero in both; terms that cancel exactly in the checker's order of
    # operations may leave a rounding residue here (8e-25 seen), far below
    # the tester's absolute tolerance of 1e-2
    assert err['jac_zero'] < 1e-15

    def planted(p, f):
        a = answer(p, f)
        a[a.shape[0] // 2] *= 1.0 + 1e-3
        return a
    bad = validate_jacobian(gri, param, phi, planted, conp=conp, order='F', jac_format=1,
                            chunk=64)
    assert bad['jac_thresholded_15'] > 1e-4


def test_validation_against_automatic_differentiation(gri_pasr):
    """Against the automatically differentiated Jacobian, as the reference's
    harness does.  Fresh-gas particles hold species at exactly zero, where the
    analytical derivatives (concentrations floored at 1e-300) and the
    automatic ones legitimately differ; a floor of 1e-30 on the mass
    fractions removes those exact zeros."""
    gri, data = gri_pasr
    data = data.copy()
    data[..., 3:] = np.maximum(data[..., 3:], 1e-30)
    param, phi = pasr.particle_states(gri, data, conp=False)

    def answer(p, f):
        return oracle.to_layout(gri, oracle.jacobian_ad(gri, False, p, f), 'F', 1)
    err = validate_jacobian(gri, param, phi, answer, conp=False, order='F', jac_format=1,
                            chunk=64)
    print({k: v for k, v in err.items() if not k.endswith('amax')})
    assert err['jac_lapack'] < 1e-10
    assert err['jac_thresholded_15'] < 1e-6


def test_validation_of_species_rates_on_pasr_states(gri_pasr):
    gri, data = gri_pasr
    param, phi = pasr.particle_states(gri, data, conp=False)
    err = validate_species_rates(gri, param, phi,
                                 lambda p, f: oracle.species_rates(gri, False, p, f), conp=False)
    # the species' rates, weighted by atol 1e-10 + rtol 1e-6 |ref|: far below
    # one.  Entries 0 and 1 (dT/dt, dP/dt) of particles at equilibrium are
    # near-total cancellations whose reference is almost zero; the condition-
    # aware tolerance of the parity tests covers them.
    assert err['dphi_store'][2:].max() < 1e-2


def test_pasr_command_line(tmp_path):
    """``python -m pyjac.functional_tester`` with the reference's own PaSR
    input (data/pasr_input.yaml: premixed H2/air, 10 particles, one residence
    time) writes particle data shaped like the reference's output
    (data/h2_pasr_output.npy: 102 steps x 10 particles x [t, T, P, 10 Y])
    and spanning the same temperatures: the pilot at Cantera's equilibrium
    and fresh gas after its first mixing near 521 K."""
    import subprocess
    import sys
    out = tmp_path / 'pasr.npy'
    # the command prints its mean temperature after every step
    res = subprocess.run([sys.executable, '-m', 'pyjac.functional_tester', '-i',
                          os.path.join(ROOT, 'tests', 'golden', 'pasr_input.yaml'), '-m',
                          os.path.join(MECH, 'h2o2.cti'), '-o', str(out), '--seed', '1',
                          '--rtol', '1e-6', '--atol', '1e-14'],
                         cwd=os.path.join(ROOT, 'pyjac-v2_amd'), timeout=400)
    assert res.returncode == 0
    data = np.load(str(out))
    assert data.shape == (102, 10, 13)
    ref = np.load(os.path.join(ROOT, 'tests', 'golden', 'h2_pasr_states.npz'))
    T = data[..., 1]
    assert abs(T[0].max() - ref['T'][0]) < 1e-3
    assert 500.0 < T.min() < 540.0 and abs(ref['T'].min() - T.min()) < 5.0
    np.testing.assert_allclose(data[..., 3:].sum(axis=-1), 1.0, rtol=1e-12)
