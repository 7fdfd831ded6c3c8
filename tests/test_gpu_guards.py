"""The reference's guards on the GPU against the checker, in C and F order:
temperatures below 100 K and above 1e4 K (rates clamp T, several Jacobian
terms read the raw T: instruction_creator.py:370-420, and e.g. the third-body
concentration P/(R T) with the raw T, examples/opencl/h2o2/species_rates.ocl
:79-100), a volume below the 1e-30 floor (VolumeGuard,
instruction_creator.py:422-440), species at exactly zero and negative moles
(the 1e-300 floor of the concentrations, rate_subs.py:640-706) and states
whose equilibrium-constant exponent exceeds the exp cap of 690.775527898
(GuardedExp, instruction_creator.py:442-462; utils.py:280-292).  The
analytical Jacobian of the reference differs from the derivative of its own
guarded rates in these regions (the clamps have zero derivative, the
Jacobian kernels do not all clamp), so the check is against the checker's
restatement of the analytical Jacobian, not against automatic
differentiation."""

import numpy as np
import pytest

import oracle
from rate_tol import rate_error, rate_scales
from jac_tol import assert_jacobian_close, per_state
from pyjac.core.states import sample_states
from pyjac.pywrap import PyJacobianKernel, PySpeciesRatesKernel

EXP_MAX = 690.775527898


def guard_states(mech, conp, seed=3):
    """Twelve states: 0-3 and 8-9 outside the temperature clamp, 4 below the
    volume floor (moles scaled so the floored concentrations stay moderate),
    5-6 species at exactly zero, 7 negative moles, 10-11 ordinary."""
    param, phi = sample_states(mech, 12, seed=seed, conp=conp)
    phi[0, 0] = 50.0
    phi[1, 0] = 2.0e4
    phi[2, 0] = 99.0
    phi[3, 0] = 1.0e4 + 5.0
    if conp:
        phi[4, 1] = 1e-35
    else:
        param[4] = 1e-35
    phi[4, 2:] *= 1e-30
    phi[5, 2:] = 0.0
    phi[6, 2:5] = 0.0
    phi[7, 2:6] = -1e-8
    phi[8, 0] = 60.0
    phi[9, 0] = 3.0e4
    return param, phi


def test_guard_states_reach_the_guards(mechs):
    """On the CPU: the chosen states do hit the exp cap (sum nu_k b_k above
    690.78 for some reversible reaction of GRI-Mech 3.0 at the clamped
    temperature) and give finite checker outputs."""
    for name in ('h2o2', 'gri30', 'test'):
        mech = mechs(name)
        for conp in (True, False):
            param, phi = guard_states(mech, conp)
            _, _, b = oracle.chem_utils(mech, conp, phi)
            A = mech.arrays
            nu = A['nu_prod'] - A['nu_reac']
            rxn = np.repeat(np.arange(mech.Nr), np.diff(A['rxn_off']))
            B = np.zeros((len(param), mech.Nr))
            np.add.at(B.T, rxn, (nu[:, None] * b[:, A['rxn_spec']].T))
            rev = A['rev_mask'] >= 0
            # at 100 K GRI-Mech 3.0 reaches 947; H2/O2 (590) and test.inp
            # (625 over its reversible reactions) stay below the cap
            if name == 'gri30':
                assert (B[:, rev].max(axis=1)[[0, 2, 8]] > EXP_MAX).all(), name
            assert np.isfinite(oracle.species_rates(mech, conp, param, phi)).all()
            assert np.isfinite(oracle.jacobian(mech, conp, param, phi)).all(), (name, conp)


@pytest.mark.gpu
@pytest.mark.parametrize('name', ['h2o2', 'gri30', 'test'])
@pytest.mark.parametrize('conp', [True, False])
@pytest.mark.parametrize('order', ['C', 'F'])
def test_guards_rates_and_jacobian(mechs, name, conp, order):
    mech = mechs(name)
    param, phi = guard_states(mech, conp)
    n = len(param)
    print('guards', name, 'conp', conp, order, flush=True)
    phi_in = phi if order == 'C' else np.ascontiguousarray(phi.T)
    rk = PySpeciesRatesKernel(n, 0, mech=mech, conp=conp, order=order)
    out = np.zeros_like(phi_in)
    rk(param, phi_in, out)
    got = out if order == 'C' else out.T
    assert np.isfinite(got).all()
    ref, scale = rate_scales(mech, conp, param, phi)
    worst = rate_error(got, ref, scale)
    assert worst <= 1.0, 'rates error {:.3g} x tolerance'.format(worst)
    for fmt in (0, 1):
        jref = oracle.to_layout(mech, oracle.jacobian(mech, conp, param, phi), order, fmt)
        jk = PyJacobianKernel(n, 0, mech=mech, conp=conp, order=order, jac_format=fmt)
        jac = np.zeros_like(jref)
        jk(param, phi_in, jac)
        assert_jacobian_close(per_state(jac, n, order), per_state(jref, n, order),
                              'guards {} conp={} {} fmt={}'.format(name, conp, order, fmt))
