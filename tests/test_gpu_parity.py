"""HIP kernels vs the CPU checker on seeded states (through the C ABI).

Tolerances: species rates within ``|gpu - ref| <= 1e-10 |ref| + 1e-13 scale
+ 1e-20`` element-wise, where ``scale`` is the sum of the absolute
contributions the element is summed from (tests/rate_tol.py: cancelling
near-equilibrium sums magnify last-bit differences of exp between the GPU
and CPU math libraries); Jacobians within a relative Frobenius error of 1e-8
per state (pyJac's Jacobian error norm, functional_tester/test.py:1131-1242)
and element-wise within ``1e-9 |ref| + 1e-14 max_state |ref|``
(tests/jac_tol.py), all measured against the checker's FP64 result on the
same inputs.
"""

import numpy as np
import pytest

import oracle
from rate_tol import rate_error, rate_scales
from jac_tol import assert_jacobian_close, per_state
from pyjac.core.states import sample_states
from pyjac.pywrap import PyJacobianKernel, PySpeciesRatesKernel, PyChemUtilsKernel

pytestmark = pytest.mark.gpu

CASES = [(m, conp, order) for m in ('h2o2', 'gri30', 'test')
         for conp in (True, False) for order in ('C', 'F')]


def _states(mech, n, conp, order, seed=7):
    param, phi = sample_states(mech, n, seed=seed, T_range=(900.0, 2400.0),
                               alpha=0.5, conp=conp)
    dev_phi = phi if order == 'C' else np.ascontiguousarray(phi.T)
    return param, phi, dev_phi


def _rate_err(got, mech, conp, param, phi):
    ref, scale = rate_scales(mech, conp, param, phi)
    return rate_error(got, ref, scale)


@pytest.mark.parametrize('name,conp,order', CASES)
def test_species_rates(mechs, name, conp, order):
    mech = mechs(name)
    n = 1000
    param, phi, dphi_in = _states(mech, n, conp, order)
    k = PySpeciesRatesKernel(n, 0, mech=mech, conp=conp, order=order)
    out = np.zeros_like(dphi_in)
    k(param, dphi_in, out)
    got = out if order == 'C' else out.T
    assert np.all(np.isfinite(got))
    worst = _rate_err(got, mech, conp, param, phi)
    assert worst <= 1.0, 'rates error {:.3g} x tolerance'.format(worst)


@pytest.mark.parametrize('name,conp,order', CASES)
@pytest.mark.parametrize('fmt', [0, 1])
def test_jacobian(mechs, name, conp, order, fmt):
    mech = mechs(name)
    n = 600
    param, phi, phi_in = _states(mech, n, conp, order)
    dense = oracle.jacobian(mech, conp, param, phi)
    ref = oracle.to_layout(mech, dense, order, fmt)
    k = PyJacobianKernel(n, 0, mech=mech, conp=conp, order=order, jac_format=fmt)
    out = np.zeros_like(ref)
    k(param, phi_in, out)
    assert_jacobian_close(per_state(out, n, order), per_state(ref, n, order),
                          '{} conp={} {} fmt={}'.format(name, conp, order, fmt))


@pytest.mark.parametrize('name,conp', [('gri30', True), ('test', False)])
def test_chem_utils(mechs, name, conp):
    mech = mechs(name)
    n = 512
    param, phi = sample_states(mech, n, seed=3, conp=conp)
    e_ref, h_ref, b_ref = oracle.chem_utils(mech, conp, phi)
    k = PyChemUtilsKernel(n, 0, mech=mech, conp=conp, order='C')
    b = np.zeros((n, mech.Ns))
    e = np.zeros_like(b)
    h = np.zeros_like(b)
    k(phi, b, e, h)
    for got, ref in ((e, e_ref), (h, h_ref), (b, b_ref)):
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


LARGE = [('s150', True, 'C', 0), ('s150', False, 'F', 1), ('s500', True, 'F', 1),
         ('s500', False, 'C', 1)]


@pytest.mark.parametrize('name,conp,order,fmt', LARGE)
def test_large_mechanisms(mechs, name, conp, order, fmt):
    """Seeded synthetic mechanisms of 150 and 500 species: species rates and
    Jacobians (whose dense accumulator exceeds LDS, so the block kernels that
    accumulate in device memory take them, through the chunked transpose in
    F order), and the lane-per-state kernels on the same states."""
    mech = mechs(name)
    n = 96
    param, phi, phi_in = _states(mech, n, conp, order, seed=17)
    rk = PySpeciesRatesKernel(n, 0, mech=mech, conp=conp, order=order)
    out = np.zeros_like(phi_in)
    rk(param, phi_in, out)
    got = out if order == 'C' else out.T
    assert _rate_err(got, mech, conp, param, phi) <= 1.0
    dense = oracle.jacobian(mech, conp, param, phi)
    jref = oracle.to_layout(mech, dense, order, fmt)
    jk = PyJacobianKernel(n, 0, mech=mech, conp=conp, order=order, jac_format=fmt)
    jac = np.zeros_like(jref)
    jk(param, phi_in, jac)
    b = per_state(jref, n, order)
    assert_jacobian_close(per_state(jac, n, order), b, name)
    jk.set_variant(1)
    jac1 = np.zeros_like(jref)
    jk(param, phi_in, jac1)
    assert_jacobian_close(per_state(jac1, n, order), b, name + ' lane-per-state')


def test_chunked_f_order(mechs, monkeypatch):
    """F-order Jacobians of the device-memory variant go through chunks of
    states and a transpose per chunk: 200 states in chunks of 64 (the last
    one partial) match the checker."""
    monkeypatch.setenv('PYJAC_HIP_CHUNK_STATES', '64')
    mech = mechs('s150')
    n = 200
    param, phi, phi_in = _states(mech, n, False, 'F', seed=23)
    jref = oracle.to_layout(mech, oracle.jacobian(mech, False, param, phi), 'F', 1)
    jk = PyJacobianKernel(n, 0, mech=mech, conp=False, order='F', jac_format=1)
    jac = np.zeros_like(jref)
    jk(param, phi_in, jac)
    assert_jacobian_close(per_state(jac, n, 'F'), per_state(jref, n, 'F'), 's150 chunks')


@pytest.mark.parametrize('fmt', [0, 1])
@pytest.mark.parametrize('n', [200, 201])
def test_pipelined_f_order_chunks(mechs, monkeypatch, fmt, n):
    """The LDS kernel's F-order path: C-order chunks and their transposes on
    a second stream, two chunk buffers reused in turn.  With chunks of 64
    states, 200 states make four chunks (both buffers reused, the last chunk
    partial) and 201 states take the single-value transpose (odd state
    count); every state matches the checker."""
    monkeypatch.setenv('PYJAC_HIP_FCHUNK', '64')
    mech = mechs('gri30')
    param, phi, phi_in = _states(mech, n, False, 'F', seed=29)
    jref = oracle.to_layout(mech, oracle.jacobian(mech, False, param, phi), 'F', fmt)
    jk = PyJacobianKernel(n, 0, mech=mech, conp=False, order='F', jac_format=fmt)
    jac = np.full_like(jref, np.nan)
    jk(param, phi_in, jac)
    assert_jacobian_close(per_state(jac, n, 'F'), per_state(jref, n, 'F'),
                          'gri30 F chunks of 64, n={} fmt={}'.format(n, fmt))


@pytest.mark.parametrize('conp,order,fmt', [(True, 'C', 0), (False, 'F', 1)])
def test_pasr_states(mechs, conp, order, fmt):
    """The reference's partially stirred reactor states (H2/air near
    equilibrium, mapped into GRI-Mech 3.0): rates and Jacobians, where the
    production rates are sums of nearly cancelling forward and reverse
    rates."""
    from test_pasr_states import pasr_states
    mech = mechs('gri30')
    param, phi = pasr_states(mech, conp)
    n = len(param)
    phi_in = phi if order == 'C' else np.ascontiguousarray(phi.T)
    rk = PySpeciesRatesKernel(n, 0, mech=mech, conp=conp, order=order)
    out = np.zeros_like(phi_in)
    rk(param, phi_in, out)
    got = out if order == 'C' else out.T
    assert np.all(np.isfinite(got))
    assert _rate_err(got, mech, conp, param, phi) <= 1.0
    jref = oracle.to_layout(mech, oracle.jacobian(mech, conp, param, phi), order, fmt)
    jk = PyJacobianKernel(n, 0, mech=mech, conp=conp, order=order, jac_format=fmt)
    jac = np.zeros_like(jref)
    jk(param, phi_in, jac)
    assert_jacobian_close(per_state(jac, n, order), per_state(jref, n, order), 'pasr')
