"""Every kernel family agrees with the checker: the block-per-state kernels
with 1, 2, 4 and 8 waves per state (variants 11, 12, 14, 18; the automatic
choice when a state's arrays fit in LDS), the wave-per-state kernels
(variant 2), the block-per-state kernels that accumulate the Jacobian in
device memory (variants 22, 24, 28; the automatic choice when the dense
accumulator exceeds LDS) and the lane-per-state kernels (variant 1)."""

import numpy as np
import pytest

import oracle
from rate_tol import rate_error, rate_scales
from jac_tol import assert_jacobian_close, per_state
from pyjac.core.states import sample_states
from pyjac.pywrap import PyJacobianKernel, PySpeciesRatesKernel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('variant', [1, 2, 11, 12, 14, 18, 22, 24, 28])
@pytest.mark.parametrize('name,conp,order,fmt', [('gri30', False, 'F', 1), ('test', True, 'C', 0),
                                                 ('h2o2', True, 'F', 0)])
def test_variant_parity(mechs, variant, name, conp, order, fmt):
    mech = mechs(name)
    n = 777  # not a multiple of the wave or grid sizes
    param, phi = sample_states(mech, n, seed=21, conp=conp)
    phi_in = phi if order == 'C' else np.ascontiguousarray(phi.T)
    rk = PySpeciesRatesKernel(n, 0, mech=mech, conp=conp, order=order)
    if variant < 20:  # 20 + W are Jacobian-only variants
        rk.set_variant(variant)
    out = np.zeros_like(phi_in)
    rk(param, phi_in, out)
    ref, scale = rate_scales(mech, conp, param, phi)
    got = out if order == 'C' else out.T
    assert rate_error(got, ref, scale) <= 1.0
    jk = PyJacobianKernel(n, 0, mech=mech, conp=conp, order=order, jac_format=fmt)
    jk.set_variant(variant)
    jref = oracle.to_layout(mech, oracle.jacobian(mech, conp, param, phi), order, fmt)
    jac = np.zeros_like(jref)
    jk(param, phi_in, jac)
    assert_jacobian_close(per_state(jac, n, order), per_state(jref, n, order),
                          '{} variant {}'.format(name, variant))


def test_unavailable_variant_is_an_error(mechs):
    """A variant the mechanism cannot run is refused, not silently replaced."""
    from pyjac.core.exceptions import HipRuntimeError
    rk = PySpeciesRatesKernel(64, 0, mech=mechs('gri30'), conp=True, order='C')
    with pytest.raises(HipRuntimeError, match='cannot run'):
        rk.set_variant(24)
    jk = PyJacobianKernel(64, 0, mech=mechs('s500'), conp=True, order='C', jac_format=1)
    with pytest.raises(HipRuntimeError, match='cannot run'):
        jk.set_variant(14)  # a 500-species dense accumulator exceeds LDS
    jk.set_variant(24)
