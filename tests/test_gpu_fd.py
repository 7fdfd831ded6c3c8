"""Finite-difference Jacobian on the GPU (jac_type finite_difference) through
the C ABI: against the analytical Jacobian within the difference quotient's
own error (the bounds of tests/test_fd_jacobian.py for the checker), and
against the checker's finite-difference Jacobian on the same states.  The
latter agree only to the quotient's roundoff level: the GPU's and the CPU's
rates differ in the last bits, and the quotient divides those by steps of
sqrt(eps) relative size, so both carry independent noise of ~1e-4."""

import numpy as np
import pytest

import oracle
from pyjac.core.enum_types import JacobianType
from pyjac.core.states import sample_states
from pyjac.pywrap import PyJacobianKernel

pytestmark = pytest.mark.gpu

CASES = [('h2o2', True, 'C', 0, 0, 1, 5e-3), ('gri30', False, 'F', 1, 0, 1, 5e-3),
         ('test', True, 'F', 0, 1, 4, 5e-3), ('gri30', True, 'C', 1, 2, 2, 5e-3),
         ('test', False, 'C', 1, 1, 8, 5e-3), ('gri30', False, 'F', 0, 0, 3, 1e-2)]


@pytest.mark.parametrize('name,conp,order,fmt,mode,fd_order,bound', CASES)
def test_fd_jacobian(mechs, name, conp, order, fmt, mode, fd_order, bound):
    mech = mechs(name)
    n = 200
    param, phi = sample_states(mech, n, seed=31, conp=conp)
    phi_in = phi if order == 'C' else np.ascontiguousarray(phi.T)
    k = PyJacobianKernel(n, 0, mech=mech, conp=conp, order=order, jac_format=fmt,
                         jac_type=JacobianType.finite_difference)
    k.set_fd_options(mode, fd_order)
    fd_ref = oracle.to_layout(mech, oracle.fd_jacobian(mech, conp, param, phi, mode=mode,
                                                       order=fd_order), order, fmt)
    an_ref = oracle.to_layout(mech, oracle.jacobian(mech, conp, param, phi), order, fmt)
    out = np.zeros_like(fd_ref)
    k(param, phi_in, out)
    assert np.all(np.isfinite(out))
    per = lambda a: a.reshape(n, -1) if order == 'C' else a.reshape(-1, n).T
    a, b, c = per(out), per(fd_ref), per(an_ref)
    fro_fd = np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
    fro_an = np.linalg.norm(a - c, axis=1) / np.linalg.norm(c, axis=1)
    # the checker's own difference-quotient error on each state: a few
    # states (steep or kinked rates, e.g. species near the 1e-300 floor) are
    # beyond any quotient's reach, and there both are equally noisy
    own = np.linalg.norm(b - c, axis=1) / np.linalg.norm(c, axis=1)
    # accuracy: as the checker's quotient (tests/test_fd_jacobian.py), on 95 %
    # of the states; the same noise level as the checker's quotient; and no
    # gross error on any state
    assert np.percentile(fro_an, 95) <= bound, np.percentile(fro_an, 95)
    assert np.median(fro_an) <= 2.0 * np.median(own) + 1e-4, (np.median(fro_an), np.median(own))
    assert np.all(fro_an <= np.maximum(10.0 * bound, 10.0 * own)), (fro_an.max(), own.max())
    assert np.median(fro_fd) <= 2.0 * np.median(own) + 1e-4, (np.median(fro_fd), np.median(own))


def test_fd_options_need_a_fd_handle(mechs):
    from pyjac.core.exceptions import HipRuntimeError
    k = PyJacobianKernel(8, 0, mech=mechs('h2o2'), conp=True, order='C')
    with pytest.raises(HipRuntimeError):
        k.set_fd_options(0, 1)
    kf = PyJacobianKernel(8, 0, mech=mechs('h2o2'), conp=True, order='C',
                          jac_type=JacobianType.finite_difference)
    with pytest.raises(HipRuntimeError):
        kf.set_fd_options(1, 3)  # central differences have even orders only


@pytest.mark.parametrize('name,conp,order', [('gri30', True, 'C'), ('gri30', False, 'F'),
                                             ('test', True, 'F'), ('s150', False, 'C')])
def test_approximate_jacobian(mechs, name, conp, order):
    """jac_type approximate on its sparser pattern (tables built with
    JacobianType.approximate) against the checker's Jacobian on that pattern
    (tests/test_approximate.py pins the checker)."""
    from pyjac._paths import mechanism_files
    from pyjac.core.mech_tables import MechTables
    mech = MechTables.from_file(*mechanism_files(name), jac_type=JacobianType.approximate)
    n = 300
    param, phi = sample_states(mech, n, seed=41, conp=conp)
    phi_in = phi if order == 'C' else np.ascontiguousarray(phi.T)
    ref = oracle.to_layout(mech, oracle.jacobian_on_pattern(mech, conp, param, phi), order, 1)
    k = PyJacobianKernel(n, 0, mech=mech, conp=conp, order=order, jac_format=1,
                         jac_type=JacobianType.approximate)
    out = np.zeros_like(ref)
    k(param, phi_in, out)
    a = out.reshape(n, -1) if order == 'C' else out.reshape(-1, n).T
    b = ref.reshape(n, -1) if order == 'C' else ref.reshape(-1, n).T
    fro = np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
    assert fro.max() <= 1e-8, fro.max()
