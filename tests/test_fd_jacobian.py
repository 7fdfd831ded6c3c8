"""Finite-difference Jacobian of the checker (finite_difference_jacobian,
core/create_jacobian.py:4523-4896, with CVODE's inverse error weights and
the textbook coefficients; see oracle/pyjac_oracle.cpp) against the
analytical Jacobian, itself pinned by dual-number AD: every mode and order
within the error a difference quotient with sqrt(eps)-sized steps leaves.
A coefficient table that does not sum to zero, or steps from
non-inverted weights (the reference's formula), miss these bounds by orders
of magnitude."""

import numpy as np
import pytest

import oracle
from pyjac.core.states import sample_states

MODES = [(0, 1, 5e-3), (2, 1, 5e-3), (0, 2, 5e-3), (0, 3, 1e-2), (0, 4, 1e-2), (0, 6, 5e-2),
         (2, 3, 1e-2), (1, 2, 5e-3), (1, 4, 5e-3), (1, 8, 5e-3)]


@pytest.mark.parametrize('name,conp', [('h2o2', True), ('gri30', False), ('test', True)])
def test_fd_matches_analytical(mechs, name, conp):
    mech = mechs(name)
    n = 12
    param, phi = sample_states(mech, n, seed=3, conp=conp)
    ref = oracle.jacobian(mech, conp, param, phi).reshape(n, -1)
    for mode, order, bound in MODES:
        fd = oracle.fd_jacobian(mech, conp, param, phi, mode=mode, order=order).reshape(n, -1)
        fro = np.linalg.norm(fd - ref, axis=1) / np.linalg.norm(ref, axis=1)
        assert fro.max() <= bound, (mode, order, fro.max())


def test_fd_of_unknown_order_is_refused(mechs):
    mech = mechs('h2o2')
    param, phi = sample_states(mech, 2, seed=1)
    with pytest.raises(AssertionError):
        oracle.fd_jacobian(mech, True, param, phi, mode=1, order=3)
