"""The checker pinned by mathematics: its analytical Jacobian (the
restatement of create_jacobian.py) must equal forward-mode automatic
differentiation of its own species rates (the reference pins its Jacobian
the same way against Adept, tests/test_jacobian.py:132-265), and every
non-zero of the AD Jacobian must lie inside the sparsity pattern
(tests/test_jacobian.py:2354-2410)."""

import numpy as np
import pytest

import oracle
from pyjac.core.states import sample_states


@pytest.mark.parametrize('name', ['h2o2', 'gri30', 'test'])
@pytest.mark.parametrize('conp', [True, False])
def test_jacobian_equals_ad(mechs, name, conp):
    mech = mechs(name)
    param, phi = sample_states(mech, 48, seed=5, T_range=(900, 2400), alpha=0.5, conp=conp)
    J = oracle.jacobian(mech, conp, param, phi)
    Jad = oracle.jacobian_ad(mech, conp, param, phi)
    fro = np.linalg.norm((J - Jad).reshape(48, -1), axis=1) / \
        np.linalg.norm(Jad.reshape(48, -1), axis=1)
    assert fro.max() < 1e-12
    # element-wise on entries that are not cancellation-dominated
    big = np.abs(Jad) > 1e-6 * np.abs(Jad).max(axis=(1, 2), keepdims=True)
    rel = np.abs(J - Jad)[big] / np.abs(Jad)[big]
    assert rel.max() < 1e-9


@pytest.mark.parametrize('name', ['h2o2', 'gri30', 'test'])
def test_pattern_covers_ad_nonzeros(mechs, name):
    mech = mechs(name)
    param, phi = sample_states(mech, 16, seed=9, conp=True)
    Jad = oracle.jacobian_ad(mech, True, param, phi)
    mask = np.zeros((mech.NN, mech.NN), dtype=bool)
    rows, cols = mech.sparse_rows_cols('C')
    mask[rows, cols] = True
    assert not np.any(np.abs(Jad[:, ~mask]) > 0)
