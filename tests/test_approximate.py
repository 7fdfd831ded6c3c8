"""The approximate Jacobian (jac_type approximate): the reference's
determine_jac_inds (core/create_jacobian.py:183-189) does not fill the rows
of species that react with the bath gas, and its kernels keep the exact
values of the stored entries, with the T and extra-variable rows summing
only the stored species entries (the conditional Jacobian reads).  Checked
on the checker: with the exact pattern the restricted Jacobian is the
Jacobian bit for bit; with the approximate pattern every stored species
entry is the exact one, the pattern is a subset, and the T row of each
species column differs from the exact one by exactly the weighted dropped
entries of that column."""

import numpy as np
import pytest

import oracle
from pyjac.core.enum_types import JacobianType
from pyjac.core.states import sample_states
from pyjac.core.chem_model import RU


def _approx(mechs, name):
    from pyjac._paths import mechanism_files
    from pyjac.core.mech_tables import MechTables
    return MechTables.from_file(*mechanism_files(name), jac_type=JacobianType.approximate)


@pytest.mark.parametrize('name', ['gri30', 'test'])
def test_exact_pattern_is_the_jacobian(mechs, name):
    mech = mechs(name)
    param, phi = sample_states(mech, 20, seed=4, conp=True)
    assert np.array_equal(oracle.jacobian_on_pattern(mech, True, param, phi),
                          oracle.jacobian(mech, True, param, phi))


@pytest.mark.parametrize('name,conp', [('gri30', True), ('gri30', False), ('test', True)])
def test_approximate_pattern(mechs, name, conp):
    exact = mechs(name)
    approx = _approx(mechs, name)
    assert approx.nnz < exact.nnz
    ra, ca = approx.sparse_rows_cols('C')
    re_, ce = exact.sparse_rows_cols('C')
    assert set(zip(ra, ca)) <= set(zip(re_, ce))
    n = 16
    param, phi = sample_states(exact, n, seed=5, conp=conp)
    J = oracle.jacobian(exact, conp, param, phi)
    Ja = oracle.jacobian_on_pattern(approx, conp, param, phi)
    mask = np.zeros((exact.NN, exact.NN), dtype=bool)
    mask[ra, ca] = True
    sp = mask.copy()
    sp[:2, :] = False
    assert np.array_equal(Ja[:, sp], J[:, sp])
    # the T row: exact minus approximate = the weighted dropped entries
    energy, heat, _ = oracle.chem_utils(exact, conp, phi)
    T = phi[:, 0]
    P = param if conp else phi[:, 1]
    V = phi[:, 1] if conp else param
    conc = np.empty((n, exact.Ns))
    conc[:, :-1] = phi[:, 2:] / V[:, None]
    conc[:, -1] = P / (RU * T) - conc[:, :-1].sum(axis=1)
    heat_tot = (conc * heat).sum(axis=1)
    wk = energy[:, :-1] - energy[:, -1:] * exact.arrays['mw_post'][None, :]
    dropped = ~mask[2:, 2:]
    diff = -np.einsum('sk,skj->sj', wk, np.where(dropped[None], J[:, 2:, 2:], 0.0)) / \
        (V * heat_tot)[:, None]
    got = J[:, 0, 2:] - Ja[:, 0, 2:]
    scale = np.abs(J[:, 0, 2:]).max(axis=1, keepdims=True)
    assert np.allclose(got, diff, rtol=1e-9, atol=1e-12 * scale.max())
