"""data.bin initial conditions through the C ABI (read_initial_conditions,
kernel_utils/common/read_initial_conditions.cpp.in:41-116), mirroring the
reference's ric_tester (tests/test_utils/ric_tester.py.in): states written
the way the reference's tests write them read back exactly in C and F
order, and the reference's exit paths are errors."""

import os

import numpy as np
import pytest

from pyjac.core.exceptions import HipRuntimeError
from pyjac.core.states import sample_states, write_data_bin
from pyjac.pywrap import read_initial_conditions, write_data


@pytest.mark.parametrize('order', ['C', 'F'])
@pytest.mark.parametrize('conp', [True, False])
def test_round_trip(mechs, tmp_path, order, conp):
    mech = mechs('gri30')
    n = 257
    param, phi = sample_states(mech, n, seed=5, conp=conp)
    path = write_data_bin(str(tmp_path / 'data.bin'), param, phi)
    assert os.path.getsize(path) == n * (mech.NN + 1) * 8
    # record layout on disk: [T, param, V|P, n_0 ...]
    raw = np.fromfile(path).reshape(n, mech.NN + 1)
    assert np.array_equal(raw[:, 0], phi[:, 0]) and np.array_equal(raw[:, 1], param)
    p_out = np.zeros(n)
    f_out = np.zeros(n * mech.NN)
    read_initial_conditions(path, n, p_out, f_out, order)
    assert np.array_equal(p_out, param)
    assert np.array_equal(f_out, phi.flatten(order))


def test_prefix_of_a_longer_file(mechs, tmp_path):
    mech = mechs('h2o2')
    param, phi = sample_states(mech, 50, seed=6)
    path = write_data_bin(str(tmp_path / 'data.bin'), param, phi)
    p, f = np.zeros(20), np.zeros(20 * mech.NN)
    read_initial_conditions(path, 20, p, f, 'F')
    assert np.array_equal(f, phi[:20].flatten('F'))


def test_errors_instead_of_exit(mechs, tmp_path):
    mech = mechs('h2o2')
    param, phi = sample_states(mech, 10, seed=7)
    path = write_data_bin(str(tmp_path / 'data.bin'), param, phi)
    p, f = np.zeros(11), np.zeros(11 * mech.NN)
    with pytest.raises(HipRuntimeError, match='incorrectly formatted'):
        read_initial_conditions(path, 11, p, f, 'C')
    with pytest.raises(HipRuntimeError, match='could not open'):
        read_initial_conditions(str(tmp_path / 'missing.bin'), 10, p[:10], f[:10 * mech.NN], 'C')


def test_write_data(tmp_path):
    a = np.arange(12.0).reshape(3, 4)
    path = str(tmp_path / 'jac.bin')
    write_data(path, a)
    assert np.array_equal(np.fromfile(path), a.ravel())
    write_data(path, np.asfortranarray(a))
    assert np.array_equal(np.fromfile(path), a.ravel('F'))
