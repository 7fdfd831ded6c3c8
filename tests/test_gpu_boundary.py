"""The C ABI's boundary behaviour on the GPU: introspection
(Kernel::speciesNames / reactionStrings / order / vectorWidth,
kernel_utils/common/kernel.hpp.in:101-147), size checks of the host and
device entry points (an undersized array is an error before anything is
written), non-finite inputs rejected, and the per-state finiteness report."""

import ctypes

import numpy as np
import pytest

from pyjac.core.exceptions import HipRuntimeError
from pyjac.core.states import sample_states
from pyjac.pywrap import (PyChemUtilsKernel, PyJacobianKernel, PySpeciesRatesKernel,
                          load_library)

pytestmark = pytest.mark.gpu
torch = pytest.importorskip('torch')

PYJAC_ERR_ARG = 1


def test_introspection(mechs):
    mech = mechs('gri30')
    for order in ('C', 'F'):
        k = PyJacobianKernel(8, 0, mech=mech, conp=False, order=order, jac_format=1)
        assert k.species_names() == mech.species_names
        assert k.species_names()[-1] == 'N2'  # the bath gas last
        assert k.reaction_strings() == mech.reaction_strings
        assert k.lib.pyjac_hip_order(k._h) == order.encode()
        assert k.vector_width() == 1
        assert k.lib.pyjac_hip_species_name(k._h, mech.Ns) is None
        assert k.lib.pyjac_hip_reaction_string(k._h, -1) is None


def test_host_size_checks(mechs):
    mech = mechs('h2o2')
    n = 16
    param, phi = sample_states(mech, n, seed=1, conp=True)
    rk = PySpeciesRatesKernel(n, 0, mech=mech, conp=True, order='C')
    with pytest.raises(ValueError, match='output'):
        rk(param, phi, np.zeros((n - 1, mech.NN)))
    with pytest.raises(ValueError, match='phi'):
        rk(param, phi[:-1], np.zeros_like(phi))
    jk = PyJacobianKernel(n, 0, mech=mech, conp=True, order='C', jac_format=0)
    with pytest.raises(ValueError, match='output'):
        jk(param, phi, np.zeros((n, mech.NN)))
    ck = PyChemUtilsKernel(n, 0, mech=mech, conp=True, order='C')
    with pytest.raises(ValueError, match='specific_heat'):
        ck(phi, np.zeros((n, mech.Ns)), np.zeros((n, mech.Ns)), np.zeros((n, 2)))
    # the C entry points check the element counts themselves: a short
    # output is PYJAC_ERR_ARG and nothing is written
    L = load_library()
    out = np.full(n * mech.NN, 7.0)
    vp = lambda a: ctypes.c_void_p(a.ctypes.data)
    st = L.pyjac_hip_species_rates_sized(rk._h, n, vp(param), n, vp(phi), phi.size, vp(out),
                                         out.size - 1)
    assert st == PYJAC_ERR_ARG and b'dphi holds' in L.pyjac_hip_last_error()
    assert (out == 7.0).all()
    st = L.pyjac_hip_jacobian_sized(jk._h, n, vp(param), n - 1, vp(phi), phi.size, vp(out),
                                    out.size)
    assert st == PYJAC_ERR_ARG and b'param holds' in L.pyjac_hip_last_error()
    # and the right counts run, resizing the handle to n_states
    jac = np.zeros(n * mech.NN * mech.NN)
    assert L.pyjac_hip_jacobian_sized(jk._h, n // 2, vp(param), n, vp(phi), phi.size, vp(jac),
                                      jac.size) == 0
    assert np.isfinite(jac[:n // 2 * mech.NN ** 2]).all()
    assert (jac[n // 2 * mech.NN ** 2:] == 0).all()


def test_device_size_checks(mechs):
    mech = mechs('gri30')
    n = 64
    param, phi = sample_states(mech, n, seed=2, conp=False, order='F')
    p = torch.from_numpy(param).cuda()
    f = torch.from_numpy(phi).cuda()
    jk = PyJacobianKernel(n, 0, mech=mech, conp=False, order='F', jac_format=1)
    with pytest.raises(ValueError, match='output'):
        jk(p, f, torch.empty(mech.nnz * n - 1, dtype=torch.float64, device='cuda'))
    with pytest.raises(ValueError, match='phi'):
        jk(p, f[:, :-1].contiguous(), torch.empty(mech.nnz * n, dtype=torch.float64,
                                                 device='cuda'))


def test_nonfinite_inputs_rejected(mechs):
    mech = mechs('gri30')
    n = 32
    param, phi = sample_states(mech, n, seed=3, conp=False)
    phi[5, 7] = np.nan
    rk = PySpeciesRatesKernel(n, 0, mech=mech, conp=False, order='C')
    out = np.full_like(phi, 3.0)
    with pytest.raises(HipRuntimeError, match='state 5'):
        rk(param, phi, out)
    assert (out == 3.0).all()
    param2 = param.copy()
    param2[9] = np.inf
    phi[5, 7] = 1e-3
    with pytest.raises(HipRuntimeError, match='param of state 9'):
        rk(param2, phi, out)
    # device arrays: checked when the handle's input check is on
    phif = np.ascontiguousarray(phi.T)
    phif[0, 11] = np.nan
    kf = PyJacobianKernel(n, 0, mech=mech, conp=False, order='F', jac_format=1)
    p = torch.from_numpy(param).cuda()
    f = torch.from_numpy(phif).cuda()
    jac = torch.zeros(mech.nnz * n, dtype=torch.float64, device='cuda')
    kf.set_input_check(True)
    with pytest.raises(HipRuntimeError, match='state 11'):
        kf(p, f, jac)
    assert bool((jac == 0).all())


def test_nonfinite_report(mechs):
    """Per-state flags of a device array in C and F order."""
    mech = mechs('h2o2')
    n = 300
    for order in ('C', 'F'):
        k = PySpeciesRatesKernel(n, 0, mech=mech, conp=True, order=order)
        a = torch.ones((n, mech.NN) if order == 'C' else (mech.NN, n), dtype=torch.float64,
                       device='cuda')
        bad = [3, 64, 65, 299]
        for j, s in enumerate(bad):
            v = [np.nan, np.inf, -np.inf, np.nan][j]
            if order == 'C':
                a[s, j % mech.NN] = v
            else:
                a[j % mech.NN, s] = v
        flags, count, first = k.nonfinite_states(a)
        assert count == len(bad) and first == 3
        want = np.zeros(n, dtype=np.int32)
        want[bad] = 1
        assert np.array_equal(flags.cpu().numpy(), want)
        flags, count, first = k.nonfinite_states(torch.ones_like(a))
        assert count == 0 and first == n and int(flags.sum()) == 0
