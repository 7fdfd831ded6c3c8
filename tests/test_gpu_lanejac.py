"""The mechanism-specialised analytical Jacobian modules (variant 31, and
variant 32: C order through F-order lines and a transpose) on the GPU:
parity with the oracle through the C ABI for every bundled mechanism, CONP
and CONV, both data orders and both formats; the guard states; calls larger
than one scratch chunk; run-to-run reproducibility; the variants and the
autotuner; the refusal of a module generated from other tables.  The modules
are built with the library (__graft_entry__.build())."""

import numpy as np
import pytest

import oracle
from jac_tol import assert_jacobian_close, per_state
from test_gpu_guards import guard_states
from pyjac.core.enum_types import JacobianType
from pyjac.core.exceptions import HipRuntimeError
from pyjac.core.states import sample_states
from pyjac.libgen import lanegen, lanejac
from pyjac.pywrap import PyJacobianKernel

pytestmark = pytest.mark.gpu

CASES = [(m, conp, order, fmt) for m in ('h2o2', 'gri30', 'test')
         for conp in (True, False) for order in ('C', 'F') for fmt in (0, 1)]


def _module(mech):
    path = lanejac.ensure_module(mech, build_missing=False)
    assert path, 'Jacobian module not built (run __graft_entry__.build())'
    return path


def _jac(k, mech, param, phi, order, fmt):
    x = phi if order == 'C' else np.ascontiguousarray(phi.T)
    n = len(param)
    width = mech.NN * mech.NN if fmt == 0 else mech.nnz
    out = np.full(n * width, np.nan)
    k(param, x, out)
    return out


@pytest.mark.parametrize('name,conp,order,fmt', CASES)
def test_module_jacobian_parity(mechs, name, conp, order, fmt):
    mech = mechs(name)
    n = 1500
    param, phi = sample_states(mech, n, seed=41, T_range=(900.0, 2800.0), alpha=0.3, conp=conp)
    k = PyJacobianKernel(n, 0, mech=mech, conp=conp, order=order, jac_format=fmt,
                         module=_module(mech))
    assert k.jac_module and k.lib.pyjac_hip_get_variant(k._h, 1) == 0
    k.set_variant(31)  # the module in every layout (C order with many entries is not automatic)
    got = _jac(k, mech, param, phi, order, fmt)
    ref = oracle.to_layout(mech, oracle.jacobian(mech, conp, param, phi), order, fmt)
    assert_jacobian_close(per_state(got, n, order), per_state(ref.reshape(-1), n, order),
                          '%s conp=%s %s fmt=%d' % (name, conp, order, fmt))


@pytest.mark.parametrize('name,conp,fmt', [(m, conp, fmt) for m in ('h2o2', 'gri30', 'test')
                                            for conp in (True, False) for fmt in (0, 1)])
def test_module_c_order_through_lines(mechs, name, conp, fmt):
    """Variant 32: the F-order store kernel into a chunk buffer, then the
    transpose into C-order rows; the same entries as the oracle."""
    mech = mechs(name)
    n = 1500
    param, phi = sample_states(mech, n, seed=53, T_range=(900.0, 2800.0), alpha=0.3, conp=conp)
    k = PyJacobianKernel(n, 0, mech=mech, conp=conp, order='C', jac_format=fmt,
                         module=_module(mech))
    k.set_variant(32)
    got = _jac(k, mech, param, phi, 'C', fmt)
    ref = oracle.to_layout(mech, oracle.jacobian(mech, conp, param, phi), 'C', fmt)
    assert_jacobian_close(per_state(got, n, 'C'), per_state(ref.reshape(-1), n, 'C'),
                          '%s conp=%s C fmt=%d through lines' % (name, conp, fmt))


def test_module_c_order_lines_two_chunks(mechs):
    """C order through lines over more than one chunk (the second partial):
    states on both sides of the chunk boundary and at both ends against the
    oracle, the automatic choice for GRI-Mech 3.0's 1984 entries, two calls
    equal bit for bit, and variant 31 (each lane storing its own state's
    entries) equal to it bit for bit."""
    import torch
    mech = mechs('gri30')
    n = lanejac.STRIDE + 1003
    param, phi = sample_states(mech, n, seed=59, T_range=(1000.0, 2800.0), alpha=0.3, conp=False)
    dev = torch.device('cuda:0')
    d_param, d_phi = torch.from_numpy(param).to(dev), torch.from_numpy(phi).to(dev)
    k = PyJacobianKernel(n, 0, mech=mech, conp=False, order='C', jac_format=1,
                         module=_module(mech))
    assert k.lib.pyjac_hip_get_variant(k._h, 1) == 0
    a = torch.full((n, mech.nnz), float('nan'), dtype=torch.float64, device=dev)
    b = torch.full_like(a, float('nan'))
    c = torch.full_like(a, float('nan'))
    k(d_param, d_phi, a)
    k(d_param, d_phi, b)
    k.set_variant(31)
    k(d_param, d_phi, c)
    torch.cuda.synchronize()
    assert torch.equal(a, b) and torch.equal(a, c)
    s0 = lanejac.STRIDE
    sel = np.r_[0:200, s0 - 200:s0 + 200, n - 200:n]
    got = a[torch.from_numpy(sel).to(dev)].cpu().numpy()
    ref = oracle.to_layout(mech, oracle.jacobian(mech, False, param[sel], phi[sel]), 'C', 1)
    assert_jacobian_close(got, ref.reshape(len(sel), -1), 'C order through lines, two chunks')


@pytest.mark.parametrize('name', ['h2o2', 'gri30', 'test'])
def test_module_jacobian_guards(mechs, name):
    mech = mechs(name)
    for conp in (True, False):
        param, phi = guard_states(mech, conp)
        n = len(param)
        ref = oracle.jacobian(mech, conp, param, phi)
        fin = np.isfinite(ref.reshape(n, -1)).all(axis=1)
        for order, variant in (('C', 31), ('C', 32), ('F', 31)):
            k = PyJacobianKernel(n, 0, mech=mech, conp=conp, order=order, jac_format=0,
                                 module=_module(mech))
            k.set_variant(variant)
            got = per_state(_jac(k, mech, param, phi, order, 0), n, order)
            want = per_state(oracle.to_layout(mech, ref, order, 0).reshape(-1), n, order)
            assert (np.isfinite(got).all(axis=1) == fin).all()
            assert_jacobian_close(got[fin], want[fin], '%s guards conp=%s %s' % (name, conp, order))


def test_module_two_chunks_reproducible(mechs):
    """Resident F-order GRI states, more than one scratch chunk of
    lanejac.STRIDE states (the second chunk partial): states on both sides of
    the chunk boundary and at both ends against the oracle, and two calls
    equal bit for bit (each entry is summed in a fixed order by one lane)."""
    import torch
    mech = mechs('gri30')
    n = lanejac.STRIDE + 1003
    param, phi = sample_states(mech, n, seed=43, T_range=(1000.0, 2800.0), alpha=0.3,
                               conp=False, order='F')
    dev = torch.device('cuda:0')
    d_param, d_phi = torch.from_numpy(param).to(dev), torch.from_numpy(phi).to(dev)
    k = PyJacobianKernel(n, 0, mech=mech, conp=False, order='F', jac_format=1,
                         module=_module(mech))
    a = torch.full((mech.nnz, n), float('nan'), dtype=torch.float64, device=dev)
    b = torch.full_like(a, float('nan'))
    k(d_param, d_phi, a)
    k(d_param, d_phi, b)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    s0 = lanejac.STRIDE
    sel = np.r_[0:200, s0 - 200:s0 + 200, n - 200:n]
    got = a[:, torch.from_numpy(sel).to(dev)].cpu().numpy().T
    ref = oracle.to_layout(mech, oracle.jacobian(mech, False, param[sel],
                                                 np.ascontiguousarray(phi.T[sel])), 'F', 1).T
    assert_jacobian_close(got, ref, 'two chunks')


def test_module_variant_and_autotune(mechs):
    import torch
    mech = mechs('gri30')
    n = 65536
    param, phi = sample_states(mech, n, seed=47, conp=False, order='F')
    dev = torch.device('cuda:0')
    d_param, d_phi = torch.from_numpy(param).to(dev), torch.from_numpy(phi).to(dev)
    d_out = torch.empty((mech.nnz, n), dtype=torch.float64, device=dev)
    plain = PyJacobianKernel(n, 0, mech=mech, conp=False, order='F', jac_format=1)
    with pytest.raises(HipRuntimeError):
        plain.set_variant(31)  # no module loaded
    with pytest.raises(HipRuntimeError):
        plain.set_variant(32)
    fk = PyJacobianKernel(n, 0, mech=mech, conp=False, order='F', jac_format=1,
                          module=_module(mech))
    with pytest.raises(HipRuntimeError):
        fk.set_variant(32)  # C order only
    k = PyJacobianKernel(n, 0, mech=mech, conp=False, order='F', jac_format=1,
                         module=_module(mech))
    k.set_variant(31)
    v, ms = k.autotune(d_param, d_phi, d_out, cache=False)
    print('autotune with the Jacobian module: variant', v, 'ms', ms)
    d_out.fill_(float('nan'))
    k(d_param, d_phi, d_out)
    torch.cuda.synchronize()
    got = d_out[:, :512].cpu().numpy().T
    ref = oracle.to_layout(mech, oracle.jacobian(mech, False, param[:512],
                                                 np.ascontiguousarray(phi.T[:512])), 'F', 1).T
    assert_jacobian_close(got, ref, 'after autotune')


def test_module_refuses_other_tables(mechs, tmp_path):
    gri, h2 = mechs('gri30'), mechs('h2o2')
    k = PyJacobianKernel(8, 0, mech=gri, conp=True, order='C')
    with pytest.raises(HipRuntimeError, match='other mechanism tables'):
        k.load_module(_module(h2))
    assert k.lib.pyjac_hip_get_variant(k._h, 1) == 0
    ga = mechs('gri30', jac_type=JacobianType.approximate)
    ka = PyJacobianKernel(8, 0, mech=ga, conp=True, order='C',
                          jac_type=JacobianType.approximate)
    with pytest.raises(HipRuntimeError):
        ka.load_module(_module(gri))


def test_auto_loads_both_modules(mechs):
    """module='auto' on a Jacobian kernel attaches the species-rate module
    and the Jacobian module of the tables; a finite-difference kernel only
    the first."""
    mech = mechs('gri30')
    k = PyJacobianKernel(8, 0, mech=mech, conp=True, order='F', module='auto')
    assert k.module == lanegen.ensure_module(mech, build_missing=False)
    assert k.jac_module == lanejac.ensure_module(mech, build_missing=False)
    fd = PyJacobianKernel(8, 0, mech=mech, conp=True, order='F', module='auto',
                          jac_type=JacobianType.finite_difference)
    assert fd.jac_module is None and fd.module
