"""The mechanism-specialised species-rate modules (variant 30) on the GPU:
parity with the oracle through the C ABI for every bundled mechanism, option
set and data order, the guard states, run-to-run reproducibility, the
autotuner's and the finite-difference Jacobian's use of the module, and the
refusal of a module generated from other tables.  The modules are built with
the library (__graft_entry__.build())."""

import numpy as np
import pytest

import oracle
from rate_tol import rate_error, rate_scales
from test_gpu_guards import guard_states
from pyjac.core.enum_types import JacobianType
from pyjac.core.exceptions import HipRuntimeError
from pyjac.core.states import sample_states
from pyjac.libgen import lanegen
from pyjac.pywrap import PyJacobianKernel, PySpeciesRatesKernel

pytestmark = pytest.mark.gpu

CASES = [(m, conp, order) for m in ('h2o2', 'gri30', 'test')
         for conp in (True, False) for order in ('C', 'F')]


def _module(mech):
    path = lanegen.ensure_module(mech, build_missing=False)
    assert path, 'module not built (run __graft_entry__.build())'
    return path


def _rates(k, param, phi, order):
    x = phi if order == 'C' else np.ascontiguousarray(phi.T)
    out = np.full_like(x, np.nan)
    k(param, x, out)
    return out if order == 'C' else np.ascontiguousarray(out.T)


@pytest.mark.parametrize('name,conp,order', CASES)
def test_module_rates_parity(mechs, name, conp, order):
    mech = mechs(name)
    n = 2000
    param, phi = sample_states(mech, n, seed=17, T_range=(900.0, 2800.0), alpha=0.3, conp=conp)
    k = PySpeciesRatesKernel(n, 0, mech=mech, conp=conp, order=order, module=_module(mech))
    assert k.lib.pyjac_hip_get_variant(k._h, 0) in (0, 30)
    got = _rates(k, param, phi, order)
    ref, scale = rate_scales(mech, conp, param, phi)
    err = rate_error(got, ref, scale)
    plain = np.abs(got - ref) > 1e-10 * np.abs(ref) + 1e-20
    print(name, conp, order, 'error / bound', err, 'elements over the plain bound',
          int(plain.sum()), 'of', plain.size)
    assert np.isfinite(got).all()
    assert err <= 1.0


@pytest.mark.parametrize('name', ['h2o2', 'gri30', 'test'])
def test_module_guards(mechs, name):
    mech = mechs(name)
    for conp in (True, False):
        param, phi = guard_states(mech, conp)
        for order in ('C', 'F'):
            k = PySpeciesRatesKernel(len(param), 0, mech=mech, conp=conp, order=order,
                                     module=_module(mech))
            got = _rates(k, param, phi, order)
            ref, scale = rate_scales(mech, conp, param, phi)
            assert np.isfinite(got).all()
            assert rate_error(got, ref, scale) <= 1.0, (name, conp, order)


def test_module_device_arrays_and_reproducible(mechs):
    """Resident F-order states through the device entry point: two calls
    agree bit for bit (each lane sums its production rates in reaction
    order), and an odd state count covers the partial last workgroup."""
    import torch
    mech = mechs('gri30')
    n = 100003
    param, phi = sample_states(mech, n, seed=23, T_range=(1000.0, 2800.0), alpha=0.3,
                               conp=False, order='F')
    dev = torch.device('cuda:0')
    d_param = torch.from_numpy(param).to(dev)
    d_phi = torch.from_numpy(phi).to(dev)
    k = PySpeciesRatesKernel(n, 0, mech=mech, conp=False, order='F', module=_module(mech))
    a = torch.full_like(d_phi, float('nan'))
    b = torch.full_like(d_phi, float('nan'))
    k(d_param, d_phi, a)
    k(d_param, d_phi, b)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    got = a.cpu().numpy().T
    sel = np.r_[0:256, n - 256:n]
    ref, scale = rate_scales(mech, False, param[sel], np.ascontiguousarray(phi.T[sel]))
    assert rate_error(got[sel], ref, scale) <= 1.0


def test_module_explicit_variant_and_autotune(mechs):
    import torch
    mech = mechs('gri30')
    n = 65536
    param, phi = sample_states(mech, n, seed=29, conp=True, order='F')
    dev = torch.device('cuda:0')
    d_param, d_phi = torch.from_numpy(param).to(dev), torch.from_numpy(phi).to(dev)
    d_out = torch.empty_like(d_phi)
    plain = PySpeciesRatesKernel(n, 0, mech=mech, conp=True, order='F')
    with pytest.raises(HipRuntimeError):
        plain.set_variant(30)  # no module loaded
    k = PySpeciesRatesKernel(n, 0, mech=mech, conp=True, order='F', module=_module(mech))
    k.set_variant(30)
    v, ms = k.autotune(d_param, d_phi, d_out, cache=False)
    print('autotune with the module: variant', v, 'ms', ms)
    k(d_param, d_phi, d_out)
    torch.cuda.synchronize()
    got = d_out.cpu().numpy().T[:512]
    ref, scale = rate_scales(mech, True, param[:512], np.ascontiguousarray(phi.T[:512]))
    assert rate_error(got, ref, scale) <= 1.0


def test_module_refuses_other_tables(mechs, tmp_path):
    gri, h2 = mechs('gri30'), mechs('h2o2')
    k = PySpeciesRatesKernel(8, 0, mech=gri, conp=True, order='C')
    with pytest.raises(HipRuntimeError, match='other mechanism tables'):
        k.load_module(_module(h2))
    with pytest.raises(HipRuntimeError, match='could not load'):
        k.load_module(str(tmp_path / 'missing.hsaco'))
    assert k.lib.pyjac_hip_get_variant(k._h, 0) == 0


def test_fd_jacobian_with_module(mechs):
    """The finite-difference Jacobian evaluates its perturbed states with the
    module when one is loaded: the bounds of tests/test_gpu_fd.py against the
    oracle's analytical Jacobian and its own difference quotient."""
    mech = mechs('gri30')
    n = 200
    conp, order, fmt, mode, fd_order, bound = True, 'C', 1, 2, 2, 5e-3
    param, phi = sample_states(mech, n, seed=31, conp=conp)
    k = PyJacobianKernel(n, 0, mech=mech, conp=conp, order=order, jac_format=fmt,
                         jac_type=JacobianType.finite_difference, module=_module(mech))
    k.set_fd_options(mode, fd_order)
    fd_ref = oracle.to_layout(mech, oracle.fd_jacobian(mech, conp, param, phi, mode=mode,
                                                       order=fd_order), order, fmt)
    an_ref = oracle.to_layout(mech, oracle.jacobian(mech, conp, param, phi), order, fmt)
    out = np.zeros_like(fd_ref)
    k(param, phi, out)
    assert np.all(np.isfinite(out))
    a, b, c = out.reshape(n, -1), fd_ref.reshape(n, -1), an_ref.reshape(n, -1)
    fro_an = np.linalg.norm(a - c, axis=1) / np.linalg.norm(c, axis=1)
    own = np.linalg.norm(b - c, axis=1) / np.linalg.norm(c, axis=1)
    print('FD with module: 95th percentile error', np.percentile(fro_an, 95))
    assert np.percentile(fro_an, 95) <= bound
    assert np.median(fro_an) <= 2.0 * np.median(own) + 1e-4
    assert np.all(fro_an <= np.maximum(10.0 * bound, 10.0 * own))
