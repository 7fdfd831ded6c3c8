"""The Jacobian entry point that allocates nothing
(pyjac_hip_jacobian_dev_noalloc, include/pyjac_hip.h), through the C ABI,
against the regular device-pointer entry point on the same states: two calls
in a row on one handle, the second with fewer states, for the block-kernel
F-order path (C-order chunks through the fixed scratch area and a transpose
per chunk), the C-order path (no scratch) and the variant that gathers the
Jacobian in device memory (its contribution buffer in the scratch area);
sizes whose 64-state buffers exceed the area are refused."""

import ctypes

import numpy as np
import pytest

from jac_tol import assert_jacobian_close
from pyjac.core.exceptions import HipRuntimeError
from pyjac.core.states import sample_states
from pyjac.pywrap import PyJacobianKernel, check, load_library

pytestmark = pytest.mark.gpu
torch = pytest.importorskip('torch')

PYJAC_ERR_UNSUPPORTED = 4


def _lib():
    L = load_library()
    P, S = ctypes.c_void_p, ctypes.c_size_t
    L.pyjac_hip_jacobian_dev_noalloc.argtypes = [P, S, P, P, P, P]
    return L


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _run_both(k, mech, n, order, fmt, seed):
    L = _lib()
    param, phi = sample_states(mech, n, seed=seed, conp=False, order=order)
    p = torch.from_numpy(param).cuda()
    f = torch.from_numpy(phi).cuda()
    nv = mech.nnz if fmt else mech.NN * mech.NN
    ref = torch.full((nv * n,), float('nan'), dtype=torch.float64, device='cuda')
    got = torch.full_like(ref, float('nan'))
    check(L.pyjac_hip_jacobian_dev(k._h, n, _vp(p), _vp(f), _vp(ref), _stream()))
    check(L.pyjac_hip_jacobian_dev_noalloc(k._h, n, _vp(p), _vp(f), _vp(got), _stream()))
    torch.cuda.synchronize()
    a = got.view(n, -1) if order == 'C' else got.view(-1, n).t()
    b = ref.view(n, -1) if order == 'C' else ref.view(-1, n).t()
    assert_jacobian_close(a.cpu().numpy(), b.cpu().numpy(),
                          'no-alloc {} n={} {} fmt={}'.format(mech.name, n, order, fmt))


@pytest.mark.parametrize('name,order,fmt,sizes', [
    ('gri30', 'F', 1, (20000, 5001)),    # chunks of 4224 states, last partial; odd n
    ('gri30', 'C', 0, (3000, 1000)),     # no scratch in C order
    ('s150', 'F', 1, (700, 130)),        # device-memory gather variant, chunked
    ('s150', 'C', 1, (300, 64)),
])
def test_noalloc_matches_regular(mechs, name, order, fmt, sizes):
    mech = mechs(name)
    k = PyJacobianKernel(sizes[0], 0, mech=mech, conp=False, order=order, jac_format=fmt)
    for i, n in enumerate(sizes):
        print('no-alloc', name, order, fmt, n, flush=True)
        _run_both(k, mech, n, order, fmt, seed=60 + i)


def test_noalloc_with_jacobian_module(mechs):
    """With the mechanism-specialised Jacobian module loaded, the regular
    entry point runs the module and the one that allocates nothing runs the
    block kernels (the module's scratch array does not fit the fixed area):
    the two agree, twice in a row, the second call smaller."""
    from pyjac.libgen import lanejac
    mech = mechs('gri30')
    path = lanejac.ensure_module(mech, build_missing=False)
    if not path:
        pytest.skip('Jacobian module not built')
    k = PyJacobianKernel(9000, 0, mech=mech, conp=False, order='F', jac_format=1, module=path)
    assert k.jac_module
    for i, n in enumerate((9000, 1001)):
        print('no-alloc with module', n, flush=True)
        _run_both(k, mech, n, 'F', 1, seed=80 + i)


def test_noalloc_refuses_oversized(mechs):
    """A full-format F-order Jacobian of 500 species needs 2 MB per state
    of chunk buffer: 64 states exceed the 64 MiB area."""
    mech = mechs('s500')
    k = PyJacobianKernel(64, 0, mech=mech, conp=True, order='F', jac_format=0)
    L = _lib()
    n = 64
    p = torch.ones(n, dtype=torch.float64, device='cuda')
    f = torch.ones(mech.NN * n, dtype=torch.float64, device='cuda')
    out = torch.zeros(mech.NN * mech.NN * n, dtype=torch.float64, device='cuda')
    st = L.pyjac_hip_jacobian_dev_noalloc(k._h, n, _vp(p), _vp(f), _vp(out), _stream())
    assert st == PYJAC_ERR_UNSUPPORTED
    assert b'scratch' in L.pyjac_hip_last_error()
    torch.cuda.synchronize()
    assert bool((out == 0).all())
