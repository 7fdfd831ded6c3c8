"""Ragged batches over several mechanisms in one launch (and one HIP graph),
and the autotuner: results equal the single-handle calls."""

import numpy as np
import pytest

import oracle
from rate_tol import rate_error, rate_scales
from pyjac.core.states import sample_states
from pyjac.pywrap import Batch, PyJacobianKernel, PySpeciesRatesKernel

pytestmark = pytest.mark.gpu
torch = pytest.importorskip('torch')

MEMBERS = [('h2o2', True, 'C', 0, 3000), ('gri30', False, 'F', 1, 2500),
           ('test', False, 'C', 1, 1111), ('s150', True, 'F', 0, 300)]


def _device_states(mech, n, conp, order, seed):
    param, phi = sample_states(mech, n, seed=seed, conp=conp)
    ph = phi if order == 'C' else np.ascontiguousarray(phi.T)
    return torch.from_numpy(param).cuda(), torch.from_numpy(ph).cuda()


def _setup(mechs, jac):
    kernels, arrays = [], []
    for i, (name, conp, order, fmt, n) in enumerate(MEMBERS):
        mech = mechs(name)
        K = PyJacobianKernel if jac else PySpeciesRatesKernel
        k = K(n, 0, mech=mech, conp=conp, order=order, jac_format=fmt)
        p, f = _device_states(mech, n, conp, order, seed=40 + i)
        nv = (mech.nnz if fmt else mech.NN ** 2) if jac else mech.NN
        out = torch.empty(nv * n, dtype=torch.float64, device='cuda')
        kernels.append(k)
        arrays.append((p, f, out))
    return kernels, arrays


# states whose entries come this close to the largest double overflow or not
# depending on the order of their sums
NEAR_OVERFLOW = 1e300


def _same(o, r, jac, n=None, order='C'):
    """Rates: bit-identical.  Jacobians: the dense kernel adds into LDS with
    atomics whose order depends on the waves per state (and varies from run
    to run with more than one), so each entry agrees to 1e-13 of the
    member's largest entry instead.  test.inp's random states include some
    whose entries reach 1e304 or overflow (the checker's too); whether such a
    state overflows depends on the order of its sums, so states near
    overflow in either result need only be near overflow in both."""
    if not jac or torch.equal(o, r):
        return torch.equal(o, r)
    if n is not None:
        rs = r.view(n, -1) if order == 'C' else r.view(-1, n).t()
        os_ = o.view(n, -1) if order == 'C' else o.view(-1, n).t()

        def near(a):
            big = torch.where(torch.isfinite(a), a.abs(), torch.full_like(a, float('inf')))
            return big.max(dim=1).values > NEAR_OVERFLOW
        nr, no = near(rs), near(os_)
        if not bool((nr == no).all()):
            # one result near overflow, the other not: a real difference
            return False
        r, o = rs[~nr], os_[~nr]
    diff = o != r
    if not bool(torch.isfinite(r[diff]).all()):
        return False  # a non-finite entry that differs
    d = (o[diff] - r[diff]).abs().max().item()
    return d <= 1e-13 * r[torch.isfinite(r)].abs().max().item()


def _eager(kernels, arrays):
    res = []
    for k, (p, f, _) in zip(kernels, arrays):
        o = torch.empty_like(arrays[kernels.index(k)][2])
        k(p, f, o)
        res.append(o)
    torch.cuda.synchronize()
    return res


@pytest.mark.parametrize('jac,one_launch', [(False, False), (True, False), (True, True)])
def test_batch_equals_single_calls(mechs, jac, one_launch):
    kernels, arrays = _setup(mechs, jac)
    ref = _eager(kernels, arrays)
    b = Batch(kernels)
    if one_launch:
        b.set_queue(True)
    for i, (p, f, o) in enumerate(arrays):
        b.set(i, p, f, o)
    for graph in (False, True):
        for o in (a[2] for a in arrays):
            o.fill_(np.nan)
        b.run(graph=graph)
        torch.cuda.synchronize()
        for (_, _, o), r, mem in zip(arrays, ref, MEMBERS):
            assert _same(o, r, jac, mem[4], mem[2]), mem


@pytest.mark.parametrize('one_launch', [False, True])
def test_graph_replays_follow_new_inputs(mechs, one_launch):
    """Replays of the captured graph, with eager work in between, see the
    inputs as they are at replay time."""
    kernels, arrays = _setup(mechs, True)
    b = Batch(kernels)
    b.set_queue(one_launch)
    for i, (p, f, o) in enumerate(arrays):
        b.set(i, p, f, o)
    for rep in range(3):
        # perturb the temperatures in place (eager work), then replay
        for (p, f, o), (name, conp, order, fmt, n) in zip(arrays, MEMBERS):
            if order == 'C':
                f.view(n, -1)[:, 0] *= 1.0 + 0.01 * (rep + 1)
            else:
                f.view(-1, n)[0, :] *= 1.0 + 0.01 * (rep + 1)
        torch.cuda.synchronize()
        ref = _eager(kernels, arrays)
        b.run(graph=True)
        torch.cuda.synchronize()
        for (_, _, o), r, mem in zip(arrays, ref, MEMBERS):
            assert _same(o, r, True, mem[4], mem[2]), (rep, mem)


def test_autotune_keeps_results_right(mechs, tmp_path, monkeypatch):
    """The autotuner's choice (kernel family, waves per state, occupancy
    target and, for the F-order Jacobian, the chunk size) is cached, picked
    up by a new kernel of the same key, and gives results within the parity
    bounds."""
    monkeypatch.setenv('PYJAC_HIP_TUNE_CACHE', str(tmp_path / 'tune.json'))
    mech = mechs('gri30')
    n = 16384
    p, f = _device_states(mech, n, False, 'F', seed=77)
    for K, nv in ((PySpeciesRatesKernel, mech.NN), (PyJacobianKernel, mech.nnz)):
        k = K(n, 0, mech=mech, conp=False, order='F', jac_format=1)
        out = torch.empty(nv * n, dtype=torch.float64, device='cuda')
        v, ms = k.autotune(p, f, out)
        assert v in (1, 2, 11, 12, 14, 18, 22, 24, 28) and ms > 0
        assert (tmp_path / 'tune.json').exists()
        tuned = k.tuning()
        print('autotune', K.__name__, 'variant', v, 'ms', ms, '(occupancy, chunk, waves)', tuned)
        k2 = K(n, 0, mech=mech, conp=False, order='F', jac_format=1)  # picks the tuned choice
        assert k2.tuning() == tuned
        out2 = torch.empty_like(out)
        k2(p, f, out2)
        torch.cuda.synchronize()
        param = p.cpu().numpy()
        phi = f.cpu().numpy().T.copy()
        if K is PySpeciesRatesKernel:
            ref, scale = rate_scales(mech, False, param, phi)
            got = out2.view(-1, n).cpu().numpy().T
            assert rate_error(got, ref, scale) <= 1.0
        else:
            m = 256
            ref = oracle.to_layout(mech, oracle.jacobian(mech, False, param[:m], phi[:m]), 'F', 1)
            got = out2.view(-1, n).cpu().numpy()[:, :m]
            fro = np.linalg.norm(got - ref, axis=0) / np.linalg.norm(ref, axis=0)
            assert fro.max() <= 1e-8
