"""The benchmark workloads at their full sizes against the checker.

* GRI-Mech 3.0, CONV, F order, sparse CCS: all 1,048,576 states of the
  headline workload's own sampler (T ~ U(1000, 2800) K, Dirichlet(0.3)),
  species rates (the table-driven block kernel and the mechanism-specialised
  module) and Jacobians evaluated on the device in one call each (the
  pipelined chunked path), every state compared with the checker (OpenMP)
  in slabs of 65,536 states.
* The ragged batch: H2/O2 200,000 + GRI-Mech 3.0 600,000 + the 150-species
  synthetic mechanism 200,000 states, full Jacobians in F order, one batch
  call replayed from a HIP graph; every state finite and zero outside the
  exact sparsity pattern, and a seeded sample of 512 states per member
  against the checker.
"""

import numpy as np
import pytest

import oracle
from jac_tol import assert_jacobian_close
from rate_tol import rate_error, rate_scales
from pyjac.core.states import sample_states
from pyjac.pywrap import Batch, PyJacobianKernel, PySpeciesRatesKernel

pytestmark = pytest.mark.gpu
torch = pytest.importorskip('torch')

SLAB = 65536


@pytest.mark.timeout(900)
def test_headline_workload_every_state(mechs):
    mech = mechs('gri30')
    n = 1048576
    print('c1: sampling', flush=True)
    param, phi = sample_states(mech, n, seed=20240, T_range=(1000.0, 2800.0), alpha=0.3,
                               conp=False, order='F')
    d_param = torch.from_numpy(param).cuda()
    d_phi = torch.from_numpy(phi).cuda()
    d_jac = torch.empty((mech.nnz, n), dtype=torch.float64, device='cuda')
    d_dphi = torch.empty_like(d_phi)
    jk = PyJacobianKernel(n, 0, mech=mech, conp=False, order='F', jac_format=1)
    rk = PySpeciesRatesKernel(n, 0, mech=mech, conp=False, order='F')
    # the species rates the bench times: the mechanism-specialised module
    rm = PySpeciesRatesKernel(n, 0, mech=mech, conp=False, order='F', module='auto')
    assert rm.module, 'specialised module not built (run __graft_entry__.build())'
    d_dphi_m = torch.empty_like(d_phi)
    jk(d_param, d_phi, d_jac)
    rk(d_param, d_phi, d_dphi)
    rm(d_param, d_phi, d_dphi_m)
    torch.cuda.synchronize()
    _, bad, _ = jk.nonfinite_states(d_jac)
    assert bad == 0
    phi_c = np.ascontiguousarray(phi.T)
    worst_rates = worst_module = 0.0
    for s0 in range(0, n, SLAB):
        sl = slice(s0, s0 + SLAB)
        print('c1: states', s0, flush=True)
        ref, scale = rate_scales(mech, False, param[sl], phi_c[sl])
        got = d_dphi[:, sl].cpu().numpy().T
        worst_rates = max(worst_rates, rate_error(got, ref, scale))
        worst_module = max(worst_module, rate_error(d_dphi_m[:, sl].cpu().numpy().T, ref, scale))
        jref = oracle.jacobian(mech, False, param[sl], phi_c[sl])
        rows, cols = mech.sparse_rows_cols('F')
        assert_jacobian_close(d_jac[:, sl].cpu().numpy().T, jref[:, rows, cols],
                              'c1 states {}..'.format(s0))
    print('c1 species rates: block kernel %.3g, specialised module %.3g of the bound'
          % (worst_rates, worst_module))
    assert worst_rates <= 1.0, worst_rates
    assert worst_module <= 1.0, worst_module


MEMBERS = [('h2o2', 200000), ('gri30', 600000), ('s150', 200000)]


@pytest.mark.timeout(900)
@pytest.mark.parametrize('one_launch', [False, True])
def test_ragged_batch_at_config_sizes(mechs, one_launch):
    kernels, arrays, host = [], [], []
    for i, (name, n) in enumerate(MEMBERS):
        mech = mechs(name)
        print('c2: member', name, n, flush=True)
        param, phi = sample_states(mech, n, seed=300 + i, conp=False, order='F')
        k = PyJacobianKernel(n, 0, mech=mech, conp=False, order='F', jac_format=0)
        out = torch.empty((mech.NN, mech.NN, n), dtype=torch.float64, device='cuda')
        out.fill_(np.nan)
        kernels.append(k)
        arrays.append((torch.from_numpy(param).cuda(), torch.from_numpy(phi).cuda(), out))
        host.append((mech, param, phi))
    b = Batch(kernels)
    b.set_queue(one_launch)
    for i, (p, f, o) in enumerate(arrays):
        b.set(i, p, f, o)
    b.run(graph=True)   # capture and replay
    torch.cuda.synchronize()
    for k, (_, _, o) in zip(kernels, arrays):
        o.fill_(np.nan)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    b.run(graph=True)   # replay
    e1.record()
    torch.cuda.synchronize()
    print('c2: one launch' if one_launch else 'c2: per-variant launches',
          'replay %.1f ms' % e0.elapsed_time(e1), flush=True)
    assert b.slot_failures() == 0
    rng = np.random.default_rng(99)
    for (name, n), k, (mech, param, phi), (_, _, out) in zip(MEMBERS, kernels, host, arrays):
        print('c2: checking', name, flush=True)
        _, bad, first = k.nonfinite_states(out, width=mech.NN * mech.NN)
        assert bad == 0, (name, bad, first)
        # entries outside the exact pattern are zero ([col][row][state])
        rows, cols = mech.sparse_rows_cols('C')
        mask = np.ones((mech.NN, mech.NN), dtype=bool)
        mask[rows, cols] = False          # mask[row, col]: not stored
        off = torch.from_numpy(np.ascontiguousarray(mask.T)).cuda()   # [col][row]
        for c0 in range(0, mech.NN, 8):
            blk = out[c0:c0 + 8]
            m = off[c0:c0 + 8]
            assert bool((blk[m] == 0).all()), (name, c0)
        pick = np.sort(rng.choice(n, 512, replace=False))
        idx = torch.from_numpy(pick).cuda()
        got = out.index_select(2, idx).cpu().numpy()          # [col][row][512]
        got = np.transpose(got, (2, 1, 0)).reshape(512, -1)   # [state][row * NN + col]
        phi_c = np.ascontiguousarray(phi[:, pick].T)
        jref = oracle.jacobian(mech, False, param[pick], phi_c).reshape(512, -1)
        assert_jacobian_close(got, jref, 'c2 ' + name)
