"""Benchmark: GRI-Mech 3.0 analytical Jacobians/s and species-rate states/s
on one MI355X.

Workload (the configuration the headline metric is quoted on): GRI-Mech 3.0
(53 species, 325 reactions), constant volume, F (state-fastest) data order,
1,048,576 seeded states (T ~ U(1000, 2800) K, P log-uniform 1-50 atm,
Dirichlet(0.3) mass fractions, V = 1 m^3), species rates plus the sparse
(CCS, 1984 non-zeros) analytical Jacobian, all FP64.  A step is one pass of
the Jacobian and one pass of the species rates over the whole batch, inputs
and outputs resident in HBM, by the mechanism-specialised modules when they
are built (the table-driven block kernels are timed beside them).

Prints one JSON line (see the keys below).  ``--gpus`` above 1 is rejected:
this build runs on one GPU.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, 'pyjac-v2_amd'))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--states', type=int, default=1048576)
    ap.add_argument('--cpu-seconds', type=float, default=15.0,
                    help='approximate CPU work of the CPU-baseline sample')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-surrogate', action='store_true')
    ap.add_argument('--surrogate-states', type=int, default=262144)
    ap.add_argument('--copy-bytes', type=int, default=0,
                    help='run one device-to-device copy of this many bytes first (a kernel of '
                         'known traffic, to calibrate hardware byte counters against)')
    return ap.parse_args()


# counter summaries of the two Jacobian paths (mechanism-specialised module;
# table-driven block kernel)
TRAFFIC_FILE_MODULE = os.path.join(ROOT, 'profiles', 'traffic_ljac_gri30_conv_f_sparse.txt')
TRAFFIC_FILE_BLOCK = os.path.join(ROOT, 'profiles', 'traffic_jacobian_gri30_conv_f_sparse.txt')


def measured_traffic(n, path):
    """HBM bytes per Jacobian call from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE per state, scaled by the copy-kernel factor),
    or None when no summary is there."""
    try:
        for line in open(path):
            if line.startswith('traffic_bytes_per_state'):
                return float(line.split('=')[1]) * n
    except OSError:
        pass
    return None


def cpu_baseline(mech, budget_s):
    """Time the CPU checker (OpenMP, all host cores) on a bounded sample of
    the same workload: Jacobians (which include the species rates)."""
    import oracle
    from pyjac.core.states import sample_states
    threads = os.cpu_count() or 1
    env_threads = os.environ.get('OMP_NUM_THREADS')
    if env_threads and env_threads.isdigit():
        threads = int(env_threads)
    # calibrate on a small sample, then size the timed sample to the budget
    n0 = 2048
    param, phi = sample_states(mech, n0, seed=1234, T_range=(1000.0, 2800.0),
                               alpha=0.3, conp=False)
    oracle.jacobian(mech, False, param, phi, nthreads=threads)
    t0 = time.perf_counter()
    oracle.jacobian(mech, False, param, phi, nthreads=threads)
    dt = time.perf_counter() - t0
    n = int(min(262144, max(n0, n0 * budget_s / max(dt, 1e-6))))
    param, phi = sample_states(mech, n, seed=1235, T_range=(1000.0, 2800.0),
                               alpha=0.3, conp=False)
    t0 = time.perf_counter()
    oracle.jacobian(mech, False, param, phi, nthreads=threads)
    dt = time.perf_counter() - t0
    return {'value': n / dt, 'unit': 'Jacobians/s', 'cores': threads, 'kind': 'port',
            'sample': '{} GRI-3.0 CONV states, analytical Jacobian incl. species rates, '
                      'C++ restatement with OpenMP over states'.format(n)}


SURROGATE_MODEL = os.path.join(ROOT, 'pyjac-v2_amd', 'pyjac', 'surrogate',
                               'gri30_conv.safetensors')


def surrogate_leg(mech, n):
    """The species-rate surrogate with its error gate and FP64 fallback
    against the FP64 species-rate kernel on n fresh states of the same
    workload (F order), the network always in the timed path: at the 1e-3
    tolerance, and at each tolerance of a sweep (gate calibrated on other
    fresh states), reporting the first tolerance at which the gate accepts
    more than half of the states."""
    if not os.path.exists(SURROGATE_MODEL):
        return None
    from pyjac.surrogate import Surrogate
    from pyjac.surrogate.train import error_percentiles, sweep
    sur = Surrogate(mech, conp=False, order='F').load(SURROGATE_MODEL)
    rows = sweep(sur, n, n, reps=3)
    keep = ('tol', 'accepted_fraction', 'speedup', 'max_error_all', 'fp64_ms',
            'surrogate_with_fallback_ms', 'network_only_ms')
    first = rows[0]
    half = next((r for r in rows if r['accepted_fraction'] > 0.5), None)
    out = {k: first[k] for k in keep}
    out.update({'states': n,
                'error': 'max_k |w_k - w_ref_k| / max_k |w_ref_k| per state',
                'network_error': error_percentiles(sur, n),
                'by_tolerance': [{k: r[k] for k in keep} for r in rows],
                'acceptance_above_half': ({k: half[k] for k in keep} if half else None)})
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py runs on one GPU only (this build has no multi-GPU path); '
              'got --gpus {}'.format(args.gpus), file=sys.stderr)
        sys.exit(2)
    import torch
    from pyjac._paths import mechanism_files
    from pyjac.core.mech_tables import MechTables
    from pyjac.core.states import sample_states
    from pyjac.pywrap import PyJacobianKernel, PySpeciesRatesKernel

    dev = torch.device('cuda:0')
    if args.copy_bytes:
        src = torch.empty(args.copy_bytes // 8, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    mech = MechTables.from_file(*mechanism_files('gri30'))
    n = args.states
    param, phi = sample_states(mech, n, seed=20240, T_range=(1000.0, 2800.0),
                               alpha=0.3, conp=False, order='F')
    d_param = torch.from_numpy(param).to(dev)
    d_phi = torch.from_numpy(phi).to(dev)
    d_dphi = torch.empty_like(d_phi)
    d_jac = torch.empty((mech.nnz, n), dtype=torch.float64, device=dev)
    # the Jacobian and the species rates by the mechanism-specialised modules
    # when they are built (__graft_entry__.build()), else by the block
    # kernels; the block kernels are timed beside them either way
    jk = PyJacobianKernel(n, 0, mech=mech, conp=False, order='F', jac_format=1, module='auto')
    jk_block = PyJacobianKernel(n, 0, mech=mech, conp=False, order='F', jac_format=1)
    rk = PySpeciesRatesKernel(n, 0, mech=mech, conp=False, order='F', module='auto')
    rk_block = PySpeciesRatesKernel(n, 0, mech=mech, conp=False, order='F')

    for _ in range(args.warmup):
        jk(d_param, d_phi, d_jac)
        jk_block(d_param, d_phi, d_jac)
        rk(d_param, d_phi, d_dphi)
        rk_block(d_param, d_phi, d_dphi)
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1, e2 in ev:
        e0.record(stream)
        jk(d_param, d_phi, d_jac)
        e1.record(stream)
        rk(d_param, d_phi, d_dphi)
        e2.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    jac_ms = float(np.mean([a.elapsed_time(b) for a, b, _ in ev]))
    rates_ms = float(np.mean([b.elapsed_time(c) for _, b, c in ev]))
    # the block kernels, outside the timed steps
    b0, b1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    b0.record(stream)
    for _ in range(args.steps):
        rk_block(d_param, d_phi, d_dphi)
    b1.record(stream)
    torch.cuda.synchronize()
    rates_block_ms = b0.elapsed_time(b1) / args.steps
    b0.record(stream)
    for _ in range(args.steps):
        jk_block(d_param, d_phi, d_jac)
    b1.record(stream)
    torch.cuda.synchronize()
    jac_block_ms = b0.elapsed_time(b1) / args.steps

    jac_states_s = n / (jac_ms * 1e-3)
    rates_states_s = n / (rates_ms * 1e-3)
    bytes_per_state = (mech.NN + 1) * 8 + mech.nnz * 8
    achieved = bytes_per_state * n / (jac_ms * 1e-3) / 1e9
    ms_per_step = wall * 1e3 / args.steps

    line = {
        'metric': 'GRI-30 Jacobians/s & species-rate states/s, 1 GPU; surrogate speedup @1e-3 err',
        'value': jac_states_s,
        'unit': 'Jacobians/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': ms_per_step,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic seeded states',
        'config': {'workload': 'gri30_conv_f_sparse_jacobian_and_rates',
                   'mechanism': 'GRI-Mech 3.0 (53 species, 325 reactions)',
                   'states': n, 'conp': False, 'order': 'F',
                   'jac_format': 'sparse CCS', 'jac_nnz': mech.nnz},
        'species_rates_states_per_s': rates_states_s,
        'species_rates_kernel': ('pyjac_lane_rates (mechanism-specialised module, one state '
                                 'per lane)' if rk.module else 'rates_block_kernel<1>'),
        'jacobian_kernel': ('pyjac_ljac_a + pyjac_ljac_b1 (mechanism-specialised module, one '
                            'state per lane, chunks of 131072 states)' if jk.jac_module
                            else 'jacobian_block_kernel<2> + entries_to_f2_small_kernel'),
        'jacobian_ms': jac_ms,
        'jacobian_block_kernel_ms': jac_block_ms,
        'jacobian_block_kernel_per_s': n / (jac_block_ms * 1e-3),
        'species_rates_ms': rates_ms,
        'species_rates_block_kernel_ms': rates_block_ms,
        'species_rates_block_kernel_states_per_s': n / (rates_block_ms * 1e-3),
        'roofline': {'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
                     'traffic': measured_traffic(n, TRAFFIC_FILE_MODULE if jk.jac_module
                                                 else TRAFFIC_FILE_BLOCK),
                     'kernel': ('Jacobian path: pyjac_ljac_a (rates, per-reaction values and '
                                'the temperature column into scratch) + pyjac_ljac_b1 (every '
                                'entry, F-order lines), per chunk of 131072 states'
                                if jk.jac_module else
                                'Jacobian path: jacobian_block_kernel<2> (C-order chunks of '
                                '65536 states) + entries_to_f2_small_kernel (transpose to F '
                                'order, on a second stream)'),
                     'algorithmic_bytes_per_state': bytes_per_state},
    }
    if not args.no_surrogate:
        del d_jac, d_dphi
        torch.cuda.empty_cache()
        line['surrogate'] = surrogate_leg(mech, args.surrogate_states)
    if not args.no_cpu:
        line['cpu_baseline'] = cpu_baseline(mech, args.cpu_seconds)
    print(json.dumps(line))


if __name__ == '__main__':
    main()
