"""Training, gate calibration and evaluation of the species-rate surrogate.

    python -m pyjac.surrogate.train --states 10000000 --epochs 24 \\
        --out pyjac/surrogate/gri30_conv.safetensors

Data: seeded states drawn like the GRI-3.0 CONV workload (T ~ U(1000, 2800)
K, P log-uniform in [1, 50] atm, Dirichlet(0.3) mass fractions, V = 1 m^3),
their production rates from the FP64 species-rate kernel.  Training runs
entirely in the surrogate kernels; the gate threshold is then calibrated
on fresh states so that every accepted state of the calibration set is
within the tolerance of the FP64 kernel.
"""

import argparse
import math
import os
import time

import numpy as np

from . import Surrogate, gather_states, pad_to, scatter_rates, _torch, OUT_PAD
from ..core.states import sample_states

T_RANGE = (1000.0, 2800.0)
ALPHA = 0.3


def host_states(mech, n, seed, conp=False):
    """Seeded states of the workload, C order (host arrays)."""
    return sample_states(mech, n, seed=seed, T_range=T_RANGE, alpha=ALPHA, conp=conp, order='C')


_FP64_KERNELS = {}


def fp64_rates(mech, conp, param, phi, order):
    """dphi of device states from the FP64 species-rate kernel (one kernel
    object per mechanism and options, made on first use: device calls take
    any number of states)."""
    from ..pywrap import PySpeciesRatesKernel
    torch = _torch()
    dphi = torch.empty_like(phi)
    key = (id(mech), bool(conp), order)
    k = _FP64_KERNELS.get(key)
    if k is None or k.mech is not mech:
        # the fastest FP64 kernel: the mechanism-specialised module when built
        k = PySpeciesRatesKernel(int(param.numel()), 0, mech=mech, conp=conp, order=order,
                                 module='auto')
        _FP64_KERNELS[key] = k
    k(param, phi, dphi)
    return dphi


def production_rates(sur, param, phi, dphi, colmajor):
    """wdot of all Ns species of every state, from dphi (FP64 kernel)."""
    torch = _torch()
    n = int(param.numel())
    idx = torch.arange(n, dtype=torch.int64, device=param.device)
    shape = (sur.Ns, n) if colmajor else (n, sur.Ns)
    w = torch.empty(shape, dtype=torch.float64, device=param.device)
    scatter_rates(param, phi, dphi, n, sur.mech.NN, sur.conp, colmajor, sur.mw_ratio, idx, w, n)
    return w


def relative_error(w, ref, colmajor):
    """max_k |w_k - ref_k| / max_k |ref_k| per state."""
    if colmajor:
        w, ref = w.t(), ref.t()
    return (w - ref).abs().amax(dim=1) / ref.abs().amax(dim=1).clamp_min(1e-300)


def train(sur, n_train, epochs, batch=32768, lr=2e-3, seed=1000, log=print):
    """Train on n_train fresh states; returns the per-epoch mean losses."""
    torch = _torch()
    mech = sur.mech
    param_h, phi_h = host_states(mech, n_train, seed, sur.conp)
    dev = sur.device
    param = torch.from_numpy(param_h).to(dev)
    phi = torch.from_numpy(phi_h).to(dev)
    del param_h, phi_h
    dphi = fp64_rates(mech, sur.conp, param, phi, 'C')
    if sur.mu is None:
        k = min(n_train, 200000)
        sur.fit_scales(param[:k].cpu().numpy(), phi[:k].cpu().numpy(), dphi[:k].cpu().numpy())
    y_all = sur.targets(param, phi, dphi, n_train, False)[:n_train]
    del dphi
    steps_per_epoch = max(1, n_train // batch)
    total = steps_per_epoch * epochs
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    history = []
    step = 0
    for ep in range(epochs):
        perm = torch.randperm(n_train, device=dev, generator=gen)
        p_e, f_e, y_e = param[perm], phi[perm], y_all[perm]
        acc = torch.zeros(1, dtype=torch.float32, device=dev)
        for i in range(steps_per_epoch):
            sl = slice(i * batch, (i + 1) * batch)
            warm = min(1.0, (step + 1) / max(1.0, min(200.0, total / 10.0)))
            lr_t = lr * warm * 0.5 * (1.0 + math.cos(math.pi * step / total))
            y = y_e[sl].contiguous()
            acc += sur.train_step(p_e[sl].contiguous(), f_e[sl].contiguous(), y, batch, lr_t)
            step += 1
        history.append(float(acc.item()) / steps_per_epoch)
        if log:
            log('epoch {:3d}  loss {:.3e}'.format(ep, history[-1]))
    return history


def calibrate(sur, n, seed, tol=1e-3, order='F', margin=0.5):
    """Largest gate threshold such that every accepted state of n fresh
    states is within margin * tol of the FP64 kernel (the margin keeps the
    states of other samples the gate accepts within tol).  Returns
    (threshold, accepted fraction)."""
    torch = _torch()
    colmajor = order == 'F'
    param_h, phi_h = host_states(sur.mech, n, seed, sur.conp)
    param = torch.from_numpy(param_h).to(sur.device)
    phi = torch.from_numpy(np.ascontiguousarray(phi_h.T) if colmajor else phi_h).to(sur.device)
    ref = production_rates(sur, param, phi, fp64_rates(sur.mech, sur.conp, param, phi, order),
                           colmajor)
    saved = sur.log10_tol
    sur.log10_tol = 1e30
    out_pred = _head(sur, param, phi, n, colmajor)
    w, _ = sur.predict(param, phi, n)
    sur.log10_tol = saved
    err = relative_error(w, ref, colmajor)
    order_idx = torch.argsort(out_pred)
    bad = (err[order_idx] > margin * tol).to(torch.int64)
    first_bad = int(torch.nonzero(bad)[0].item()) if bool(bad.any()) else n
    if first_bad == 0:
        thr = -1e30
    elif first_bad == n:
        thr = float(out_pred[order_idx[-1]].item()) + 1e-6
    else:
        thr = float(out_pred[order_idx[first_bad]].item())
    sur.log10_tol = thr
    return thr, first_bad / n


def _head(sur, param, phi, n, colmajor):
    """The error head's output for n device states."""
    x, _, M = sur.features(param, phi, n, colmajor)
    out = sur.net.forward(x, M)
    return out[:n, sur.Ns].clone()


def rates_with_fallback(sur, param, phi):
    """Surrogate production rates with the FP64 kernel for rejected states:
    (wdot, accept) of device states in the surrogate's order."""
    torch = _torch()
    colmajor = sur.order == 'F'
    n = int(param.numel())
    # the network and its gate always run, whatever the calibration gave
    w, accept = sur.predict(param, phi, n)
    idx = torch.nonzero(accept == 0).flatten()
    m = int(idx.numel())
    if m:
        p_c, f_c = gather_states(param, phi, n, sur.mech.NN, colmajor, idx)
        d_c = fp64_rates(sur.mech, sur.conp, p_c, f_c, sur.order)
        scatter_rates(p_c, f_c, d_c, m, sur.mech.NN, sur.conp, colmajor, sur.mw_ratio, idx, w, n)
    return w, accept


def evaluate(sur, n, seed, tol=1e-3, reps=5):
    """Accuracy and speed of surrogate + fallback against the FP64 kernel on
    n fresh states (the surrogate's order)."""
    torch = _torch()
    colmajor = sur.order == 'F'
    param_h, phi_h = host_states(sur.mech, n, seed, sur.conp)
    param = torch.from_numpy(param_h).to(sur.device)
    phi = torch.from_numpy(np.ascontiguousarray(phi_h.T) if colmajor else phi_h).to(sur.device)
    from ..pywrap import PySpeciesRatesKernel
    k = PySpeciesRatesKernel(n, 0, mech=sur.mech, conp=sur.conp, order=sur.order,
                             module='auto')
    dphi = torch.empty_like(phi)
    k(param, phi, dphi)
    ref = production_rates(sur, param, phi, dphi, colmajor)
    w, accept = rates_with_fallback(sur, param, phi)
    err = relative_error(w, ref, colmajor)
    acc = accept.to(torch.bool)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    t_plain = timed(lambda: k(param, phi, dphi))
    t_sur = timed(lambda: rates_with_fallback(sur, param, phi))
    t_net = timed(lambda: sur.predict(param, phi, n))
    return {
        'states': n,
        'accepted_fraction': float(acc.float().mean().item()),
        'max_error_all': float(err.max().item()),
        'max_error_accepted_surrogate': float(err[acc].max().item()) if bool(acc.any()) else 0.0,
        'p99_error_network': float(torch.quantile(err[acc].float(), 0.99).item()) if bool(acc.any()) else 0.0,
        'fp64_ms': t_plain,
        'surrogate_with_fallback_ms': t_sur,
        'network_only_ms': t_net,
        'speedup': t_plain / t_sur,
        'fp64_kernel': 'specialised module' if k.module else 'block kernel',
        'tol': tol,
        'log10_gate': sur.log10_tol,
    }


TOLERANCES = (1e-3, 3e-3, 1e-2, 3e-2, 1e-1, 2e-1, 3e-1, 5e-1)


def sweep(sur, n_calib, n_eval, tols=TOLERANCES, calib_seed=2000, eval_seed=3000, reps=5):
    """Calibrate the gate and evaluate surrogate + fallback at each
    tolerance; the gate is left at the first (tightest) tolerance."""
    out = []
    for tol in tols:
        thr, frac = calibrate(sur, n_calib, seed=calib_seed, tol=tol)
        r = evaluate(sur, n_eval, seed=eval_seed, tol=tol, reps=reps)
        r['calibration_accepted'] = frac
        out.append(r)
    calibrate(sur, n_calib, seed=calib_seed, tol=tols[0])
    return out


def error_percentiles(sur, n, seed=3000, q=(0.1, 1, 10, 50, 90)):
    """Percentiles of the network's own error (every state through the
    network, no gate) on n fresh states."""
    torch = _torch()
    colmajor = sur.order == 'F'
    param_h, phi_h = host_states(sur.mech, n, seed, sur.conp)
    param = torch.from_numpy(param_h).to(sur.device)
    phi = torch.from_numpy(np.ascontiguousarray(phi_h.T) if colmajor else phi_h).to(sur.device)
    ref = production_rates(sur, param, phi, fp64_rates(sur.mech, sur.conp, param, phi, sur.order),
                           colmajor)
    saved = sur.log10_tol
    sur.log10_tol = 1e30
    w, _ = sur.predict(param, phi, n)
    sur.log10_tol = saved
    err = relative_error(w, ref, colmajor).cpu().numpy()
    return {'percentiles': list(q), 'error': [float(v) for v in np.percentile(err, q)]}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--states', type=int, default=10_000_000)
    ap.add_argument('--epochs', type=int, default=40)
    ap.add_argument('--batch', type=int, default=32768)
    ap.add_argument('--lr', type=float, default=2e-3)
    ap.add_argument('--calib', type=int, default=262144)
    ap.add_argument('--eval', type=int, default=1_048_576)
    ap.add_argument('--out', type=str, default=None)
    args = ap.parse_args(argv)
    from .._paths import mechanism_files
    from ..core.mech_tables import MechTables
    import json
    mech = MechTables.from_file(*mechanism_files('gri30'))
    sur = Surrogate(mech, conp=False, order='F')
    t0 = time.time()
    hist = train(sur, args.states, args.epochs, args.batch, args.lr)
    _torch().cuda.synchronize()
    t_train = time.time() - t0
    res = {'train_states': args.states, 'epochs': args.epochs, 'train_seconds': t_train,
           'final_loss': hist[-1], 'by_tolerance': sweep(sur, args.calib, args.eval)}
    print(json.dumps(res), flush=True)
    if args.out:
        sur.save(args.out)
    return res


if __name__ == '__main__':
    main()
