"""Condition-aware tolerance of species-rate parity.

A production rate is a sum of reaction contributions that can cancel almost
exactly (reactions near equilibrium, fast forward/reverse pairs).  Two
correctly rounded FP64 evaluations of the same rate constants then differ by
a few ulps of the largest contribution, not of the result: the GPU's exp
and log (ROCm device library) and the CPU's libm round differently in the
last bit, and the cancellation magnifies that.  The bound per element is

    |gpu - ref| <= RTOL |ref| + COND * scale + ATOL

with ``scale`` the sum of the absolute values of the terms the element is
summed from:

* dn_k/dt = V wdot_k: V * sum_r |nu_kr| (|rop_fwd_r| + |rop_rev_r|) |pres_mod_r|;
* dT/dt = -sum_k e_k wdot_k / sum_k c_k C_k: sum_k |e_k| scale(wdot_k) / |sum_k c_k C_k|;
* dV/dt (CONP) or dP/dt (CONV): the same propagation through
  rate_subs.py:709-1076's formula, with |1 - mw_k/mw_ns| weights.

COND = 1e-13 is ~450 ulps of the largest term: the GPU/CPU difference
measured on cancelling elements is ~1e-16 of the term sum.
"""

import numpy as np

import oracle

RTOL, COND, ATOL = 1e-10, 1e-13, 1e-20
RU = 8314.4621  # J/(kmol K), the gas constant of the kernels


def rate_scales(mech, conp, param, phi):
    """(ref, scale): the checker's dphi ``[n, NN]`` and the per-element sum
    of absolute contributions."""
    param = np.ascontiguousarray(param, dtype=np.float64).reshape(-1)
    phi = np.ascontiguousarray(phi, dtype=np.float64).reshape(-1, mech.NN)
    ref, fwd, rev, pmod, _ = oracle.species_rates_full(mech, conp, param, phi)
    A = mech.arrays
    off, spec = A['rxn_off'], A['rxn_spec']
    nu = np.abs(A['nu_prod'].astype(np.float64) - A['nu_reac'])
    n, Ns = phi.shape[0], mech.Ns
    term = (np.abs(fwd) + np.abs(rev)) * np.abs(pmod)            # [n, Nr]
    rxn = np.repeat(np.arange(mech.Nr), np.diff(off))
    sw = np.zeros((n, Ns))
    np.add.at(sw.T, spec, (term[:, rxn] * nu).T)
    T = np.clip(phi[:, 0], 100.0, 1e4)
    if conp:
        P, V = param, np.maximum(1e-30, phi[:, 1])
    else:
        P, V = phi[:, 1], np.maximum(1e-30, param)
    energy, heat, _ = oracle.chem_utils(mech, conp, phi)
    conc = np.empty((n, Ns))
    conc[:, :-1] = phi[:, 2:] / V[:, None]
    conc[:, -1] = P / (RU * T) - conc[:, :-1].sum(axis=1)
    lower = np.abs((conc * heat).sum(axis=1))
    sT = (np.abs(energy) * sw).sum(axis=1) / lower
    w = np.abs(1.0 - mech.arrays['mw_post'])
    sE = (w * sw[:, :-1]).sum(axis=1) * T * RU
    scale = np.empty_like(ref)
    scale[:, 0] = sT
    scale[:, 1] = (V * sT / T + V * sE / P) if conp else (P * sT / T + sE)
    scale[:, 2:] = V[:, None] * sw[:, :-1]
    return ref, scale


def rate_error(got, ref, scale):
    """Largest |got - ref| / bound (<= 1 passes), states in rows."""
    return float(np.max(np.abs(got - ref) / (RTOL * np.abs(ref) + COND * scale + ATOL)))
