"""The CPU checker's species rates against the textbook and against invariants.

The checker restates the reference's kernels on the flattened rate tables.
This file computes the same quantities a second, independent way: numpy
straight from the parsed reactions and the standard definitions --

* mass action with k = A T^b exp(-Ta/T);
* Kc = exp(-sum nu g/RT) (P_atm / R T)^(sum nu), with h/RT, s/R, cp/R from
  the NASA-7 polynomials in their textbook form (not the reference's
  combined Gibbs polynomial);
* third-body concentrations [M] = sum eff_k C_k, Lindemann / Troe / SRI
  blending in the Gilbert form, chemically activated k0 F / (1 + Pr);
* PLOG: ln k linear in ln P between the bracketing pressures, clamped
  outside; Chebyshev: log10 k = sum a_tp T_t(T~) T_p(P~) with numpy's
  Chebyshev series (numpy.polynomial.chebyshev.chebval2d);
* dT/dt from the energy balance, dV/dt (CONP) and dP/dt (CONV) from the
  ideal-gas law with the bath gas eliminated by mass conservation --

and checks conservation laws on the checker's output directly: element
and mass conservation of the production rates, and equal production
rates for CONP and CONV at the same (T, P, V, n).
"""

import numpy as np
import pytest
from numpy.polynomial import chebyshev as npcheb

import oracle
from pyjac.core.states import sample_states

RU = 8314.4621
PA = 101325.0


def _thermo(specs, T):
    """cp/R, h/RT, s/R for every species at temperatures T (n, Ns)."""
    lo = np.array([s.lo for s in specs])
    hi = np.array([s.hi for s in specs])
    Tmid = np.array([s.Trange[1] for s in specs])
    a = np.where((T[:, None] < Tmid[None, :])[:, :, None], lo[None], hi[None])
    T1 = T[:, None]
    cp = a[..., 0] + a[..., 1] * T1 + a[..., 2] * T1**2 + a[..., 3] * T1**3 + a[..., 4] * T1**4
    h = (a[..., 0] + a[..., 1] * T1 / 2 + a[..., 2] * T1**2 / 3 + a[..., 3] * T1**3 / 4 +
         a[..., 4] * T1**4 / 5 + a[..., 5] / T1)
    s = (a[..., 0] * np.log(T1) + a[..., 1] * T1 + a[..., 2] * T1**2 / 2 + a[..., 3] * T1**3 / 3 +
         a[..., 4] * T1**4 / 4 + a[..., 6])
    return cp, h, s


def _arr(A, b, Ta, T):
    return A * T**b * np.exp(-Ta / T)


def _falloff_F(r, Pr, T):
    """Blending function of a falloff / chemically activated reaction."""
    if r.troe:
        alpha, T3, T1 = r.troe_par[:3]
        Fcent = (1 - alpha) * np.exp(-T / T3) + alpha * np.exp(-T / T1)
        if len(r.troe_par) > 3:
            Fcent = Fcent + np.exp(-r.troe_par[3] / T)
        lF = np.log10(Fcent)
        c = -0.4 - 0.67 * lF
        nn = 0.75 - 1.27 * lF
        lP = np.log10(Pr)
        f1 = (lP + c) / (nn - 0.14 * (lP + c))
        return 10.0 ** (lF / (1 + f1**2))
    if r.sri:
        a, b, c = r.sri_par[:3]
        d = r.sri_par[3] if len(r.sri_par) > 3 else 1.0
        e = r.sri_par[4] if len(r.sri_par) > 4 else 0.0
        X = 1 / (1 + np.log10(Pr) ** 2)
        return d * (a * np.exp(-b / T) + np.exp(-T / c)) ** X * T**e
    return np.ones_like(T)


def textbook_rates(mech, conp, param, phi):
    """dphi = [dT/dt, dE/dt, V wdot_0 .. V wdot_{Ns-2}] and the net rate of
    progress of every reaction, from the parsed mechanism."""
    specs, reacs = mech.specs, mech.reacs
    Ns = len(specs)
    T = phi[:, 0]
    P, V = (param, phi[:, 1]) if conp else (phi[:, 1], param)
    n_k = phi[:, 2:]
    C = np.empty((len(T), Ns))
    C[:, :-1] = n_k / V[:, None]
    # bath gas by difference; the reference sums the moles in species order,
    # and the difference cancels when the bath gas is rare, so keep that order
    nsum = np.zeros_like(T)
    for k in range(Ns - 1):
        nsum = nsum + n_k[:, k]
    C[:, -1] = P / (RU * T) - nsum / V
    cp, h_RT, s_R = _thermo(specs, T)
    g_RT = h_RT - s_R
    wdot = np.zeros_like(C)
    scale = np.zeros_like(C)
    qnet = np.zeros((len(T), len(reacs)))
    for i, r in enumerate(reacs):
        nu_f = np.zeros(Ns)
        nu_r = np.zeros(Ns)
        for sp, nu in zip(r.reac, r.reac_nu):
            nu_f[sp] += nu
        for sp, nu in zip(r.prod, r.prod_nu):
            nu_r[sp] += nu
        if r.plog:
            pars = sorted(r.plog_par, key=lambda p: p[0])
            lk = np.array([np.log(_arr(A, b, Ta, T)) for (_, A, b, Ta) in pars])
            lp = np.log([p[0] for p in pars])
            lnP = np.log(P)
            k = np.empty_like(T)
            for s in range(len(T)):
                k[s] = np.exp(np.interp(lnP[s], lp, lk[:, s]))
            kf = k
        elif r.cheb:
            Tmin, Tmax = r.cheb_tlim
            Pmin, Pmax = r.cheb_plim
            Tr = (2 / T - 1 / Tmin - 1 / Tmax) / (1 / Tmax - 1 / Tmin)
            Pr_ = ((2 * np.log10(P) - np.log10(Pmin) - np.log10(Pmax)) /
                   (np.log10(Pmax) - np.log10(Pmin)))
            kf = 10.0 ** npcheb.chebval2d(Tr, Pr_, np.asarray(r.cheb_par))
        else:
            kf = _arr(r.A, r.b, r.E, T)
        # third-body concentration
        if r.thd_body or r.low or r.high:
            if r.pdep_sp not in (None, ''):
                M = C[:, r.pdep_sp]
            else:
                eff = np.ones(Ns)
                for sp, e in r.thd_body_eff:
                    eff[sp] = e
                M = C @ eff
        factor = np.ones_like(T)
        if r.thd_body:
            factor = M
        elif r.low:   # falloff: main = high-pressure limit
            k0 = _arr(r.low[0], r.low[1], r.low[2], T)
            Pr = k0 * M / kf
            kf = kf * Pr / (1 + Pr) * _falloff_F(r, Pr, T)
        elif r.high:  # chemically activated: main = low-pressure limit
            kinf = _arr(r.high[0], r.high[1], r.high[2], T)
            Pr = kf * M / kinf
            kf = kf / (1 + Pr) * _falloff_F(r, Pr, T)
        fwd = kf * np.prod(C ** nu_f, axis=1)
        rev = np.zeros_like(T)
        if r.rev:
            dnu = nu_r - nu_f
            if r.rev_par:
                kr = _arr(r.rev_par[0], r.rev_par[1], r.rev_par[2], T)
            else:
                Kc = np.exp(-(g_RT @ dnu)) * (PA / (RU * T)) ** dnu.sum()
                kr = kf / Kc
            rev = kr * np.prod(C ** nu_r, axis=1)
        q = factor * (fwd - rev)
        qnet[:, i] = q
        wdot += np.outer(q, nu_r - nu_f)
        scale += np.outer(np.abs(factor) * (np.abs(fwd) + np.abs(rev)), np.abs(nu_r - nu_f))
    mw = np.array([s.mw for s in specs])
    if conp:
        e_k, c_k = RU * T[:, None] * h_RT, RU * cp
    else:
        e_k, c_k = RU * T[:, None] * (h_RT - 1), RU * (cp - 1)
    heat = (C * c_k).sum(axis=1)
    Tdot = -(e_k * wdot).sum(axis=1) / heat
    Tdot_scale = (np.abs(e_k) * scale).sum(axis=1) / heat
    dn = ((1 - mw[:-1] / mw[-1])[None, :] * wdot[:, :-1]).sum(axis=1)
    dn_scale = (np.abs(1 - mw[:-1] / mw[-1])[None, :] * scale[:, :-1]).sum(axis=1)
    if conp:
        Edot = V * Tdot / T + V * RU * T / P * dn
        Edot_scale = V * Tdot_scale / T + V * RU * T / P * dn_scale
    else:
        Edot = P * Tdot / T + RU * T * dn
        Edot_scale = P * Tdot_scale / T + RU * T * dn_scale
    dphi = np.column_stack([Tdot, Edot, V[:, None] * wdot[:, :-1]])
    dphi_scale = np.column_stack([Tdot_scale, Edot_scale, V[:, None] * scale[:, :-1]])
    return dphi, dphi_scale, qnet, wdot, scale


CASES = [(m, conp) for m in ('h2o2', 'gri30', 'test') for conp in (True, False)] + \
    [('s150', True), ('s500', False)]


@pytest.mark.parametrize('name,conp', CASES)
def test_rates_match_textbook(mechs, name, conp):
    mech = mechs(name)
    param, phi = sample_states(mech, 48, seed=11, T_range=(900.0, 2400.0), alpha=0.5, conp=conp)
    ref, ref_scale, _, _, _ = textbook_rates(mech, conp, param, phi)
    got = oracle.species_rates(mech, conp, param, phi)
    err = np.abs(got - ref) / (1e-11 * ref_scale + 1e-300)
    assert err.max() <= 1.0, 'max error {:.3g} of tolerance at {}'.format(
        err.max(), np.unravel_index(err.argmax(), err.shape))


@pytest.mark.parametrize('name,conp', CASES)
def test_rates_of_progress_match_textbook(mechs, name, conp):
    """Per-reaction net rates (rop_fwd - rop_rev) * pres_mod, so that no
    error can cancel in the sum over reactions."""
    mech = mechs(name)
    param, phi = sample_states(mech, 32, seed=12, T_range=(900.0, 2400.0), alpha=0.5, conp=conp)
    _, _, qnet, _, _ = textbook_rates(mech, conp, param, phi)
    _, fwd, rev, pmod, net = oracle.species_rates_full(mech, conp, param, phi)
    mag = (np.abs(fwd) + np.abs(rev)) * np.abs(pmod)
    err = np.abs(net - qnet) / (1e-11 * mag + 1e-300)
    assert err.max() <= 1.0, 'reaction {} off by {:.3g} of tolerance'.format(
        np.unravel_index(err.argmax(), err.shape)[1], err.max())


def _elements(specs):
    names = sorted({e[0].lower() for s in specs for e in s.elem})
    A = np.zeros((len(names), len(specs)))
    for k, s in enumerate(specs):
        for e, cnt in s.elem:
            A[names.index(e.lower()), k] += cnt
    return A


@pytest.mark.parametrize('name,conp', CASES)
def test_conservation(mechs, name, conp):
    """Element and mass conservation of the checker's production rates (the
    bath gas's rate follows from the others by mass conservation)."""
    mech = mechs(name)
    param, phi = sample_states(mech, 32, seed=13, T_range=(900.0, 2400.0), alpha=0.5, conp=conp)
    got = oracle.species_rates(mech, conp, param, phi)
    V = phi[:, 1] if conp else param
    wdot = got[:, 2:] / V[:, None]
    mw = np.array([s.mw for s in mech.specs])
    w_ns = -(mw[:-1] * wdot).sum(axis=1) / mw[-1]
    full = np.column_stack([wdot, w_ns])
    _, _, _, _, scale = textbook_rates(mech, conp, param, phi)
    # the bath gas's rate is derived, so its rounding scale is the mass sum
    scale = scale.copy()
    scale[:, -1] = (mw[:-1] * scale[:, :-1]).sum(axis=1) / mw[-1]
    A = _elements(mech.specs)
    resid = full @ A.T
    bound = scale @ np.abs(A).T
    assert np.all(np.abs(resid) <= 1e-10 * bound + 1e-300)


@pytest.mark.parametrize('name', ['h2o2', 'gri30', 'test'])
def test_conp_conv_same_production_rates(mechs, name):
    """The production rates depend on (T, C) only: a CONV state and the CONP
    state with the same T, P, V and n give the same V * wdot."""
    mech = mechs(name)
    param_v, phi_v = sample_states(mech, 32, seed=14, conp=False)
    phi_p = phi_v.copy()
    param_p = phi_v[:, 1].copy()    # P
    phi_p[:, 1] = param_v           # V
    conv = oracle.species_rates(mech, False, param_v, phi_v)
    conp = oracle.species_rates(mech, True, param_p, phi_p)
    np.testing.assert_allclose(conp[:, 2:], conv[:, 2:], rtol=1e-12, atol=1e-12 * np.abs(conv[:, 2:]).max())


def test_textbook_detects_planted_mistakes(mechs):
    """The comparison is sharp enough to see the mistakes it is there for:
    a reverse rate built with the wrong sign of sum(nu) in Kc, or a third
    body efficiency dropped, moves the textbook result off the checker."""
    mech = mechs('h2o2')
    param, phi = sample_states(mech, 16, seed=15, conp=True)
    got = oracle.species_rates(mech, True, param, phi)
    ref, ref_scale, _, _, _ = textbook_rates(mech, True, param, phi)
    assert (np.abs(got - ref) / (1e-11 * ref_scale)).max() <= 1.0
    import copy
    import types

    def variant(change):
        reacs = [copy.deepcopy(r) for r in mech.reacs]
        change(reacs)
        return types.SimpleNamespace(specs=mech.specs, reacs=reacs)

    def worst(m):
        ref_bad, _, _, _, _ = textbook_rates(m, True, param, phi)
        return (np.abs(got - ref_bad) / (1e-11 * ref_scale)).max()

    def swap_sides(reacs):
        # reactants and products exchanged: Kc -> 1/Kc and the net rate of
        # progress changes sign
        r = next(r for r in reacs if r.rev and sum(r.prod_nu) != sum(r.reac_nu))
        r.reac, r.reac_nu, r.prod, r.prod_nu = r.prod, r.prod_nu, r.reac, r.reac_nu

    def drop_efficiency(reacs):
        r = next(r for r in reacs if r.thd_body_eff)
        r.thd_body_eff = r.thd_body_eff[1:]

    def troe_T1_T3_swapped(reacs):
        r = next(r for r in reacs if r.troe)
        r.troe_par = [r.troe_par[0], r.troe_par[2], r.troe_par[1]] + list(r.troe_par[3:])

    for change in (swap_sides, drop_efficiency, troe_T1_T3_swapped):
        assert worst(variant(change)) > 1e3, change.__name__
