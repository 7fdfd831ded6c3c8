"""Mechanism-specialised species-rate kernel: HIP source generated from the
tables of one mechanism, one state per lane.

The library's block kernels read every rate parameter from tables at run time
and give each lane of a workgroup a different reaction of one state, so most
of their instructions are table addressing, selects and lanes idle in the
short phases.  The reference instead generates code per mechanism with the
parameters as literals (``examples/opencl/ch4/species_rates.oclh``,
``pyjac/core/rate_subs.py``), vectorised over states.  This module does the
same for gfx950: every lane evaluates one whole state in straight-line code in
which each reaction's Arrhenius parameters, stoichiometry, third-body
efficiencies, falloff parameters and each species' NASA-7 coefficients are
constants, so the instructions that remain are the chemistry's own.  In F
order the lanes of a wave read and write 64 consecutive states of each array
entry: whole cache lines.

The arithmetic is the block kernel's (``csrc/pyjac_block.hip``), expression
for expression, with each table value replaced by its literal:

* concentrations and the bath gas by difference in the reference's
  sequential order (rate_subs.py:597-706);
* the Gibbs terms b_k of the NASA-7 polynomials (rate_subs.py:3281-3388);
* rate constants in the reference's specialised Arrhenius forms
  (rate_subs.py:2803-2870), PLOG (:2084-2248) and Chebyshev (:1904-2081);
* third-body concentrations and falloff blending: Lindemann, Troe, SRI,
  chemically activated (rate_subs.py:1788-1901, :2251-2639, :1510-1624);
* rates of progress with ``fast_powi`` product chains and the reverse rate
  from the equilibrium constant (rate_subs.py:1371-1507, :1627-1785);
* production rates in reaction order (:1079-1167) and the temperature and
  extra-variable rates with the three sums in the reference's sequential
  order (rate_subs.py:709-1076).

The production rates are summed in reaction order here and in the order of the
block kernel's LDS additions there, so the two agree to rounding (within the
rate tolerance's condition term), not bit for bit.

The generated module exports ``pyjac_lane_rates`` and ``pyjac_lane_sig`` (the
tables' fingerprint, which the runtime compares with the handle's before it
accepts the module; see :func:`tables_hash`).
"""

import os
import subprocess

import numpy as np

from ..core.chem_model import RU as RU_, PA as PA_

LANE_VERSION = 1
KERNEL = 'pyjac_lane_rates'
SIG = 'pyjac_lane_sig'
BLOCK = 64

# the arrays of pyjac_mech_tables the generated code depends on, with their
# lengths from the struct's counts (runtime.cpp hashes the same sequence)
_HASHED = [
    ('rxn_off', lambda c: c['Nr'] + 1), ('rxn_spec', lambda c: c['nrs']),
    ('nu_prod', lambda c: c['nrs']), ('nu_reac', lambda c: c['nrs']),
    ('nu_sum', lambda c: c['Nr']), ('rev_mask', lambda c: c['Nr']),
    ('thd_mask', lambda c: c['Nr']), ('simple_mask', lambda c: c['Nr']),
    ('plog_mask', lambda c: c['Nr']), ('cheb_mask', lambda c: c['Nr']),
    ('fall_mask', lambda c: c['Nr']),
    ('simple_form', lambda c: c['n_simple']), ('simple_A', lambda c: c['n_simple']),
    ('simple_b', lambda c: c['n_simple']), ('simple_Ta', lambda c: c['n_simple']),
    ('fall_type', lambda c: c['nf']), ('fall_form', lambda c: c['nf']),
    ('fall_blend', lambda c: c['nf']), ('fall_param', lambda c: c['nf']),
    ('fall_A', lambda c: c['nf']), ('fall_b', lambda c: c['nf']),
    ('fall_Ta', lambda c: c['nf']),
    ('troe_a', lambda c: c['n_troe']), ('troe_T1inv', lambda c: c['n_troe']),
    ('troe_T2', lambda c: c['n_troe']), ('troe_T3inv', lambda c: c['n_troe']),
    ('sri_a', lambda c: c['n_sri']), ('sri_b', lambda c: c['n_sri']),
    ('sri_c', lambda c: c['n_sri']), ('sri_d', lambda c: c['n_sri']),
    ('sri_e', lambda c: c['n_sri']),
    ('thd_map', lambda c: c['nt']), ('thd_type', lambda c: c['nt']),
    ('thd_off', lambda c: c['nt'] + 1), ('thd_spec', lambda c: c['nts']),
    ('thd_eff', lambda c: c['nts']),
    ('plog_np', lambda c: c['n_plog']),
    ('plog_params', lambda c: 4 * c['n_plog'] * c['plog_max_p']),
    ('cheb_nt', lambda c: c['n_cheb']), ('cheb_np', lambda c: c['n_cheb']),
    ('cheb_params', lambda c: c['n_cheb'] * c['cheb_max_t'] * c['cheb_max_p']),
    ('cheb_plim', lambda c: 2 * c['n_cheb']), ('cheb_tlim', lambda c: 2 * c['n_cheb']),
    ('a_lo', lambda c: 7 * c['Ns']), ('a_hi', lambda c: 7 * c['Ns']),
    ('T_mid', lambda c: c['Ns']), ('mw_post', lambda c: c['Ns'] - 1),
]

_FNV_OFFSET = 0xcbf29ce484222325
_FNV_PRIME = 0x100000001b3
_M64 = (1 << 64) - 1


def _counts(mt):
    s = mt.cstruct
    return dict(Ns=s.n_species, Nr=s.n_reactions, nrs=s.n_rxn_spec, n_simple=s.n_simple,
                nf=s.n_fall, n_troe=s.n_troe, n_sri=s.n_sri, nt=s.n_thd, nts=s.n_thd_spec,
                n_plog=s.n_plog, plog_max_p=s.plog_max_p, n_cheb=s.n_cheb,
                cheb_max_t=s.cheb_max_t, cheb_max_p=s.cheb_max_p)


def tables_hash(mt):
    """64-bit FNV-1a over the species and reaction counts (int32) and then the
    bytes of every table the generated code reads, in a fixed order and with
    the lengths the struct's counts give (``pyjac_hip_lane_hash`` computes the
    same number from a ``pyjac_mech_tables``)."""
    c = _counts(mt)
    h = _FNV_OFFSET
    chunks = [np.array([c['Ns'], c['Nr']], dtype=np.int32).tobytes()]
    for name, length in _HASHED:
        a = mt.arrays[name]
        chunks.append(np.ascontiguousarray(a[:length(c)]).tobytes())
    buf = np.frombuffer(b''.join(chunks), dtype=np.uint8)
    for byte in buf.tolist():
        h = ((h ^ byte) * _FNV_PRIME) & _M64
    return h


def lit(x):
    """An exact C++ literal of a double (the shortest repr round-trips)."""
    x = float(x)
    if not np.isfinite(x):
        raise ValueError('non-finite mechanism parameter')
    r = repr(x)
    if 'e' not in r and '.' not in r:
        r += '.0'
    return '(' + r + ')' if x < 0 else r


def _chain(v, e):
    """c^e as the reference's left-to-right product chain (fast_powi)."""
    if e == 0:
        return '1.0'
    out = v
    for _ in range(e - 1):
        out = '(%s * %s)' % (out, v)
    return out


def _powi(val, p):
    """powi of the block kernel: 1/val first for negative exponents."""
    if p < 0:
        return _chain('(1.0 / %s)' % val, -p)
    return _chain(val, p)


def kr_product(slots, nsum, kf, gibbs='b%d', kf_exponent=None):
    """The reverse rate constant kf / (P_sum exp(B_sum)) of a reversible
    reaction (rate_subs.py:1627-1785; P_sum = (PA / RU)^nsum, B_sum the
    stoichiometric sum of the Gibbs terms `gibbs`) as the product kf (1 /
    P_sum) exp(-B_sum), or, given the exponent of an Arrhenius kf, as (1 /
    P_sum) exp(kf_exponent - B_sum): the quotient to a few units in the last
    place, without a division."""
    nb = ['%s * %s' % (lit(float(nr - np_)), gibbs % k) for k, nr, np_ in slots if np_ - nr != 0]
    pinv = 1.0
    for _ in range(abs(nsum)):
        pinv = pinv * (RU_ / PA_ if nsum > 0 else PA_ / RU_)
    fac = '' if nsum == 0 else '%s * ' % lit(pinv)
    if kf_exponent is not None:
        return '%sex(%s)' % (fac, ' + '.join([kf_exponent] + nb))
    return '%s * %sex(%s)' % (kf, fac, ' + '.join(['0.0'] + nb))


def _arrhenius(form, A, b, Ta):
    """The specialised Arrhenius forms (A is log A for forms 2-4); terms that
    are exactly zero are left out, which changes no value since log T and 1/T
    are finite and positive at the clamped temperature."""
    if form == 0:
        return lit(A)
    if form == 1:
        return '%s * %s' % (lit(A), _powi('Tc', int(b)))
    terms = [lit(A)]
    if form in (2, 3) and b != 0.0:
        terms.append('logT * %s' % lit(b))
    if form in (2, 4) and Ta != 0.0:
        terms.append('%s * Tinv' % lit(-Ta))
    return 'ex(%s)' % ' + '.join(terms)


# the two targets: the gfx950 module, and a host build of the same code that
# the CPU tests compile with g++ to check the generator against the oracle
TARGET_HIP = r'''// generated by pyjac.libgen.lanegen -- mechanism-specialised species rates
#include <hip/hip_runtime.h>
#define PJ_FN __device__ __forceinline__
#define PJ_CONST __device__ const
#define PJ_OPAQUE(x) __asm__ volatile("" : "+v"(x))
'''

TARGET_HOST = r'''// generated by pyjac.libgen.lanegen -- host build of the specialised species rates
#include <math.h>
#define PJ_FN static inline
#define PJ_CONST static const
#define PJ_OPAQUE(x) ((void)0)
static inline double __dadd_rn(double a, double b) { return a + b; }
static inline double __dmul_rn(double a, double b) { return a * b; }
'''

PRELUDE = r'''
namespace {
constexpr double RU = 8314.4621;
constexpr double PA = 101325.0;
constexpr double EXP_MAX = 690.775527898;
constexpr double SMALL = 1e-300;
constexpr double T_MIN = 100.0, T_MAX = 10000.0;
constexpr double V_MIN = 1e-30;
constexpr double LN10 = 2.302585092994046;

PJ_FN double ex(double x) { return exp(fmin(EXP_MAX, x)); }

// b_k = g_k / RT of the NASA-7 polynomial (rate_subs.py:3281-3388)
PJ_FN double gibbs(double T, double logT, double Tinv, double a0, double a1, double a2,
                    double a3, double a4, double a5, double a6) {
    return T * (T * (T * (T * a4 / 20.0 + a3 / 12.0) + a2 / 6.0) + a1 / 2.0) +
           (a0 + -1.0) * logT + -1.0 * a0 + a6 + -1.0 * a5 * Tinv;
}

// energy (h or u) and heat capacity (cp or cv) of the NASA-7 polynomial;
// the three sums of the temperature rate advance together, species by species
PJ_FN void energy_sums(double T, int conp, double a0, double a1, double a2, double a3,
                       double a4, double a5, double c, double w, double& lower,
                       double& upper) {
    const double x = T * (T * (T * (T * (T * a4 / 5.0 + a3 / 4.0) + a2 / 3.0) + a1 / 2.0) + a0);
    const double y = T * (T * (T * (T * a4 + a3) + a2) + a1) + a0;
    const double energy = RU * (conp ? x + a5 : x - T + a5);
    const double heat = RU * (conp ? y : y - 1.0);
    lower = __dadd_rn(lower, __dmul_rn(c, heat));
    upper = __dadd_rn(upper, __dmul_rn(energy, w));
}
'''

PLOG_HELPER = r'''
// PLOG rate constant (rate_subs.py:2084-2248); par[q * mp + k] holds
// (log P, log A, b, Ta) of pressure k for q = 0..3
template <int MP>
PJ_FN double plog_k(const double* par, int numP, double logT, double Tinv, double logP) {
    auto P = [&](int q, int k) { return par[q * MP + k]; };
    const bool lower = logP <= P(0, 0);
    const bool upper = logP > P(0, numP);
    int lo = 0, hi = 0;
    if (upper) { lo = numP; hi = numP; }
#pragma unroll
    for (int k = 0; k < MP - 1; ++k)
        if (k < numP && logP > P(0, k) && logP <= P(0, k + 1)) {
            lo = k;
            hi = k + 1;
        }
    const double logk1 = P(1, lo) + logT * P(2, lo) - P(3, lo) * Tinv;
    const double logk2 = P(1, hi) + logT * P(2, hi) - P(3, hi) * Tinv;
    double kt = logk1;
    if (!(lower || upper))
        kt = (-logk1 + logk2) * (logP - P(0, lo)) / (P(0, hi) - P(0, lo)) + kt;
    return exp(fmin(EXP_MAX, kt));
}
'''

CHEB_HELPER = r'''
// Chebyshev rate constant (rate_subs.py:1904-2081); recurrences
// T_p = 2 x T_{p-1} - T_{p-2}
template <int MT, int MPP>
PJ_FN double cheb_k(const double* par, int numT, int numP, double Tmin, double Tmax,
                    double Pmin, double Pmax, double Tinv, double logP) {
    const double Pred = (-Pmax - Pmin + 2.0 * logP) / (Pmax - Pmin);
    const double Tred = (-Tmax - Tmin + 2.0 * Tinv) / (Tmax - Tmin);
    double kt = 0.0;
    double tm2 = 1.0, tm1 = Tred;
#pragma unroll
    for (int t = 0; t < MT; ++t) {
        double tt;
        if (t == 0) tt = 1.0;
        else if (t == 1) tt = tm1;
        else { tt = 2.0 * Tred * tm1 - tm2; tm2 = tm1; tm1 = tt; }
        double temp = 0.0;
        double pm2 = 1.0, pm1 = Pred;
#pragma unroll
        for (int k = 0; k < MPP; ++k) {
            double pk;
            if (k == 0) pk = 1.0;
            else if (k == 1) pk = pm1;
            else { pk = 2.0 * Pred * pm1 - pm2; pm2 = pm1; pm1 = pk; }
            if (k < numP) temp = temp + pk * par[t * MPP + k];
        }
        if (t < numT) kt = kt + tt * temp;
    }
    return exp(LN10 * fmin(EXP_MAX, kt));
}
'''


class LaneRatesSource(object):
    """The generated source for one mechanism's tables (``MechTables``)."""

    def __init__(self, mt):
        self.mt = mt
        self.A = mt.arrays
        self.c = _counts(mt)
        self.Ns = self.c['Ns']
        self.Nr = self.c['Nr']
        self.hash = tables_hash(mt)

    # ---------------------------------------------------------------- parts
    def _species(self):
        A, Ns = self.A, self.Ns
        out = ['    // concentrations; the bath gas by difference (rate_subs.py:597-706).',
               '    // The state\'s moles are loaded together and pinned by empty asm',
               '    // statements before their first use, so that the loads are in flight',
               '    // together instead of one after another',
               '    double nsum = 0.0;']
        for k in range(Ns - 1):
            out.append('    double n%d = ps[%d * sI];' % (k, k + 2))
        for k in range(Ns - 1):
            out.append('    PJ_OPAQUE(n%d);' % k)
        for k in range(Ns - 1):
            out.append('    n%d = fmax(SMALL, n%d);' % (k, k))
            out.append('    nsum = nsum + n%d;' % k)
            out.append('    const double c%d = n%d * V_inv;' % (k, k))
        out.append('    const double c%d = __dadd_rn(P / (RU * Tc), '
                   '__dmul_rn(-1.0 * fmax(SMALL, nsum), V_inv));' % (Ns - 1))
        mids = sorted(set(float(x) for x in A['T_mid'][:Ns]))
        self._mid = {m: i for i, m in enumerate(mids)}
        for m, i in self._mid.items():
            out.append('    const bool lo%d = Tc < %s;' % (i, lit(m)))
        out.append('    // Gibbs terms (rate_subs.py:3281-3388)')
        for k in range(Ns):
            out.append('    const double b%d = gibbs(Tc, logT, Tinv, %s);' % (k, self._coeffs(k, 7)))
        return out

    def _coeffs(self, k, n, flag='lo'):
        A = self.A
        lo = A['a_lo'][7 * k:7 * k + 7]
        hi = A['a_hi'][7 * k:7 * k + 7]
        sel = '%s%d' % (flag, self._mid[float(A['T_mid'][k])])
        parts = []
        for i in range(n):
            if lo[i] == hi[i]:
                parts.append(lit(lo[i]))
            else:
                parts.append('%s ? %s : %s' % (sel, lit(lo[i]), lit(hi[i])))
        return ', '.join(parts)

    def _pdep(self, r):
        """Statements leaving the reaction's pressure modification in `ci`
        (third body or falloff) and its PLOG / Chebyshev rate constant in
        `ksp`; empty when it has neither."""
        A = self.A
        out = []
        t = int(A['thd_mask'][r])
        if t >= 0:
            tt = int(A['thd_type'][t])
            ns = 1 if tt != 2 else 0
            out.append('        double thd = P * %s / (RU * T);' % ('1.0' if ns else '0.0'))
            for e in range(int(A['thd_off'][t]), int(A['thd_off'][t + 1])):
                eff = float(A['thd_eff'][e]) + -1.0 * ns
                out.append('        thd = thd + %s * c%d;' % (lit(eff), int(A['thd_spec'][e])))
            f = int(A['fall_mask'][r])
            if f < 0:
                out.append('        const double ci = thd;')
            else:
                out.extend(self._falloff(r, f))
        return out

    def _falloff(self, r, f):
        A = self.A
        ftype = int(A['fall_type'][f])
        kff = _arrhenius(int(A['fall_form'][f]), A['fall_A'][f], A['fall_b'][f], A['fall_Ta'][f])
        out = ['        const double kff = %s;' % kff]
        if ftype:
            out.append('        const double Pr = thd * kf / kff;')
        else:
            out.append('        const double Pr = thd * kff / kf;')
        blend = int(A['fall_blend'][f])
        i = int(A['fall_param'][f])
        if blend == 2:
            a = float(A['troe_a'][i])
            T1 = float(A['troe_T1inv'][i])
            T3 = float(A['troe_T3inv'][i])
            T2 = float(A['troe_T2'][i])
            out.append('        double Fc = %s * ex(-1.0 * Tc * %s) + %s * ex(-1.0 * Tc * %s);'
                       % (lit(a), lit(T1), lit(1.0 + -1.0 * a), lit(T3)))
            if T2 != 0.0:
                out.append('        Fc = Fc + ex(-1.0 * %s / Tc);' % lit(T2))
            out.append('        const double logFc = log10(fmax(SMALL, Fc));')
            out.append('        const double logPr = log10(fmax(SMALL, Pr));')
            out.append('        const double B = -1.1762 * logFc + -1.0 * 0.14 * logPr + 0.806;')
            out.append('        const double Af = -0.67 * logFc + logPr + -0.4;')
            out.append('        const double Fi = pow(fmax(SMALL, Fc), 1.0 / (Af * Af / (B * B) + 1.0));')
        elif blend == 3:
            a, b, cc, d, e = (float(A['sri_' + x][i]) for x in 'abcde')
            out.append('        const double logPr = log10(fmax(SMALL, Pr));')
            out.append('        const double X = 1.0 / (logPr * logPr + 1.0);')
            out.append('        const double base = %s * ex(%s * Tinv) + ex(-Tc / %s);'
                       % (lit(a), lit(-b), lit(cc)))
            out.append('        double Fi = pow(base, X);')
            if d != 1.0:
                out.append('        Fi = Fi * %s;' % lit(d))
            if e != 0.0:
                out.append('        Fi = Fi * pow(Tc, %s);' % lit(e))
        else:
            out.append('        const double Fi = 1.0;')
        if ftype:
            out.append('        const double ci = Fi / (1.0 + Pr);')
        else:
            out.append('        const double ci = Fi / (1.0 + Pr) * Pr;')
        return out

    def _kf(self, r):
        A = self.A
        si = int(A['simple_mask'][r])
        if si >= 0:
            return _arrhenius(int(A['simple_form'][si]), A['simple_A'][si],
                              A['simple_b'][si], A['simple_Ta'][si])
        p = int(A['plog_mask'][r])
        c = self.c
        if p >= 0:
            mp = c['plog_max_p']
            return 'plog_k<%d>(plog_tab + %d, %d, logT, Tinv, logP)' % (
                mp, 4 * mp * p, int(A['plog_np'][p]) - 1)
        q = int(A['cheb_mask'][r])
        mt_, mp_ = c['cheb_max_t'], c['cheb_max_p']
        Tl = A['cheb_tlim'][2 * q:2 * q + 2]
        Pl = A['cheb_plim'][2 * q:2 * q + 2]
        return 'cheb_k<%d, %d>(cheb_tab + %d, %d, %d, %s, %s, %s, %s, Tinv, logP)' % (
            mt_, mp_, q * mt_ * mp_, int(A['cheb_nt'][q]), int(A['cheb_np'][q]),
            lit(Tl[0]), lit(Tl[1]), lit(Pl[0]), lit(Pl[1]))

    def _reaction(self, r):
        A, Ns = self.A, self.Ns
        lo, hi = int(A['rxn_off'][r]), int(A['rxn_off'][r + 1])
        slots = [(int(A['rxn_spec'][i]), int(A['nu_reac'][i]), int(A['nu_prod'][i]))
                 for i in range(lo, hi)]
        out = ['    {  // reaction %d' % r,
               '        const double kf = %s;' % self._kf(r)]
        rf = 'kf'
        for k, nr, _ in slots:
            if nr:
                rf = '%s * %s' % (rf, _chain('c%d' % k, nr) if nr == 1 else '(%s)' % _chain('c%d' % k, nr))
        out.append('        const double rf = %s;' % rf)
        rr = None
        if int(A['rev_mask'][r]) >= 0:
            out.append('        const double kr = %s;'
                       % kr_product(slots, int(A['nu_sum'][r]), 'kf'))
            rr = 'kr'
            for k, _, np_ in slots:
                if np_:
                    rr = '%s * %s' % (rr, _chain('c%d' % k, np_) if np_ == 1 else '(%s)' % _chain('c%d' % k, np_))
            out.append('        const double rr = %s;' % rr)
        out.extend(self._pdep(r))
        net = '(rf - rr)' if rr else '(rf - 0.0)'
        if int(A['thd_mask'][r]) >= 0:
            out.append('        const double net = %s * ci;' % net)
        else:
            out.append('        const double net = %s * 1.0;' % net)
        for k, nr, np_ in slots:
            nu = np_ - nr
            if nu:
                out.append('        w%d = w%d + %s * net;' % (k, k, lit(float(nu))))
        out.append('    }')
        return out

    def _energy(self):
        A, Ns = self.A, self.Ns
        out = ['    // temperature and extra-variable rates (rate_subs.py:709-1076); the',
               '    // coefficients are selected again from an opaque copy of the',
               '    // temperature, so that the compiler does not keep the selections of',
               '    // the Gibbs terms alive through the whole reaction pass',
               '    double Tq = Tc;',
               '    PJ_OPAQUE(Tq);']
        for m, i in self._mid.items():
            out.append('    const bool hq%d = Tq < %s;' % (i, lit(m)))
        out.append('    double lower = 0.0, upper = 0.0, dE = 0.0;')
        for k in range(Ns):
            out.append('    energy_sums(Tq, conp, %s, c%d, w%d, lower, upper);'
                       % (self._coeffs(k, 6, 'hq'), k, k))
            if k < Ns - 1:
                mwf = 1.0 + -1.0 * float(A['mw_post'][k])
                out.append('    dE = __dadd_rn(dE, __dmul_rn(%s, w%d));' % (lit(mwf), k))
        out += ['    const double V_val = fmax(V_MIN, V);',
                '    const double Tdot = 0.0 + -1.0 * upper / lower;',
                '    double Edot;',
                '    if (conp) {',
                '        Edot = V_val * Tdot / Tc;',
                '        Edot = Edot + V_val * dE * Tc * RU / P;',
                '    } else {',
                '        Edot = P * Tdot / Tc;',
                '        Edot = Edot + Tc * RU * dE;',
                '    }',
                '    double* out = dphi + s * sN;',
                '    out[0] = Tdot;',
                '    out[sI] = Edot;']
        for k in range(Ns - 1):
            out.append('    out[%d * sI] = V_val * w%d;' % (k + 2, k))
        return out

    def source(self, target='hip'):
        """The module's source; ``target`` 'hip' (the gfx950 kernel) or
        'host' (the same per-state code in a loop, for the CPU tests)."""
        A, c, Ns, Nr = self.A, self.c, self.Ns, self.Nr
        out = [TARGET_HIP if target == 'hip' else TARGET_HOST, PRELUDE]
        if c['n_plog']:
            out.append(PLOG_HELPER)
            n = 4 * c['n_plog'] * c['plog_max_p']
            # PLOG reaction p's block: par[q * mp + k] = plog_params[(q * n_plog + p) * mp + k]
            mp, npl = c['plog_max_p'], c['n_plog']
            vals = []
            for p in range(npl):
                for q in range(4):
                    for k in range(mp):
                        vals.append(A['plog_params'][(q * npl + p) * mp + k])
            out.append('PJ_CONST double plog_tab[%d] = {%s};' % (n, ', '.join(lit(v) for v in vals)))
        if c['n_cheb']:
            out.append(CHEB_HELPER)
            n = c['n_cheb'] * c['cheb_max_t'] * c['cheb_max_p']
            out.append('PJ_CONST double cheb_tab[%d] = {%s};'
                       % (n, ', '.join(lit(v) for v in A['cheb_params'][:n])))
        out.append('')
        out.append('// species rates of state s; phi / dphi entry i of state s at [s * sN + i * sI]')
        out.append('PJ_FN void lane_state(long s, const double* __restrict__ param,\n'
                   '                      const double* __restrict__ phi, double* __restrict__ dphi,\n'
                   '                      long sN, long sI, int conp) {')
        out += ['    const double* ps = phi + s * sN;',
                '    const double T = ps[0], x1 = ps[sI], pv = param[s];',
                '    const double P = conp ? pv : x1;',
                '    const double V = conp ? x1 : pv;',
                '    const double Tc = fmin(T_MAX, fmax(T_MIN, T));',
                '    const double logT = log(Tc);',
                '    const double Tinv = 1.0 / Tc;',
                '    const double logP = log(fmax(SMALL, P));',
                '    const double V_inv = 1.0 / fmax(V_MIN, V);',
                '    (void)logP;']
        out += self._species()
        out.append('    double %s;' % ', '.join('w%d = 0.0' % k for k in range(Ns)))
        for r in range(Nr):
            out += self._reaction(r)
        out += self._energy()
        out.append('}')
        out.append('}  // namespace\n')
        sig = '{0x%016xULL, %dULL, %dULL, %dULL}' % (self.hash, LANE_VERSION, Ns, Nr)
        if target == 'hip':
            out.append('extern "C" __device__ const unsigned long long %s[4] = %s;\n' % (SIG, sig))
            out.append('// one state per lane, %d states per workgroup' % BLOCK)
            out.append('extern "C" __global__ void __launch_bounds__(%d) %s(\n'
                       '    long n, const double* __restrict__ param, const double* __restrict__ phi,\n'
                       '    double* __restrict__ dphi, long sN, long sI, int conp) {' % (BLOCK, KERNEL))
            out += ['    const long s = (long)blockIdx.x * %d + threadIdx.x;' % BLOCK,
                    '    if (s < n) lane_state(s, param, phi, dphi, sN, sI, conp);',
                    '}']
        else:
            out.append('extern "C" const unsigned long long %s[4] = %s;\n' % (SIG, sig))
            out.append('extern "C" void %s_host(long n, const double* param, const double* phi,\n'
                       '                          double* dphi, long sN, long sI, int conp) {' % KERNEL)
            out += ['    for (long s = 0; s < n; ++s) lane_state(s, param, phi, dphi, sN, sI, conp);',
                    '}']
        return '\n'.join(out) + '\n'


def generate(mt, target='hip'):
    """Source of the specialised species-rate module of ``mt``."""
    return LaneRatesSource(mt).source(target)


def build(mt, out_path, verbose=False):
    """Generate and compile the module of ``mt`` into the code object
    ``out_path`` (hipcc --genco for gfx950); returns its path.  The source is
    written beside it (``.hip``)."""
    from . import hipcc, HIPCC_FLAGS
    if mt.Ns > MAX_SPECIES:
        raise ValueError('{} species: the specialised module keeps every concentration, '
                         'Gibbs term and production rate of a state in registers, which '
                         'holds up to {} species'.format(mt.Ns, MAX_SPECIES))
    src = generate(mt)
    base = os.path.splitext(out_path)[0]
    src_path = base + '.hip'
    os.makedirs(os.path.dirname(os.path.abspath(out_path)) or '.', exist_ok=True)
    with open(src_path, 'w') as f:
        f.write(src)
    flags = [x for x in HIPCC_FLAGS if x != '-fPIC']
    tmp = out_path + '.tmp'
    cmd = [hipcc(), '--genco'] + flags + ['-o', tmp, src_path]
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError('hipcc failed on %s:\n%s' % (src_path, res.stdout[-4000:]))
    if verbose:
        print(res.stdout)
    os.replace(tmp, out_path)
    return out_path


# modules of the bundled mechanisms, built by __graft_entry__.build() / the
# pyjac --lang hip build, are cached here under their tables' fingerprint
MODULE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                          'lane_modules')
MAX_SPECIES = 64
BUNDLED = ('h2o2', 'gri30', 'test')


def module_path(mt, directory=None):
    """Where the module of ``mt``'s tables is cached (it may not exist)."""
    return os.path.join(directory or MODULE_DIR,
                        'lane_%016x_v%d.hsaco' % (tables_hash(mt), LANE_VERSION))


def ensure_module(mt, build_missing=True, directory=None):
    """Path of ``mt``'s module, compiled first if missing and
    ``build_missing``; None if missing and not built."""
    path = module_path(mt, directory)
    if os.path.exists(path):
        return path
    if not build_missing:
        return None
    return build(mt, path)


def build_bundled(jobs=3, names=BUNDLED):
    """Compile the modules of the bundled mechanisms (always rebuilds)."""
    import concurrent.futures
    from .._paths import mechanism_files
    from ..core.mech_tables import MechTables

    def one(name):
        mt = MechTables.from_file(*mechanism_files(name))
        return build(mt, module_path(mt))
    with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        return list(ex.map(one, names))
