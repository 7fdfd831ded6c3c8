"""Mechanism-specialised analytical Jacobian: HIP source generated from the
tables of one mechanism, one state per lane, written column by column.

The block kernels accumulate each state's Jacobian in a dense LDS array that
one workgroup owns, reading every rate parameter from tables; most of their
instructions are table addressing, LDS atomics and lanes idle in the short
phases.  The reference generates code per mechanism with the parameters as
literals (``pyjac/core/create_jacobian.py``), vectorised over states.  Here a
lane evaluates one whole state in straight-line code in two kernels:

* kernel A (``pyjac_ljac_a``) evaluates the species rates exactly as the
  specialised species-rate module (``lanegen``) does and, per reaction, every
  value the Jacobian's sub-kernels need from it: the PLOG and Chebyshev rate
  constants (kernel B recomputes the Arrhenius ones), the pressure
  modification, the third-body/falloff derivative product, and the
  reaction's total contributions to the temperature and extra-variable
  columns (``create_jacobian.py:362-3230``); per species, the energy weights
  of the temperature row and the terms of the two corner entries.  They go to
  a scratch array laid out ``[value][state]``, so that the lanes of a wave
  write and read 64 consecutive states of a value: whole cache lines.
* kernel B (``pyjac_ljac_b``) forms the entries in column order: the
  temperature and extra-variable columns as row sums of the per-reaction
  column contributions, then each species column j from the reactions that
  contain species j (``dRopidnj``, ``create_jacobian.py:4169-4520``), the
  third-body and falloff derivatives (``dci_*_dnj``, ``:3558-4166``) and the
  bath-gas terms that every column of a row shares, with the temperature and
  extra-variable rows of the column (``dTdot_dnj``, ``dEdot_dnj``,
  ``:3266-3484``) summed over the column's rows in row order.  Every entry is
  stored once, at its position in the handle's format and data order, with
  the output strides the runtime gives: in F order a wave's store is one
  whole line; in C order the runtime has the entries written as lines of a
  chunk (at their C-order positions) and transposes them into C-order rows.

The expressions are the checker's restatement of the reference's sub-kernels
(``oracle/pyjac_oracle.cpp``), term for term with each table value replaced
by its literal; the order in which an entry's contributions are added differs
(per reaction, then per row), so entries agree to rounding.  The module
carries the same fingerprint check as the species-rate module and exports its
entry list, ``pyjac_ljac_entries`` (row << 16 | column, in the order the
kernel stores them), and the scratch size per state, ``pyjac_ljac_sig``.
"""

import os
import re
import subprocess

import numpy as np

from . import lanegen
from .lanegen import lit, _powi, tables_hash

LJAC_VERSION = 2
KERNEL_A = 'pyjac_ljac_a'
KERNEL_B = 'pyjac_ljac_b'
SIG = 'pyjac_ljac_sig'
ENTRIES = 'pyjac_ljac_entries'
BLOCK = 64
MAX_SPECIES = 64
# states per chunk of the scratch array: its stride, a literal in the module
# so that every scratch offset is a constant
STRIDE = 1 << 17

TARGET_HIP = r'''// generated by pyjac.libgen.lanejac -- mechanism-specialised analytical Jacobian
#include <hip/hip_runtime.h>
#define PJ_FN __device__ __forceinline__
#define PJ_CONST __device__ const
#define PJ_OPAQUE(x) __asm__ volatile("" : "+v"(x))
#define PJ_OPQ(x) __asm__ volatile("" : "+s"(x))
#define PJ_SS %d
#define PJ_LANES 64
#define PJ_NTST(p, v) __builtin_nontemporal_store((double)(v), (p))
#define PJ_NTLD(p) __builtin_nontemporal_load(p)
'''

TARGET_HOST = r'''// generated by pyjac.libgen.lanejac -- host build of the specialised Jacobian
#include <math.h>
#define PJ_FN static inline
#define PJ_CONST static const
#define PJ_OPAQUE(x) ((void)0)
#define PJ_OPQ(x) ((void)0)
static long pj_ss = 1;
#define PJ_SS pj_ss
#define PJ_LANES 1
#define PJ_NTST(p, v) (*(p) = (v))
#define PJ_NTLD(p) (*(p))
static inline double __dadd_rn(double a, double b) { return a + b; }
static inline double __dmul_rn(double a, double b) { return a * b; }
'''

JAC_HELPERS = r'''
constexpr double LN10SQ = LN10 * LN10;
constexpr double RU_INV = 1.0 / RU;

// derivative of the Gibbs term b_k with respect to T (create_jacobian.py
// :3234-3555)
PJ_FN double dgibbs(double T, double Tinv, double a0, double a1, double a2, double a3,
                    double a4, double a5) {
    return T * (T * (T * a4 / 5.0 + a3 / 4.0) + a2 / 3.0) + a1 / 2.0 +
           Tinv * (a0 + -1.0 + a5 * Tinv);
}

// derivative of cp with respect to T
PJ_FN double dcp(double T, double a1, double a2, double a3, double a4) {
    return RU * (T * (T * (4.0 * T * a4 + 3.0 * a3) + 2.0 * a2) + a1);
}

// energy (h or u) and heat capacity (cp or cv) of the NASA-7 polynomial
PJ_FN void thermo(double T, int conp, double a0, double a1, double a2, double a3, double a4,
                  double a5, double& energy, double& heat) {
    const double x = T * (T * (T * (T * (T * a4 / 5.0 + a3 / 4.0) + a2 / 3.0) + a1 / 2.0) + a0);
    const double y = T * (T * (T * (T * a4 + a3) + a2) + a1) + a0;
    energy = RU * (conp ? x + a5 : x - T + a5);
    heat = RU * (conp ? y : y - 1.0);
}
'''

PLOG_JAC = r'''
// PLOG interval search of the derivative kernels (create_jacobian.py
// :2823-2850); par[q * MP + k] holds (log P, log A, b, Ta) of pressure k
template <int MP>
PJ_FN void plog_interval(const double* par, int numP, double logP, int& lo, int& hi) {
    lo = 0;
    hi = numP;
#pragma unroll
    for (int k = 0; k < MP - 1; ++k)
        if (k <= numP && logP > par[k] && logP <= par[k + 1]) {
            lo = k;
            hi = k + 1;
        }
}

// d ln kf / dT of a PLOG reaction (create_jacobian.py:2823-2899)
template <int MP>
PJ_FN double plog_dT(const double* par, int numP, double logP, double Tinv) {
    auto P = [&](int q, int k) { return par[q * MP + k]; };
    int lo, hi;
    plog_interval<MP>(par, numP, logP, lo, hi);
    double dkf;
    if (logP > P(0, numP))
        dkf = (P(2, hi) + P(3, hi) * Tinv) * Tinv;
    else
        dkf = (P(2, lo) + P(3, lo) * Tinv) * Tinv;
    if (logP > P(0, 0) && logP <= P(0, numP))
        dkf = dkf + Tinv * (logP - P(0, lo)) *
                        (P(2, hi) - P(2, lo) + (P(3, hi) - P(3, lo)) * Tinv) /
                        (P(0, hi) - P(0, lo));
    return dkf;
}

// d kf / dP over kf of a PLOG reaction, constant volume (create_jacobian.py
// :1000-1140)
template <int MP>
PJ_FN double plog_dP(const double* par, int numP, double logP, double logT, double Tinv,
                     double P_) {
    auto P = [&](int q, int k) { return par[q * MP + k]; };
    int lo, hi;
    plog_interval<MP>(par, numP, logP, lo, hi);
    double dkf = 0;
    if (logP > P(0, 0) && logP <= P(0, numP))
        dkf = (P(1, hi) - P(1, lo) + logT * (P(2, hi) - P(2, lo)) -
               (P(3, hi) - P(3, lo)) * Tinv) /
              (P_ * (P(0, hi) - P(0, lo)));
    return dkf;
}
'''

CHEB_JAC = r'''
// d ln kf / dT of a Chebyshev reaction (create_jacobian.py:2902-2975)
template <int MT, int MPP>
PJ_FN double cheb_dT(const double* par, int nt, int np_, double Tmin, double Tmax, double Pmin,
                     double Pmax, double Tinv, double logP) {
    constexpr int PM = (MPP > MT - 1 ? MPP : MT - 1);
    constexpr int SZ = (PM + 1 > 2 ? PM + 1 : 2);
    double pp[SZ], tp[SZ];
    const int numP = np_, numT = nt - 1;
    const double Tred = (2 * Tinv - Tmax - Tmin) / (Tmax - Tmin);
    const double Pred = (2 * logP - Pmax - Pmin) / (Pmax - Pmin);
    pp[0] = 1;
    pp[1] = Pred;
    tp[0] = 1;
    tp[1] = 2 * Tred;
#pragma unroll
    for (int p = 2; p < PM; ++p) {
        if (p < numP) pp[p] = 2 * Pred * pp[p - 1] - pp[p - 2];
        if (p < numT) tp[p] = 2 * Tred * tp[p - 1] - tp[p - 2];
    }
    double dkf = 0;
#pragma unroll
    for (int t = 0; t < MT - 1; ++t) {
        double temp = 0;
#pragma unroll
        for (int k = 0; k < MPP; ++k)
            if (k < numP) temp = temp + pp[k] * par[(t + 1) * MPP + k];
        if (t < numT) dkf = dkf + (t + 1) * tp[t] * temp;
    }
    return -dkf * 2 * LN10 * Tinv * Tinv / (Tmax - Tmin);
}

// d kf / dP over kf of a Chebyshev reaction, constant volume
// (create_jacobian.py:1143-1290)
template <int MT, int MPP>
PJ_FN double cheb_dP(const double* par, int nt, int np_, double Tmin, double Tmax, double Pmin,
                     double Pmax, double Tinv, double logP, double P_) {
    constexpr int PM = (MPP - 1 > MT ? MPP - 1 : MT);
    constexpr int SZ = (PM + 1 > 2 ? PM + 1 : 2);
    double pp[SZ], tp[SZ];
    const int numP = np_ - 1, numT = nt;
    const double Tred = (2 * Tinv - Tmax - Tmin) / (Tmax - Tmin);
    const double Pred = (2 * logP - Pmax - Pmin) / (Pmax - Pmin);
    pp[0] = 1;
    pp[1] = 2 * Pred;
    tp[0] = 1;
    tp[1] = Tred;
#pragma unroll
    for (int p = 2; p < PM; ++p) {
        if (p < numP) pp[p] = 2 * Pred * pp[p - 1] - pp[p - 2];
        if (p < numT) tp[p] = 2 * Tred * tp[p - 1] - tp[p - 2];
    }
    double dkf = 0;
#pragma unroll
    for (int t = 0; t < MT; ++t) {
        double temp = 0;
#pragma unroll
        for (int k = 0; k < MPP - 1; ++k)
            if (k < numP) temp = temp + (k + 1) * pp[k] * par[t * MPP + k + 1];
        if (t < numT) dkf = dkf + tp[t] * temp;
    }
    return dkf * 2 * LN10 / (P_ * (Pmax - Pmin));
}
'''


def _nu(nr, np_):
    """(nu_prod + -1.0 * nu_reac) as the literal the checker multiplies by."""
    return float(np_) + -1.0 * float(nr)


class LaneJacSource(lanegen.LaneRatesSource):
    """The generated Jacobian source for one mechanism's tables."""

    def __init__(self, mt, nt_store=True, u_lds=True, db_depth=3, nt_loads=True,
                 early_loads=True):
        super(LaneJacSource, self).__init__(mt)
        self.early_loads = early_loads
        self.db_depth = db_depth
        self.nt_loads = nt_loads
        # the entries are stored non-temporally (they are not read again, and
        # the state that the columns reload stays in L2); the bath-gas terms
        # of the rows live in LDS beside the Gibbs terms instead of registers
        self.nt_store = nt_store
        self.u_lds = u_lds
        self._nU = 0
        A, Ns, Nr = self.A, self.Ns, self.Nr
        self.ns = Ns - 1
        self.NN = Ns + 1
        self.slots = []
        for r in range(Nr):
            lo, hi = int(A['rxn_off'][r]), int(A['rxn_off'][r + 1])
            self.slots.append([(int(A['rxn_spec'][i]), int(A['nu_reac'][i]), int(A['nu_prod'][i]))
                               for i in range(lo, hi)])
        self.rev = [int(A['rev_mask'][r]) >= 0 for r in range(Nr)]
        self.thd = [int(A['thd_mask'][r]) for r in range(Nr)]
        self.fall = [int(A['fall_mask'][r]) for r in range(Nr)]
        self.has_ns = set(int(x) for x in A['rxn_has_ns'][:int(mt.cstruct.n_rxn_has_ns)])
        # scratch slots: per reaction and per species values kernel B reads
        self.sv = {}
        self.table_kf = [int(A['plog_mask'][r]) >= 0 or int(A['cheb_mask'][r]) >= 0
                         for r in range(Nr)]
        # reactions with extra-variable-column terms: at constant volume only
        # those with the bath gas, a third body or a PLOG / Chebyshev rate
        self.ee_conv = [r in self.has_ns or self.thd[r] >= 0 or self.table_kf[r]
                        for r in range(Nr)]
        for r in range(Nr):
            if self.table_kf[r]:
                self._slot('KF', r)
            if self.thd[r] >= 0:
                self._slot('CI', r)
                self._slot('G', r)
            self._slot('EE', r)
        for k in range(Ns - 1):
            for name in ('EW', 'DCP', 'WD', 'JT'):
                self._slot(name, k)
        for k in range(Ns):
            self._slot('DB', k)
        for name in ('DTDT', 'CPTOT', 'DTSUM2', 'CPNS', 'JTT', 'JET'):
            self._slot(name, 0)
        self.nscr = len(self.sv)
        self.pattern = set()
        rows, cols = mt.sparse_rows_cols('C')
        for a, b in zip(rows, cols):
            self.pattern.add((int(a), int(b)))
        self.entries = []

    def _slot(self, name, i):
        self.sv[(name, i)] = len(self.sv)

    # ------------------------------------------------------------ kernel A
    def _thd_info(self, r):
        """(thd index t, type, last efficiency species, its efficiency)."""
        A = self.A
        t = self.thd[r]
        e = int(A['thd_off'][t + 1]) - 1
        last = int(A['thd_spec'][e]) if e >= int(A['thd_off'][t]) else -1
        eff = float(A['thd_eff'][e]) if e >= int(A['thd_off'][t]) else 0.0
        return t, int(A['thd_type'][t]), last, eff

    def _ns_products(self, r):
        """S'_ns and S''_ns of the column kernels (create_jacobian.py:3027-3060)."""
        slots = self.slots[r]
        _, nf_last, nr_last = slots[-1]
        sf = [lit(float(nf_last))]
        sr = [lit(float(nr_last))]
        for k, nf, nr in slots:
            if k == self.ns:
                nf, nr = nf - 1, nr - 1
            if nf:
                sf.append(_powi(self.cgfmt % k, nf))
            if nr:
                sr.append(_powi(self.cgfmt % k, nr))
        return ' * '.join(sf), ' * '.join(sr)

    def _db_species(self, r):
        """Species whose dB/dT the reaction's temperature term reads."""
        if not self.rev[r] or not (int(self.A['simple_mask'][r]) >= 0 or self.table_kf[r]):
            return []
        return [k for k, nr, np_ in self.slots[r] if _nu(nr, np_) != 0.0]

    def _db_loads(self, r):
        """(loads, pins) of the reaction's dB/dT values: issued before the
        previous reaction's block and pinned after it, so that waiting for
        them does not wait for that block's stores."""
        ks = self._db_species(r)
        loads = ['    PJ_OPAQUE(oa);']
        loads += ['    double dbr%d_%d = LA(%d);' % (r, k, self.sv[('DB', k)]) for k in ks]
        pins = ['    PJ_OPAQUE(dbr%d_%d);' % (r, k) for k in ks]
        return loads, pins

    def _dB(self, r):
        terms = []
        for k, nr, np_ in self.slots[r]:
            nu = _nu(nr, np_)
            if nu != 0.0:
                terms.append('%s * dbr%d_%d' % (lit(nu), r, k))
        return ' + '.join(terms) if terms else '0.0'

    def _dropi_dT(self, r, dkf):
        """dropi_dT_core of the checker: the reaction's d(rop)/dT term."""
        out = ['            double d = rf * %s;' % dkf]
        if self.rev[r]:
            out.append('            d = d + -1.0 * rr * (%s + -1.0 * (%s));' % (dkf, self._dB(r)))
        if self.thd[r] >= 0:
            out.append('            d = d * ci;')
        out.append('            tt = tt + d * V;')
        return out

    def _column_T(self, r):
        """Statements adding reaction r's temperature-column contributions to
        `tt` (create_jacobian.py:1847-3230, in the checker's kernel order)."""
        A, c = self.A, self.c
        out = ['        double tt = 0.0;']
        s = int(A['simple_mask'][r])
        if s >= 0:
            b, Ta = float(A['simple_b'][s]), float(A['simple_Ta'][s])
            out.append('        {')
            out.append('            const double dkf = (%s + %s * Tr_inv) * Tr_inv;' % (lit(b), lit(Ta)))
            out += self._dropi_dT(r, 'dkf')
            out.append('        }')
        if r in self.has_ns:
            Sf, Sr = self._ns_products(r)
            kr = 'kr' if self.rev[r] else '0.0'
            ci = ' * ci' if self.thd[r] >= 0 else ''
            out.append('        {')
            out.append('            const double Sf = %s, Sr = %s;' % (Sf, Sr))
            out.append('            tt = tt + (Sr * %s + -1.0 * Sf * kf) * V%s * P / (RU * T * T);'
                       % (kr, ci))
            out.append('        }')
        p = int(A['plog_mask'][r])
        if p >= 0:
            mp = c['plog_max_p']
            out.append('        {')
            out.append('            const double dkf = plog_dT<%d>(plog_tab + %d, %d, logPr_, Tr_inv);'
                       % (mp, 4 * mp * p, int(A['plog_np'][p]) - 1))
            out += self._dropi_dT(r, 'dkf')
            out.append('        }')
        q = int(A['cheb_mask'][r])
        if q >= 0:
            mt_, mp_ = c['cheb_max_t'], c['cheb_max_p']
            Tl = A['cheb_tlim'][2 * q:2 * q + 2]
            Pl = A['cheb_plim'][2 * q:2 * q + 2]
            out.append('        {')
            out.append('            const double dkf = cheb_dT<%d, %d>(cheb_tab + %d, %d, %d, %s, %s, '
                       '%s, %s, Tr_inv, logPr_);'
                       % (mt_, mp_, q * mt_ * mp_, int(A['cheb_nt'][q]), int(A['cheb_np'][q]),
                          lit(Tl[0]), lit(Tl[1]), lit(Pl[0]), lit(Pl[1])))
            out += self._dropi_dT(r, 'dkf')
            out.append('        }')
        if self.thd[r] >= 0:
            out += self._dci_dT(r)
        return out

    def _mod_T(self, r):
        t, ttype, last, eff = self._thd_info(r)
        mod = 1.0
        if ttype == 2:
            mod = float(last == self.ns)
        if ttype == 1 and last == self.ns:
            mod = eff
        return mod

    def _fall_params(self, r):
        A = self.A
        f = self.fall[r]
        s = int(A['simple_mask'][r])
        if int(A['fall_type'][f]):
            return ('kf', float(A['simple_b'][s]), float(A['simple_Ta'][s]),
                    'kff', float(A['fall_b'][f]), float(A['fall_Ta'][f]))
        return ('kff', float(A['fall_b'][f]), float(A['fall_Ta'][f]),
                'kf', float(A['simple_b'][s]), float(A['simple_Ta'][s]))

    def _dci_dT(self, r):
        """dci_thd_dT and dci_{lind,sri,troe}_dT (create_jacobian.py:2117-2588)."""
        A = self.A
        mod = self._mod_T(r)
        out = ['        {']
        f = self.fall[r]
        if f < 0:
            out.append('            tt = tt + -1.0 * P * %s * RU_INV * Tr_inv * Tr_inv * V * ropi;'
                       % lit(mod))
            out.append('        }')
            return out
        k0, b0, Ta0, kinf, binf, Tainf = self._fall_params(r)
        blend = int(A['fall_blend'][f])
        out.append('            const double thd_fac = -1.0 * P * %s * RU_INV * Tr_inv * Tr_inv;'
                   % lit(mod))
        out.append('            const double theta_Pr = Tr_inv * (%s - %s + (%s - %s) * Tr_inv);'
                   % (lit(b0), lit(binf), lit(Ta0), lit(Tainf)))
        out.append('            const double theta_no_Pr = thd_fac * %s / %s;' % (k0, kinf))
        if blend == 2:
            i = int(A['fall_param'][f])
            a = float(A['troe_a'][i])
            T1 = float(A['troe_T1inv'][i])
            T2 = float(A['troe_T2'][i])
            T3 = float(A['troe_T3inv'][i])
            out.append('            const double dFcent = -1.0 * %s * %s * exp(fmin(EXP_MAX, -1.0 * T * %s))'
                       ' + (%s + -1.0) * %s * exp(fmin(EXP_MAX, -1.0 * T * %s))'
                       ' + %s * Tr_inv * Tr_inv * exp(fmin(EXP_MAX, -1.0 * %s * Tr_inv));'
                       % (lit(a), lit(T1), lit(T1), lit(a), lit(T3), lit(T3), lit(T2), lit(T2)))
            out.append('            const double logFcent = log(fmax(SMALL, Fc));')
            out.append('            const double absq = Af * Af + B * B;')
            out.append('            const double absqsq = absq * absq;')
            out.append('            const double dFi = -1.0 * B * (2.0 * Af * Fc * (0.14 * Af + B) * '
                       '(Pr * theta_Pr + theta_no_Pr) * logFcent + Pr * dFcent * (2.0 * Af * '
                       '(1.1762 * Af + -1.0 * 0.67 * B) * logFcent + -1.0 * B * absq * LN10)) / '
                       '(Fc * fmax(SMALL, Pr) * absqsq * LN10);')
        elif blend == 3:
            i = int(A['fall_param'][f])
            a, b, cc, e_ = (float(A['sri_' + x][i]) for x in 'abce')
            out.append('            const double cinv = 1.0 / %s;' % lit(cc))
            out.append('            const double ec = exp(fmin(EXP_MAX, -T * cinv));')
            out.append('            const double eb = exp(fmin(EXP_MAX, -%s * Tr_inv));' % lit(b))
            out.append('            const double lv = log(fmax(SMALL, %s * eb + ec));' % lit(a))
            out.append('            const double dFi = -X * (ec * cinv - %s * %s * Tr_inv * Tr_inv * eb) / '
                       '(%s * eb + ec) + %s * Tr_inv - 2 * X * X * lv * (Pr * theta_Pr + '
                       'theta_no_Pr) * log(fmax(SMALL, Pr)) / (fmax(SMALL, Pr) * LN10SQ);'
                       % (lit(a), lit(b), lit(a), lit(e_)))
        else:
            out.append('            const double dFi = 0.0;')
        out.append('            double d = ci * (-1.0 * (Pr * theta_Pr + theta_no_Pr) / (Pr + 1.0) + dFi);')
        if not int(A['fall_type'][f]):
            out.append('            d = d + theta_Pr * ci + Fi * theta_no_Pr / (Pr + 1.0);')
        out.append('            tt = tt + d * V * ropi;')
        out.append('        }')
        return out

    def _conp_term(self, r, with_pmod):
        nu_f = -1 + sum(nr for _, nr, _ in self.slots[r])
        nu_r = -1 + sum(np_ for _, _, np_ in self.slots[r])
        out = ['            double d = %s * rf;' % lit(-1.0 * nu_f)]
        if self.rev[r]:
            out.append('            d = d + %s * rr;' % lit(float(nu_r)))
        if with_pmod and self.thd[r] >= 0:
            out.append('            d = d * ci;')
        out.append('            ee = ee + d;')
        return out

    def _column_E(self, r):
        """Statements adding reaction r's extra-variable-column contributions
        to `ee` (create_jacobian.py:362-1844)."""
        A, c = self.A, self.c
        out = ['        double ee = 0.0;', '        if (conp) {']
        s = int(A['simple_mask'][r])
        if s >= 0:
            out.append('          {')
            out += self._conp_term(r, True)
            out.append('          }')
        if r in self.has_ns:
            Sf, Sr = self._ns_products(r)
            kr = 'kr' if self.rev[r] else '0.0'
            ci = ' * ci' if self.thd[r] >= 0 else ''
            out.append('          {')
            out.append('            const double Sf = %s, Sr = %s;' % (Sf, Sr))
            out.append('            ee = ee + (Sf * kf - Sr * %s)%s * (P / (RU * T));' % (kr, ci))
            out.append('          }')
        if int(A['plog_mask'][r]) >= 0 or int(A['cheb_mask'][r]) >= 0:
            out.append('          {')
            out += self._conp_term(r, False)
            out.append('          }')
        out.append('        } else {')
        if r in self.has_ns:
            Sf, Sr = self._ns_products(r)
            kr = 'kr' if self.rev[r] else '0.0'
            ci = ' * ci' if self.thd[r] >= 0 else ''
            out.append('          {')
            out.append('            const double Sf = %s, Sr = %s;' % (Sf, Sr))
            out.append('            ee = ee + (Sf * kf - Sr * %s)%s * (V / (RU * T));' % (kr, ci))
            out.append('          }')
        p = int(A['plog_mask'][r])
        q = int(A['cheb_mask'][r])
        if p >= 0 or q >= 0:
            if p >= 0:
                mp = c['plog_max_p']
                dk = 'plog_dP<%d>(plog_tab + %d, %d, logPr_, logTr_, Tr_inv, P)' % (
                    mp, 4 * mp * p, int(A['plog_np'][p]) - 1)
            else:
                mt_, mp_ = c['cheb_max_t'], c['cheb_max_p']
                Tl = A['cheb_tlim'][2 * q:2 * q + 2]
                Pl = A['cheb_plim'][2 * q:2 * q + 2]
                dk = 'cheb_dP<%d, %d>(cheb_tab + %d, %d, %d, %s, %s, %s, %s, Tr_inv, logPr_, P)' % (
                    mt_, mp_, q * mt_ * mp_, int(A['cheb_nt'][q]), int(A['cheb_np'][q]),
                    lit(Tl[0]), lit(Tl[1]), lit(Pl[0]), lit(Pl[1]))
            out.append('          {')
            out.append('            const double dkf = %s;' % dk)
            d = 'rf - rr' if self.rev[r] else 'rf'
            out.append('            ee = ee + (%s) * dkf * V;' % d)
            out.append('          }')
        out.append('        }')
        if self.thd[r] >= 0:
            out += self._dci_dE(r)
        return out

    def _dci_dE(self, r):
        """dci_thd_dE and dci_{lind,sri,troe}_dE (create_jacobian.py:362-882)."""
        A = self.A
        t, ttype, last, eff = self._thd_info(r)
        f = self.fall[r]
        out = ['        {']
        # mod before the falloff scaling, CONP and CONV
        def mod_for(conp):
            mod = float(ttype == 1) if conp else 1.0
            if ttype == 1 and last == self.ns:
                mod = eff
            if ttype == 2:
                mod = float(last == self.ns)
            return mod
        mc, mv = mod_for(True), mod_for(False)
        if f < 0:
            if ttype != 3:
                conp_mod = '%s * P * rt_inv - ci' % lit(mc)
            else:
                conp_mod = lit(mc)
            out.append('            if (conp) ee = ee + (%s) * ropi;' % conp_mod)
            out.append('            else ee = ee + %s * rt_inv * V * ropi;' % lit(mv))
            out.append('        }')
            return out
        not_unity = int(ttype != 3)
        k0, _, _, kinf, _, _ = self._fall_params(r)
        blend = int(A['fall_blend'][f])
        out.append('            const double mod = conp ? %s * P * rt_inv * %s / %s : %s * rt_inv * %s / %s;'
                   % (lit(mc), k0, kinf, lit(mv), k0, kinf))
        out.append('            const double theta = conp ? (mod + -1.0 * Pr * %d) : mod;' % not_unity)
        if blend == 2:
            out.append('            double absqsq = Af * Af + B * B;')
            out.append('            absqsq = absqsq * absqsq;')
            out.append('            const double dFi = -2.0 * Af * B * log(fmax(SMALL, Fc)) * (0.14 * Af + B) '
                       '* theta / (fmax(SMALL, Pr) * absqsq * LN10);')
        elif blend == 3:
            i = int(A['fall_param'][f])
            a, b = float(A['sri_a'][i]), float(A['sri_b'][i])
            cc = float(A['sri_c'][i])
            out.append('            const double cinv = 1.0 / %s;' % lit(cc))
            out.append('            double inner;')
            out.append('            if (conp) {')
            out.append('                const double exp_cinv = exp(fmin(EXP_MAX, T * cinv));')
            out.append('                const double exp_b = exp(fmin(EXP_MAX, %s * Tr_inv));' % lit(b))
            out.append('                const double exp_bc = exp(fmin(EXP_MAX, -%s * Tr_inv - T * cinv));'
                       % lit(b))
            out.append('                inner = log(fmax(SMALL, (%s * exp_cinv + exp_b) * exp_bc));' % lit(a))
            out.append('            } else {')
            out.append('                const double exp_b = exp(fmin(EXP_MAX, -%s * Tr_inv));' % lit(b))
            out.append('                const double exp_cinv = exp(fmin(EXP_MAX, -T * cinv));')
            out.append('                inner = log(fmax(SMALL, %s * exp_b + exp_cinv));' % lit(a))
            out.append('            }')
            out.append('            const double dFi = -2 * X * X * theta * inner * log(fmax(SMALL, Pr)) / '
                       '(fmax(SMALL, Pr) * LN10SQ);')
        else:
            out.append('            const double dFi = 0.0;')
        out.append('            const double num = conp ? (-mod + Pr * %d) : -mod;' % not_unity)
        out.append('            double d = ci * (num / (Pr + 1) + dFi);')
        if not int(A['fall_type'][f]):
            out.append('            if (conp) d = d + Fi * mod / (Pr + 1) + -1.0 * %d * ci;' % not_unity)
            out.append('            else d = d + Fi * mod / (Pr + 1);')
        out.append('            ee = ee + (conp ? d * ropi : d * V * ropi);')
        out.append('        }')
        return out

    def _fall_factor(self, r):
        """fall_dnj_factor of the checker: the falloff dFi/dPr factor folded
        with the dPr/dn term (create_jacobian.py:3780-3890)."""
        A = self.A
        f = self.fall[r]
        blend = int(A['fall_blend'][f])
        out = []
        if blend == 2:
            out.append('            double dFi = Af * Af + B * B;')
            out.append('            dFi = -2.0 * Af * B * (0.14 * Af + B) * log(fmax(SMALL, Fc)) / '
                       '(fmax(SMALL, Pr) * dFi * dFi * LN10);')
        elif blend == 3:
            i = int(A['fall_param'][f])
            a, b, cc = float(A['sri_a'][i]), float(A['sri_b'][i]), float(A['sri_c'][i])
            out.append('            const double lv = log(fmax(SMALL, %s * exp(fmin(EXP_MAX, -%s / T)) + '
                       'exp(fmin(EXP_MAX, -T / %s))));' % (lit(a), lit(b), lit(cc)))
            out.append('            const double dFi = -2.0 * X * X * lv * log(fmax(SMALL, Pr)) / '
                       '(fmax(SMALL, Pr) * LN10SQ);')
        else:
            out.append('            const double dFi = 0.0;')
        k0, _, _, kinf, _, _ = self._fall_params(r)
        if int(A['fall_type'][f]):
            out.append('            const double Fi_fac = dFi;')
        else:
            out.append('            const double Fi_fac = Pr * dFi + 1.0;')
        out.append('            const double fac = %s * (Fi * Fi_fac + -1.0 * ci) / (%s * (Pr + 1.0));'
                   % (k0, kinf))
        return out

    def _store(self, name, i, val):
        if self.nt_store:
            return '        PJ_NTST(PJ_SCP(%d), %s);' % (self.sv[(name, i)], val)
        return '        SC(%d) = %s;' % (self.sv[(name, i)], val)

    def _reaction(self, r):
        """The species-rate module's reaction block (Gibbs terms from LDS),
        plus the values of the reaction that the Jacobian's column kernels
        use: the temperature-column contributions are added to the rows'
        accumulators, the rest is stored."""
        out = super(LaneJacSource, self)._reaction(r)
        assert out[-1] == '    }'
        out = [re.sub(r'\bb(\d+)\b', r'BS(\1)', l) for l in out]
        # a fresh opaque base for the reaction's scratch loads, so that loads
        # of one value by several reactions are not merged into one long-lived
        # register
        body = out[:-1]
        if self.table_kf[r]:
            body.append(self._store('KF', r, 'kf'))
        if self.thd[r] >= 0:
            body.append(self._store('CI', r, 'ci'))
            body.append('        const double ropi = %s;' % ('rf + -1.0 * rr' if self.rev[r] else 'rf'))
            if self.fall[r] >= 0:
                body.append('        {')
                body += self._fall_factor(r)
                body.append('    ' + self._store('G', r, 'ropi * fac'))
                body.append('        }')
            else:
                body.append(self._store('G', r, 'ropi'))
        body += self._column_T(r)
        for k, nr, np_ in self.slots[r]:
            if k != self.ns and _nu(nr, np_) != 0.0:
                body.append('        jt%d = jt%d + %s * tt;' % (k, k, lit(_nu(nr, np_))))
        body += self._column_E(r)
        if self.ee_conv[r]:
            body.append(self._store('EE', r, 'ee'))
        else:
            body.append('        if (conp) ' + self._store('EE', r, 'ee').lstrip())
        body.append('    }')
        return body

    def _falloff(self, r, f):
        # the Troe and SRI intermediates stay in scope for the Jacobian terms
        out = super(LaneJacSource, self)._falloff(r, f)
        A = self.A
        if int(A['fall_blend'][f]) != 2:
            out.append('        const double Fc = 0.0, Af = 0.0, B = 0.0; (void)Fc; (void)Af; (void)B;')
        if int(A['fall_blend'][f]) != 3:
            out.append('        const double X = 0.0; (void)X;')
        return out

    def _kernel_a(self):
        A, Ns = self.A, self.Ns
        ns = self.ns
        out = ['// kernel A of state s: species rates, the temperature column and the',
               '// per-reaction and per-species values of the other columns, into scratch',
               '// entry v at scr[v * PJ_SS + sl]; the Gibbs terms in LDS (bsh[k * lanes +',
               '// lane]); conp a template parameter (with a run-time flag the two forms',
               '// of the extra-variable column in every reaction kept many more values',
               '// alive)',
               'template <int conp>',
               'PJ_FN void ja_state(long s, long sl, const double* __restrict__ param,',
               '                    const double* __restrict__ phi, double* __restrict__ scr,',
               '                    double* __restrict__ bsh, int ln, long sN, long sI) {',
               '#define SC(v) (*PJ_SCP(v))',
               '#define PJ_SCP(v) (scr + (long)(v) * PJ_SS + sl)',
               '#define BS(k) bsh[(k) * PJ_LANES + ln]',
               '    const double* ps = phi + s * sN;',
               '    const double T = ps[0], x1 = ps[sI], pv = param[s];',
               '    const double P = conp ? pv : x1;',
               '    const double V = conp ? x1 : pv;',
               '    const double Tc = fmin(T_MAX, fmax(T_MIN, T));',
               '    const double logT = log(Tc);',
               '    const double Tinv = 1.0 / Tc;',
               '    const double logP = log(fmax(SMALL, P));',
               '    const double V_inv = 1.0 / fmax(V_MIN, V);',
               '    // the Jacobian kernels take T, P and V unclamped',
               '    const double Tr_inv = 1.0 / T;',
               '    const double logPr_ = log(fmax(SMALL, P));',
               '    const double logTr_ = log(fmax(SMALL, T));',
               '    const double rt_inv = 1.0 / (RU * T);',
               '    const double Vinv = 1.0 / V;',
               '    (void)logP; (void)logPr_; (void)logTr_; (void)rt_inv;']
        self.cgfmt = 'fmax(SMALL, c%d)'
        out.append('    long oa = sl;')
        out.append('#define LA(v) scr[(long)(v) * PJ_SS + oa]')
        species = [re.sub(r'const double b(\d+) = ', r'BS(\1) = ', l) for l in self._species()]
        out += species
        # d b_k / dT at the clamped temperature, kept in scratch; the
        # coefficients are selected from an opaque copy of the temperature so
        # that the selections of the Gibbs terms are not kept alive for them
        out += ['    double Td = Tc;', '    PJ_OPAQUE(Td);']
        for m, i in self._mid.items():
            out.append('    const bool dq%d = Td < %s;' % (i, lit(m)))
        for k in range(Ns):
            out.append(self._store('DB', k, 'dgibbs(Tc, Tinv, %s)' % self._coeffs(k, 6, 'dq')))
        out.append('    double %s;' % ', '.join('w%d = 0.0' % k for k in range(Ns)))
        out.append('    double %s;' % ', '.join('jt%d = 0.0' % k for k in range(Ns - 1)))
        # reaction r + D's scratch loads (through a fresh opaque offset, so
        # that loads of one value by several reactions are not merged into one
        # long-lived register) are issued before reaction r's block, and
        # reaction r + 1's are pinned after it: each reaction's loads are in
        # flight over D reaction blocks, and waiting for them does not wait
        # for the stores issued after them
        D = self.db_depth
        dl = [self._db_loads(r) for r in range(self.Nr)] + [([], [])] * D
        for r in range(D):
            out += dl[r][0]
        out += dl[0][1]
        for r in range(self.Nr):
            out += dl[r + D][0]
            out += self._reaction(r)
            out += dl[r + 1][1]
        # energies, heat capacities, temperature rate; the temperature column's
        # corner entries; the per-species terms of the other rows
        out += ['    double Tq = Tc;', '    PJ_OPAQUE(Tq);']
        for m, i in self._mid.items():
            out.append('    const bool hq%d = Tq < %s;' % (i, lit(m)))
        out.append('    long op = s * sN;')
        out.append('    double e_ns, cp_ns;')
        out.append('    thermo(Tq, conp, %s, e_ns, cp_ns);' % self._coeffs(ns, 6, 'hq'))
        out.append('    const double dcp_ns = dcp(Tq, %s);' % self._coeffs_sel(ns, (1, 2, 3, 4), 'hq'))
        out.append('    double lower = 0.0, upper = 0.0;')
        out.append('    double dTsum = (cp_ns * Tr_inv + -1.0 * dcp_ns) * c%d;' % ns)
        out.append('    double dTsum2 = 0.0, rate_sum = 0.0, esum = 0.0;')
        out.append('    const double fac = T / V;')
        # the concentrations again, from the state: loads issued together
        out.append('    PJ_OPAQUE(op);')
        for k in range(Ns - 1):
            out.append('    double nn%d = phi[op + %d * sI];' % (k, k + 2))
        for k in range(Ns - 1):
            out.append('    PJ_OPAQUE(nn%d);' % k)
        for k in range(Ns - 1):
            mw = float(A['mw_post'][k])
            mwf = lit(1.0 + -1.0 * mw)
            out.append('    {')
            out.append('        const double cc%d = fmax(SMALL, nn%d) * V_inv;' % (k, k))
            out.append('        double e, cp;')
            out.append('        thermo(Tq, conp, %s, e, cp);' % self._coeffs(k, 6, 'hq'))
            out.append('        const double dc = dcp(Tq, %s);' % self._coeffs_sel(k, (1, 2, 3, 4), 'hq'))
            out.append('        lower = __dadd_rn(lower, __dmul_rn(cc%d, cp));' % k)
            out.append('        upper = __dadd_rn(upper, __dmul_rn(e, w%d));' % k)
            out.append('        dTsum = dTsum + (cp_ns * Tr_inv + -1.0 * dc) * cc%d;' % k)
            out.append('        dTsum2 = dTsum2 + (cp + -1.0 * cp_ns) * cc%d;' % k)
            out.append('        const double ew = e + -1.0 * e_ns * %s;' % lit(mw))
            out.append(self._store('EW', k, 'ew'))
            out.append(self._store('DCP', k, 'cp + -1.0 * cp_ns'))
            out.append(self._store('WD', k, 'w%d' % k))
            out.append(self._store('JT', k, 'jt%d' % k))
            # dTdot/dT and dEdot/dT sums (create_jacobian.py:1847-2114)
            out.append('        rate_sum = rate_sum + Vinv * jt%d * (-ew);' % k)
            out.append('        rate_sum = rate_sum + w%d * (-1.0 * cp + %s * cp_ns);' % (k, lit(mw)))
            out.append('        esum = esum + (conp ? %s * (Vinv * jt%d + Tr_inv * w%d) : '
                       '%s * (jt%d * fac + w%d));' % (mwf, k, k, mwf, k, k))
            out.append('    }')
        out.append('    lower = __dadd_rn(lower, __dmul_rn(c%d, cp_ns));' % ns)
        out.append('    upper = __dadd_rn(upper, __dmul_rn(e_ns, w%d));' % ns)
        out.append('    const double dTdt = 0.0 + -1.0 * upper / lower;')
        out.append(self._store('DTDT', 0, 'dTdt'))
        out.append(self._store('CPTOT', 0, 'lower'))
        out.append(self._store('DTSUM2', 0, 'dTsum2'))
        out.append(self._store('CPNS', 0, 'cp_ns'))
        out.append('    const double JTT = (dTdt * dTsum + rate_sum) / lower;')
        out.append('    double JET;')
        out.append('    if (conp) {')
        out.append('        JET = V * Tr_inv * (JTT + -1.0 * Tr_inv * dTdt);')
        out.append('        JET = JET + RU * T * V * esum / P;')
        out.append('    } else {')
        out.append('        JET = P * (JTT - dTdt * Tr_inv) * Tr_inv;')
        out.append('        JET = JET + RU * esum;')
        out.append('    }')
        out.append(self._store('JTT', 0, 'JTT'))
        out.append(self._store('JET', 0, 'JET'))
        out.append('#undef SC')
        out.append('#undef PJ_SCP')
        out.append('#undef LA')
        out.append('#undef BS')
        out.append('}')
        return out

    def _coeffs_sel(self, k, idx, flag):
        A = self.A
        lo = A['a_lo'][7 * k:7 * k + 7]
        hi = A['a_hi'][7 * k:7 * k + 7]
        sel = '%s%d' % (flag, self._mid[float(A['T_mid'][k])])
        parts = []
        for i in idx:
            if lo[i] == hi[i]:
                parts.append(lit(lo[i]))
            else:
                parts.append('%s ? %s : %s' % (sel, lit(lo[i]), lit(hi[i])))
        return ', '.join(parts)

    # ------------------------------------------------------------ kernel B
    def _positions(self):
        """Output index of every pattern entry in the four layouts: sparse C
        (CRS), sparse F (CCS), full C ([row][col]), full F ([col][row])."""
        NN = self.NN
        crs = {}
        rows, cols = self.mt.sparse_rows_cols('C')
        for q, (a, b) in enumerate(zip(rows, cols)):
            crs[(int(a), int(b))] = q
        ccs = {}
        rows, cols = self.mt.sparse_rows_cols('F')
        for q, (a, b) in enumerate(zip(rows, cols)):
            ccs[(int(a), int(b))] = q
        return crs, ccs

    def _entry(self, row, col, val, out, indent='        '):
        """Store of entry (row, col) at its output index in the kernel's
        layout L (a template parameter, so the index is a literal)."""
        if (row, col) not in self.pattern:
            raise RuntimeError('entry (%d, %d) has contributions but is not in the '
                               'sparsity pattern' % (row, col))
        self.entries.append((row, col))
        NN = self.NN
        addr = 'o + (long)PJ_POS(%d, %d, %d, %d) * jI' % (
            self._crs[(row, col)], self._ccs[(row, col)], row * NN + col, col * NN + row)
        if self.nt_store:
            out.append('%sPJ_STO(%s, %s);' % (indent, addr, val))
        else:
            out.append('%s*(%s) = %s;' % (indent, addr, val))

    def _sf_sr(self, r, j):
        """S' and S'' of dRopidnj for column species j (create_jacobian.py
        :4169-4400): the stoichiometric factor of j and the products of the
        other concentrations, with the checker's guarded concentrations."""
        for k, nf, nr in self.slots[r]:
            if k == j:
                sf0, sr0 = nf, nr
        sf_terms = [lit(float(sf0))] if sf0 != 1 else []
        sr_terms = [lit(float(sr0))] if sr0 != 1 else []
        for k, nf, nr in self.slots[r]:
            if k == j:
                nf, nr = nf - 1, nr - 1
            if nf and sf0:
                sf_terms.append(_powi('cg%d' % k, nf))
            if nr and sr0:
                sr_terms.append(_powi('cg%d' % k, nr))
        Sf = (' * '.join(sf_terms) if sf_terms else '1.0') if sf0 else None
        Sr = (' * '.join(sr_terms) if sr_terms else '1.0') if sr0 else None
        return Sf, Sr

    def _kr_product(self, r, kf, kf_exponent=None):
        """kr of reversible reaction r, the species-rate module's product form
        with the Gibbs terms in LDS."""
        return lanegen.kr_product(self.slots[r], int(self.A['nu_sum'][r]), kf, 'BS(%d)',
                                  kf_exponent)

    def _kf_kr(self, r, need_kr, indent='        ', kf_loaded=None, need_kf=True):
        """Definitions of kf<r> (if need_kf) and kr<r> (if need_kr) in kernel
        B: the Arrhenius rate constant recomputed from its literals exactly as
        kernel A evaluates it (PLOG and Chebyshev rate constants are read back
        from scratch), and the reverse rate constant kf / (P_sum exp(B_sum))
        of the Gibbs terms (rate_subs.py:1627-1785) evaluated as kf (1 /
        P_sum) exp(-B_sum), a product instead of a division; where the column
        needs kr alone and kf is an Arrhenius exponential, the two exponents
        are added and one exponential taken.  Either form is the quotient to a
        few units in the last place."""
        A = self.A
        exp_form = False
        if self.table_kf[r]:
            kf = kf_loaded or 'LD(%d)' % self.sv[('KF', r)]
        else:
            kf = self._kf(r)
            si = int(A['simple_mask'][r])
            exp_form = si >= 0 and int(A['simple_form'][si]) >= 2
            # the column's opaque copies of the temperature terms
            kf = re.sub(r'\bTc\b', 'Tcc', kf)
            kf = re.sub(r'\blogT\b', 'logTc', kf)
            kf = re.sub(r'\bTinv\b', 'Tinvc', kf)
        need_kr = need_kr and self.rev[r]
        out = []
        if need_kf or not need_kr:
            decl, e = self._opaque_literals(kf, 'kq%d_' % r, indent)
            out += decl
            out.append('%sconst double kf%d = %s;' % (indent, r, e))
        if need_kr:
            if not need_kf and exp_form:
                assert kf.startswith('ex(') and kf.endswith(')')
                kr = self._kr_product(r, None, kf[3:-1])
            else:
                kr = self._kr_product(r, 'kf%d' % r if need_kf else '(%s)' % kf)
            decl, e = self._opaque_literals(kr, 'rq%d_' % r, indent)
            out += decl
            out.append('%sconst double kr%d = %s;' % (indent, r, e))
        return out

    # a literal of the generated code (lit()): a decimal point or an exponent,
    # negative ones in parentheses
    _LITERAL = re.compile(r'(?<![\w.])(\(-(?:\d+\.\d*(?:e[+-]?\d+)?|\d+e[+-]?\d+)\)'
                          r'|(?:\d+\.\d*(?:e[+-]?\d+)?|\d+e[+-]?\d+))')
    # literals that are inline constants of the instruction set
    _INLINE = {0.0, 0.5, 1.0, 2.0, 4.0, -0.5, -1.0, -2.0, -4.0}

    def _opaque_literals(self, expr, prefix, indent):
        """The literals of `expr` (other than inline constants) as local
        copies, each pinned in scalar registers by an empty asm statement:
        without it the compiler keeps a constant that several columns use
        (a reaction's log A, 1 / P_sum) in a vector register from the first
        column to the last, and spills it.  (declarations, expr)"""
        decl, names = [], {}

        def sub(m):
            tok = m.group(1)
            if float(tok.strip('()')) in self._INLINE:
                return tok
            if tok not in names:
                nm = '%s%d' % (prefix, len(names))
                names[tok] = nm
                decl.append('%sdouble %s = %s; PJ_OPQ(%s);' % (indent, nm, tok, nm))
            return names[tok]
        return decl, self._LITERAL.sub(sub, expr)

    def _cg_defs(self, species, indent='        '):
        """Guarded concentrations of `species` recomputed from the state (the
        bath gas is kept from the start of the kernel)."""
        out = []
        for k in sorted(species):
            if k == self.ns:
                continue
            out.append('%sconst double cg%d = fmax(SMALL, fmax(SMALL, phi[op + %d * sI]) * V_inv);'
                       % (indent, k, k + 2))
        return out

    def _kernel_b(self):
        A, Ns, ns = self.A, self.Ns, self.ns
        self.entries = []
        self._crs, self._ccs = self._positions()
        LD = lambda v: 'LD(%d)' % v
        LDN = lambda v: ('LDN(%d)' if self.nt_loads else 'LD(%d)') % v
        KF = lambda r: 'kf%d' % r
        KR = lambda r: 'kr%d' % r
        CI = lambda r: LD(self.sv[('CI', r)])
        G = lambda r: LD(self.sv[('G', r)])
        out = ['// kernel B of state s: every Jacobian entry, column by column, stored in',
               '// layout L (0 sparse C, 1 sparse F, 2 full C, 3 full F).  The rate',
               '// constants are recomputed from their literals (the Gibbs terms of the',
               '// reverse ones in LDS), the energy weights and bath-gas terms of the rows',
               '// stay in registers, and each column adds its reactions\' terms to its',
               '// rows in reaction order.  Scratch and state loads go through offsets',
               '// that an empty asm statement renews before every column, and the',
               '// temperature terms are renewed the same way, so that nothing computed or',
               '// loaded for one column is kept in a register for the next',
               '#define PJ_POS(a, b, c, d) (L == 0 ? (a) : L == 1 ? (b) : L == 2 ? (c) : (d))',
               '// F-order entries are whole lines written once: stored non-temporally; in C',
               '// order a lane fills its own state\'s lines entry by entry, and plain stores',
               '// let the L2 merge them',
               '#define PJ_STO(p, v) do { if (L & 1) PJ_NTST(p, v); else *(p) = (v); } while (0)',
               'template <int L>',
               'PJ_FN void jb_state(long s, long sl, const double* __restrict__ param,',
               '                    const double* __restrict__ phi, double* __restrict__ scr,',
               '                    double* __restrict__ jac, double* __restrict__ bsh, int ln,',
               '                    long sN, long sI, long jN, long jS, long jI, int conp) {',
               '#define BS(k) bsh[(k) * PJ_LANES + ln]',
               '    const double* ps = phi + s * sN;',
               '    double* o = jac + (s - jS) * jN;',
               '    long ob = sl, op = s * sN;',
               '#define LD(v) scr[(long)(v) * PJ_SS + ob]',
               '// scratch values read once: non-temporal loads, which leave the L2 to',
               '// the state and the values that several columns read',
               '#define LDN(v) PJ_NTLD(&scr[(long)(v) * PJ_SS + ob])',
               '    const double T = ps[0], x1 = ps[sI], pv = param[s];',
               '    const double P = conp ? pv : x1;',
               '    const double V = conp ? x1 : pv;',
               '    const double Tc = fmin(T_MAX, fmax(T_MIN, T));',
               '    const double V_inv = 1.0 / fmax(V_MIN, V);',
               '    const double Vinv = 1.0 / V;',
               '    const double dTdt = LD(%d), cptot = LD(%d);'
               % (self.sv[('DTDT', 0)], self.sv[('CPTOT', 0)])]
        # the bath gas by difference, exactly as kernel A computes it; the
        # state's moles loaded together and pinned before their first use
        out.append('    double nsum = 0.0;')
        out += ['    double nb%d = ps[%d * sI];' % (k, k + 2) for k in range(Ns - 1)]
        out += ['    PJ_OPAQUE(nb%d);' % k for k in range(Ns - 1)]
        for k in range(Ns - 1):
            out.append('    nsum = nsum + fmax(SMALL, nb%d);' % k)
        out.append('    const double cg%d = fmax(SMALL, __dadd_rn(P / (RU * Tc), '
                   '__dmul_rn(-1.0 * fmax(SMALL, nsum), V_inv)));' % ns)
        # the Gibbs terms of the reverse rate constants, as kernel A has them
        out += ['    const double logT = log(Tc);',
                '    const double Tinv = 1.0 / Tc;',
                '    const double logP = log(fmax(SMALL, P));',
                '    (void)logP; (void)logT;']
        for m, i in self._mid.items():
            out.append('    const bool lo%d = Tc < %s;' % (i, lit(m)))
        gibbs_used = sorted({kk for r in range(self.Nr) if self.rev[r]
                             for kk, nr, np_ in self.slots[r] if np_ != nr})
        for k in gibbs_used:
            out.append('    BS(%d) = gibbs(Tc, logT, Tinv, %s);' % (k, self._coeffs(k, 7)))
        # the energy weights of the temperature row, read once
        for k in range(Ns - 1):
            out.append('    double ew%d = %s;' % (k, LDN(self.sv[('EW', k)])))
        out += ['    PJ_OPAQUE(ew%d);' % k for k in range(Ns - 1)]

        # ---- the temperature column (kernel A's) and the extra-variable
        # column.  Scratch values are loaded in batches, each batch pinned by
        # an empty asm statement before its first use, so that the loads of a
        # batch are in flight together
        def batched(names_vals, consume, indent='    ', size=16):
            """Loads `name = value` in batches of `size`, each batch followed
            by consume(batch) statements; batch k + 1's loads are issued
            before batch k is pinned and consumed, so that one batch is in
            flight while the previous one is used."""
            batches = [names_vals[b0:b0 + size] for b0 in range(0, len(names_vals), size)]

            def issue(batch):
                return (['%sPJ_OPAQUE(ob);' % indent] +
                        ['%sdouble %s = %s;' % (indent, nm, val) for nm, val in batch])
            body = issue(batches[0]) if batches else []
            for k, batch in enumerate(batches):
                if k + 1 < len(batches):
                    body += issue(batches[k + 1])
                body += ['%sPJ_OPAQUE(%s);' % (indent, nm) for nm, _ in batch]
                body += consume(batch)
            return body

        out.append('    {  // column 0')
        col0 = [(i + 2, 'jt_%d' % i, LDN(self.sv[('JT', i)])) for i in range(Ns - 1)
                if (i + 2, 0) in self.pattern]
        col0 += [(0, 'jt_T', LDN(self.sv[('JTT', 0)])), (1, 'jt_E', LDN(self.sv[('JET', 0)]))]
        rowof = {nm: row for row, nm, _ in col0}

        def store0(batch):
            body = []
            for nm, _ in batch:
                self._entry(rowof[nm], 0, nm, body)
            return body
        out += batched([(nm, val) for _, nm, val in col0], store0, '        ')
        out.append('    }')
        for mode in (1, 0):
            out.append('    %s {  // column 1' % ('if (conp)' if mode else 'else'))
            rxn = [r for r in range(self.Nr) if mode or self.ee_conv[r]]
            rows = sorted({k for r in rxn for k, nr, np_ in self.slots[r]
                           if k != ns and _nu(nr, np_) != 0.0})
            for i in rows:
                out.append('        double e_%d = 0.0;' % i)

            def add(batch):
                body = []
                for nm, _ in batch:
                    r = int(nm[3:])
                    for k, nr, np_ in self.slots[r]:
                        if k != ns and _nu(nr, np_) != 0.0:
                            body.append('        e_%d = e_%d + %s * %s;' % (k, k, lit(_nu(nr, np_)), nm))
                return body
            # each value once, its rows' sums in reaction order
            out += batched([('ee_%d' % r, LDN(self.sv[('EE', r)])) for r in rxn], add, '        ')
            if mode:
                out += batched([('wd_%d' % i, LDN(self.sv[('WD', i)])) for i in range(Ns - 1)],
                               lambda b: [], '        ')
            out.append('        double specsum = 0.0, esum = 0.0;')
            for i in range(Ns - 1):
                mwf = lit(1.0 + -1.0 * float(A['mw_post'][i]))
                if i in rows:
                    self._entry(i + 2, 1, 'e_%d' % i, out)
                    if mode:
                        out.append('        specsum = specsum + ew%d * (e_%d + -1.0 * wd_%d);' % (i, i, i))
                    else:
                        out.append('        specsum = specsum + ew%d * e_%d;' % (i, i))
                    out.append('        esum = esum + %s * e_%d;' % (mwf, i))
                else:
                    if (i + 2, 1) in self.pattern:
                        self._entry(i + 2, 1, '0.0', out)
                    if mode:
                        out.append('        specsum = specsum + ew%d * (-1.0 * wd_%d);' % (i, i))
            if mode:
                out.append('        const double JTE = (dTdt * %s + -1.0 * specsum) * (1.0 / (cptot * V));'
                           % LD(self.sv[('DTSUM2', 0)]))
                out.append('        const double JEE = (V * JTE + dTdt) / T + RU * T * esum / P;')
            else:
                out.append('        const double spec_inv = 1.0 / cptot;')
                out.append('        double JTE = 0.0 - (dTdt * %s / (RU * T)) * spec_inv;'
                           % LD(self.sv[('CPNS', 0)]))
                out.append('        JTE = JTE - (specsum * Vinv) * spec_inv;')
                out.append('        const double JEE = (P * JTE + dTdt) / T + RU * T * esum / V;')
            self._entry(0, 1, 'JTE', out)
            self._entry(1, 1, 'JEE', out)
            out.append('    }')
        # the entries were listed twice (one branch each): keep one list
        seen = set()
        self.entries = [e for e in self.entries if not (e in seen or seen.add(e))]

        # ---- bath-gas terms shared by every species column of a row, kept in
        # registers
        U = {}
        Dns = []
        for r in sorted(self.has_ns):
            Sr_t = ['1.0']
            Sf_t = ['1.0']
            for k, nf, nr in self.slots[r]:
                if k == ns:
                    Sr_t.append(lit(float(nr)))
                    nr -= 1
                    Sf_t.append(lit(float(nf)))
                    nf -= 1
                if nr:
                    Sr_t.append(_powi('cg%d' % k, nr))
                if nf:
                    Sf_t.append(_powi('cg%d' % k, nf))
            kr = KR(r) if self.rev[r] else '0.0'
            ci = ' * ' + CI(r) if self.thd[r] >= 0 else ''
            Dns.append((r, '(%s * %s + -1.0 * %s * %s)%s'
                        % (kr, ' * '.join(Sr_t), KF(r), ' * '.join(Sf_t), ci)))
            for k, nr, np_ in self.slots[r]:
                if k != ns and _nu(nr, np_) != 0.0:
                    U.setdefault(k, []).append('Dns%d * %s' % (r, lit(_nu(nr, np_))))
        for r in range(self.Nr):
            t = self.thd[r]
            if t < 0:
                continue
            tt, ttype, last, eff = self._thd_info(r)
            if last != ns:
                continue
            dci = 0.0
            if ttype == 2:
                dci = -1.0
            if ttype == 1:
                dci = 1.0 + -1.0 * eff
            if dci == 0.0:
                continue
            for k, nr, np_ in self.slots[r]:
                if k != ns and _nu(nr, np_) != 0.0:
                    U.setdefault(k, []).append('%s * %s' % (lit(_nu(nr, np_) * dci), G(r)))
        uname = {}
        for i, k in enumerate(sorted(U)):
            uname[k] = ('bsh[(%d) * PJ_LANES + ln]' % (Ns + i)) if self.u_lds else 'U%d' % k
        self._nU = len(U) if self.u_lds else 0
        if U:
            if not self.u_lds:
                out.append('    double %s;' % ', '.join('U%d' % k for k in sorted(U)))
            out.append('    {')
            out.append('        const double Tcc = Tc, logTc = logT, Tinvc = Tinv;')
            out.append('        (void)Tcc; (void)logTc; (void)Tinvc;')
            sp = set()
            for r, _ in Dns:
                sp.update(k for k, _, _ in self.slots[r])
            out += self._cg_defs(sp)
            for r, expr in Dns:
                out += self._kf_kr(r, True)
                out.append('        const double Dns%d = %s;' % (r, expr))
            for k in sorted(U):
                out.append('        %s = %s;' % (uname[k], ' + '.join(U[k])))
            out.append('    }')

        # ---- species columns
        thd_cols = {}
        for r in range(self.Nr):
            t = self.thd[r]
            if t < 0:
                continue
            ttype = int(A['thd_type'][t])
            for e in range(int(A['thd_off'][t]), int(A['thd_off'][t + 1])):
                j = int(A['thd_spec'][e])
                if j == ns:
                    continue
                dci = 0.0
                if ttype == 2:
                    dci = 1.0
                if ttype == 1:
                    dci = float(A['thd_eff'][e]) + -1.0
                if dci != 0.0:
                    thd_cols.setdefault(j, []).append((r, dci))
        mwf = [1.0 + -1.0 * float(A['mw_post'][k]) for k in range(Ns - 1)]
        # each column in three parts: its loads (state and scratch values,
        # issued together and then pinned by an empty asm statement, so that
        # the column waits once), its arithmetic, and its stores; column j+1's
        # loads are issued before column j's arithmetic (early_loads; else
        # before its stores) and pinned after its stores, so that they are in
        # flight during the arithmetic and the wait for them does not include
        # those stores (one counter covers loads and stores, in issue order)
        parts = []
        for j in range(Ns - 1):
            col = j + 2
            rxns = [r for r in range(self.Nr) if any(k == j for k, _, _ in self.slots[r])]
            terms = []
            for r in rxns:
                Sf, Sr = self._sf_sr(r, j)
                if Sf is None and (Sr is None or not self.rev[r]):
                    continue
                terms.append((r, Sf, Sr))
            sp = set()
            rows = set()
            for r, Sf, Sr in terms:
                # only the concentrations the column's products use
                for expr in (Sf, Sr if self.rev[r] else None):
                    if expr is not None:
                        sp.update(int(x) for x in re.findall(r'\bcg(\d+)\b', expr))
                rows.update(k for k, nr, np_ in self.slots[r] if k != ns and _nu(nr, np_) != 0.0)
            for r, dci in thd_cols.get(j, []):
                rows.update(k for k, nr, np_ in self.slots[r] if k != ns and _nu(nr, np_) != 0.0)
            rows.update(k for k in U)
            rows = sorted(rows)
            loads, pins = ['    PJ_OPAQUE(ob);', '    PJ_OPAQUE(op);'], []
            for k in sorted(sp):
                if k != ns:
                    loads.append('    double n%d_%d = phi[op + %d * sI];' % (j, k, k + 2))
                    pins.append('n%d_%d' % (j, k))
            for r, _, _ in terms:
                if self.thd[r] >= 0:
                    loads.append('    double ci%d_%d = %s;' % (j, r, CI(r)))
                    pins.append('ci%d_%d' % (j, r))
                if self.table_kf[r]:
                    loads.append('    double kt%d_%d = LD(%d);' % (j, r, self.sv[('KF', r)]))
                    pins.append('kt%d_%d' % (j, r))
            for r, dci in thd_cols.get(j, []):
                loads.append('    double g%d_%d = %s;' % (j, r, G(r)))
                pins.append('g%d_%d' % (j, r))
            loads.append('    double dcp%d = %s;' % (j, LDN(self.sv[('DCP', j)])))
            pins.append('dcp%d' % j)
            comp = ['    // column %d' % col]
            for k in rows:
                comp.append('    double a%d_%d = %s;' % (j, k, uname[k] if k in U else '0.0'))
            comp.append('    double jt%d, je%d;' % (j, j))
            comp.append('    {')
            comp.append('        double Tcc = Tc, logTc = logT, Tinvc = Tinv;')
            comp.append('        PJ_OPAQUE(Tcc); PJ_OPAQUE(logTc); PJ_OPAQUE(Tinvc);')
            comp.append('        (void)Tcc; (void)logTc; (void)Tinvc;')
            for k in sorted(sp):
                if k != ns:
                    comp.append('        const double cg%d = fmax(SMALL, fmax(SMALL, n%d_%d) * V_inv);'
                                % (k, j, k))
            for r, Sf, Sr in terms:
                pr = []
                if Sf is not None:
                    pr.append('%s * %s' % (KF(r), Sf))
                if Sr is not None and self.rev[r]:
                    pr.append('-1.0 * %s * %s' % (KR(r), Sr))
                comp.append('        {')
                comp += self._kf_kr(r, Sr is not None, '            ',
                                    'kt%d_%d' % (j, r) if self.table_kf[r] else None,
                                    need_kf=Sf is not None)
                expr = ' + '.join(pr)
                if self.thd[r] >= 0:
                    expr = '(%s) * ci%d_%d' % (expr, j, r)
                comp.append('            const double D = %s;' % expr)
                for k, nr, np_ in self.slots[r]:
                    if k != ns and _nu(nr, np_) != 0.0:
                        comp.append('            a%d_%d = a%d_%d + %s * D;'
                                    % (j, k, j, k, lit(_nu(nr, np_))))
                comp.append('        }')
            for r, dci in thd_cols.get(j, []):
                comp.append('        {')
                comp.append('            const double g = %s * g%d_%d;' % (lit(dci), j, r))
                for k, nr, np_ in self.slots[r]:
                    if k != ns and _nu(nr, np_) != 0.0:
                        comp.append('            a%d_%d = a%d_%d + %s * g;'
                                    % (j, k, j, k, lit(_nu(nr, np_))))
                comp.append('        }')
            comp.append('        double sT = 0.0, sE = 0.0;')
            for k in rows:
                comp.append('        sT = sT + ew%d * a%d_%d;' % (k, j, k))
                comp.append('        sE = sE + %s * a%d_%d;' % (lit(mwf[k]), j, k))
            comp.append('        jt%d = -1.0 * (sT + dTdt * dcp%d) / (V * cptot);' % (j, j))
            comp.append('        je%d = T * RU * sE / (conp ? P : V) + (conp ? V : P) * jt%d * '
                        '(1.0 / T);' % (j, j))
            comp.append('    }')
            stores = []
            for k in rows:
                self._entry(k + 2, col, 'a%d_%d' % (j, k), stores, '    ')
            for k in range(Ns - 1):
                if k not in rows and (k + 2, col) in self.pattern:
                    self._entry(k + 2, col, '0.0', stores, '    ')
            self._entry(0, col, 'jt%d' % j, stores, '    ')
            self._entry(1, col, 'je%d' % j, stores, '    ')
            parts.append((loads, ['    PJ_OPAQUE(%s);' % x for x in pins], comp, stores))
        if parts:
            out += parts[0][0] + parts[0][1]
        for j, (loads, pins, comp, stores) in enumerate(parts):
            nxt = parts[j + 1] if j + 1 < len(parts) else None
            if nxt and self.early_loads:
                out += nxt[0]
            out += comp
            if nxt and not self.early_loads:
                out += nxt[0]
            out += stores
            if nxt:
                out += nxt[1]
        missing = self.pattern - set(self.entries)
        for row, col in sorted(missing):
            # stored entries that no sub-kernel writes stay zero
            self._entry(row, col, '0.0', out, '    ')
        # the full layouts: every entry outside the pattern is zero
        NN = self.NN
        out.append('    if (L >= 2) {')
        for row in range(NN):
            for col in range(NN):
                if (row, col) not in self.pattern:
                    addr = 'o + (long)(L == 2 ? %d : %d) * jI' % (row * NN + col, col * NN + row)
                    out.append(('        PJ_STO(%s, 0.0);' if self.nt_store else '        *(%s) = 0.0;')
                               % addr)
        out.append('    }')
        out.append('#undef LD')
        out.append('#undef LDN')
        out.append('#undef BS')
        out.append('}')
        out.append('#undef PJ_POS')
        out.append('#undef PJ_STO')
        return out

    def source(self, target='hip', kernels='ab'):
        A, c, Ns = self.A, self.c, self.Ns
        out = [TARGET_HIP % STRIDE if target == 'hip' else TARGET_HOST, lanegen.PRELUDE,
               JAC_HELPERS]
        if c['n_plog']:
            out.append(lanegen.PLOG_HELPER)
            out.append(PLOG_JAC)
            n = 4 * c['n_plog'] * c['plog_max_p']
            mp, npl = c['plog_max_p'], c['n_plog']
            vals = []
            for p in range(npl):
                for q in range(4):
                    for k in range(mp):
                        vals.append(A['plog_params'][(q * npl + p) * mp + k])
            out.append('PJ_CONST double plog_tab[%d] = {%s};' % (n, ', '.join(lit(v) for v in vals)))
        if c['n_cheb']:
            out.append(lanegen.CHEB_HELPER)
            out.append(CHEB_JAC)
            n = c['n_cheb'] * c['cheb_max_t'] * c['cheb_max_p']
            out.append('PJ_CONST double cheb_tab[%d] = {%s};'
                       % (n, ', '.join(lit(v) for v in A['cheb_params'][:n])))
        out.append('#define SB(v) scr[(long)(v) * PJ_SS + sl]')
        out += self._kernel_a()
        if 'b' in kernels:
            out += self._kernel_b()
        out.append('#undef SB')
        out.append('}  // namespace\n')
        ent = ', '.join('%d' % ((a << 16) | b) for a, b in self.entries)
        sig = '{0x%016xULL, %dULL, %dULL, %dULL, %dULL, %dULL, %dULL}' % (
            self.hash, LJAC_VERSION, Ns, self.Nr, self.nscr, len(self.entries), STRIDE)
        qual = 'extern "C" __device__ const' if target == 'hip' else 'extern "C" const'
        out.append('%s unsigned long long %s[7] = %s;' % (qual, SIG, sig))
        out.append('%s int %s[%d] = {%s};\n' % (qual, ENTRIES, len(self.entries), ent))
        if target == 'hip':
            out += ['// one state per lane, %d states per workgroup; state s0 + sl, scratch' % BLOCK,
                    '// column sl; the Gibbs terms of each lane in its column of the LDS array.',
                    '// Each kernel is a part of its own (PJ_PART 0: kernel A, 1 + L: the store',
                    '// kernel of layout L), so that the parts compile in parallel',
                    '#if !defined(PJ_PART) || PJ_PART == 0',
                    'extern "C" __global__ void __launch_bounds__(%d) %s(' % (BLOCK, KERNEL_A),
                    '    long n, long s0, const double* __restrict__ param, const double* __restrict__ phi,',
                    '    double* __restrict__ scr, long sN, long sI, long sS, int conp) {',
                    '    __shared__ double bsh[%d * %d];' % (Ns, BLOCK),
                    '    const long sl = (long)blockIdx.x * %d + threadIdx.x;' % BLOCK,
                    '    if (sl >= n) return;',
                    '    if (conp) ja_state<1>(s0 + sl, sl, param, phi, scr, bsh, threadIdx.x, sN, sI);',
                    '    else ja_state<0>(s0 + sl, sl, param, phi, scr, bsh, threadIdx.x, sN, sI);',
                    '}',
                    '#endif']
            for L in (range(4) if 'b' in kernels else ()):
                out += ['#if !defined(PJ_PART) || PJ_PART == %d' % (L + 1),
                        'extern "C" __global__ void __launch_bounds__(%d) %s%d(' % (BLOCK, KERNEL_B, L),
                        '    long n, long s0, const double* __restrict__ param, const double* __restrict__ phi,',
                        '    double* __restrict__ scr, double* __restrict__ jac, long sN, long sI,',
                        '    long sS, long jN, long jS, long jI, int conp) {',
                        '    __shared__ double bsh[%d * %d];' % (Ns + self._nU, BLOCK),
                        '    const long sl = (long)blockIdx.x * %d + threadIdx.x;' % BLOCK,
                        '    if (sl < n) jb_state<%d>(s0 + sl, sl, param, phi, scr, jac, bsh, threadIdx.x,'
                        % L,
                        '                              sN, sI, jN, jS, jI, conp);',
                        '}',
                        '#endif']
        else:
            out += ['extern "C" void pyjac_ljac_host(long n, const double* param, const double* phi,',
                    '    double* scr, double* jac, int layout, long sN, long sI, long jN, int conp) {',
                    '    double bsh[%d];' % (Ns + self._nU),
                    '    pj_ss = n;',
                    '    for (long s = 0; s < n; ++s) {',
                    '        if (conp) ja_state<1>(s, s, param, phi, scr, bsh, 0, sN, sI);',
                    '        else ja_state<0>(s, s, param, phi, scr, bsh, 0, sN, sI);',
                    '    }',
                    '    for (long s = 0; s < n; ++s) {',
                    '        switch (layout) {',
                    '            case 0: jb_state<0>(s, s, param, phi, scr, jac, bsh, 0, sN, sI, jN, 0, sI, conp); break;',
                    '            case 1: jb_state<1>(s, s, param, phi, scr, jac, bsh, 0, sN, sI, jN, 0, sI, conp); break;',
                    '            case 2: jb_state<2>(s, s, param, phi, scr, jac, bsh, 0, sN, sI, jN, 0, sI, conp); break;',
                    '            default: jb_state<3>(s, s, param, phi, scr, jac, bsh, 0, sN, sI, jN, 0, sI, conp);',
                    '        }',
                    '    }',
                    '}',
                    '// the store kernel of a layout with explicit output strides (entry p of',
                    '// state s at jac[(s - jS) * jN + p * jI]), as the runtime launches it to',
                    '// write C-order positions as lines of a chunk',
                    'extern "C" void pyjac_ljac_host_strided(long n, const double* param,',
                    '    const double* phi, double* scr, double* jac, int layout, long sN, long sI,',
                    '    long jN, long jS, long jI, int conp) {',
                    '    double bsh[%d];' % (Ns + self._nU),
                    '    pj_ss = n;',
                    '    for (long s = 0; s < n; ++s) {',
                    '        if (conp) ja_state<1>(s, s, param, phi, scr, bsh, 0, sN, sI);',
                    '        else ja_state<0>(s, s, param, phi, scr, bsh, 0, sN, sI);',
                    '    }',
                    '    for (long s = 0; s < n; ++s) {',
                    '        switch (layout) {',
                    '            case 0: jb_state<0>(s, s, param, phi, scr, jac, bsh, 0, sN, sI, jN, jS, jI, conp); break;',
                    '            case 1: jb_state<1>(s, s, param, phi, scr, jac, bsh, 0, sN, sI, jN, jS, jI, conp); break;',
                    '            case 2: jb_state<2>(s, s, param, phi, scr, jac, bsh, 0, sN, sI, jN, jS, jI, conp); break;',
                    '            default: jb_state<3>(s, s, param, phi, scr, jac, bsh, 0, sN, sI, jN, jS, jI, conp);',
                    '        }',
                    '    }',
                    '}']
        return '\n'.join(out) + '\n'


def generate(mt, target='hip'):
    """Source of the specialised Jacobian module of ``mt``."""
    return LaneJacSource(mt).source(target)


def entry_list(mt):
    """The (row, column) of every entry the module stores, in store order."""
    src = LaneJacSource(mt)
    src.source('host')
    return list(src.entries)


PARTS = 5   # kernel A, and the store kernel of each of the four layouts


def part_paths(out_path):
    """The code objects of a module: ``out_path`` holds kernel A, the
    siblings ``<base>.b<L>.hsaco`` the store kernel of layout L."""
    base = os.path.splitext(out_path)[0]
    return [out_path] + ['%s.b%d.hsaco' % (base, L) for L in range(4)]


def build(mt, out_path, verbose=False, jobs=PARTS):
    """Generate the Jacobian module of ``mt`` and compile its five parts
    (hipcc --genco for gfx950, in parallel) into ``part_paths(out_path)``;
    returns ``out_path``.  The source is written beside it (``.hip``)."""
    import concurrent.futures
    from . import hipcc, HIPCC_FLAGS
    if mt.Ns > MAX_SPECIES:
        raise ValueError('{} species: the specialised Jacobian keeps the concentrations and '
                         'Gibbs terms of a state in registers, which holds up to {} species'
                         .format(mt.Ns, MAX_SPECIES))
    if getattr(mt, 'jac_type', 0):
        raise ValueError('the specialised Jacobian module stores the exact Jacobian')
    src = generate(mt)
    base = os.path.splitext(out_path)[0]
    src_path = base + '.hip'
    os.makedirs(os.path.dirname(os.path.abspath(out_path)) or '.', exist_ok=True)
    with open(src_path, 'w') as f:
        f.write(src)
    flags = [x for x in HIPCC_FLAGS if x != '-fPIC']

    def one(part):
        path = part_paths(out_path)[part]
        tmp = path + '.tmp'
        cmd = [hipcc(), '--genco'] + flags + ['-DPJ_PART=%d' % part, '-o', tmp, src_path]
        res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if res.returncode != 0:
            raise RuntimeError('hipcc failed on %s (part %d):\n%s'
                               % (src_path, part, res.stdout[-4000:]))
        if verbose:
            print(res.stdout)
        return tmp, path
    with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        done = list(ex.map(one, range(PARTS)))
    for tmp, path in done:
        os.replace(tmp, path)
    return out_path


def module_path(mt, directory=None):
    """Where the Jacobian module of ``mt``'s tables is cached (it may not
    exist)."""
    return os.path.join(directory or lanegen.MODULE_DIR,
                        'ljac_%016x_v%d.hsaco' % (tables_hash(mt), LJAC_VERSION))


def ensure_module(mt, build_missing=True, directory=None):
    """Path of ``mt``'s Jacobian module, compiled first if missing and
    ``build_missing``; None if missing and not built (or not buildable)."""
    if mt.Ns > MAX_SPECIES or getattr(mt, 'jac_type', 0):
        return None
    path = module_path(mt, directory)
    if all(os.path.exists(p) for p in part_paths(path)):
        return path
    if not build_missing:
        return None
    return build(mt, path)


def build_bundled(jobs=1, names=lanegen.BUNDLED):
    """Compile the Jacobian modules of the bundled mechanisms (always
    rebuilds; each module's parts compile in parallel)."""
    import concurrent.futures
    from .._paths import mechanism_files
    from ..core.mech_tables import MechTables

    def one(name):
        mt = MechTables.from_file(*mechanism_files(name))
        return build(mt, module_path(mt))
    with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        return list(ex.map(one, names))
