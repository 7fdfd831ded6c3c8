// pyjac_oracle.cpp -- CPU restatement of the reference's species-rate and
// analytical-Jacobian arithmetic.  TEST INFRASTRUCTURE ONLY: this library is
// the checker the HIP kernels are compared against (tests, smoke check and
// the CPU-baseline leg of the benchmark).  It is never called by the
// product path.
//
// One state at a time, every sub-kernel as its own function, run in the
// reference's kernel order (core/rate_subs.py:3092-3200 for the species
// rates, core/create_jacobian.py:4960-5112 for the Jacobian).  Expressions
// keep the reference's evaluation order (as emitted in
// examples/opencl/{h2o2,ch4}/*.ocl and the loopy templates cited per
// function); the file is compiled with -ffp-contract=off so that no
// operation is fused.
//
// The species-rate chain is a template over the scalar type so that the
// same code, instantiated with a forward-mode dual number, yields an
// automatic-differentiation Jacobian (the reference's tests use Adept for
// this, tests/test_jacobian.py:132-265).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/pyjac_hip.h"

namespace {

const double RU = 8314.4621;              // chem_model.py:20
const double PA = 101325.0;               // chem_model.py:28
const double EXP_MAX = 690.775527898;     // utils.py:286
const double SMALL = 1e-300;              // utils.py:292
const double T_MIN = 100.0, T_MAX = 10000.0;  // instruction_creator.py:370
const double V_MIN = 1e-30;               // instruction_creator.py:382
const double LN10 = 2.302585092994046;    // np.log(10)
const double LN10SQ = LN10 * LN10;        // log(10) * log(10)
const double RU_INV = 1.0 / RU;

// ---------------------------------------------------------------- duals
template <int N>
struct Dual {
    double v;
    double d[N];
    Dual() : v(0) { std::fill(d, d + N, 0.0); }
    Dual(double x) : v(x) { std::fill(d, d + N, 0.0); }
};
template <int N> Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v + b.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <int N> Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v - b.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <int N> Dual<N> operator-(const Dual<N>& a) {
    Dual<N> r(-a.v);
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
}
template <int N> Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v * b.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <int N> Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v / b.v);
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
    return r;
}
template <int N> Dual<N> operator+(const Dual<N>& a, double b) { return a + Dual<N>(b); }
template <int N> Dual<N> operator+(double a, const Dual<N>& b) { return Dual<N>(a) + b; }
template <int N> Dual<N> operator-(const Dual<N>& a, double b) { return a - Dual<N>(b); }
template <int N> Dual<N> operator-(double a, const Dual<N>& b) { return Dual<N>(a) - b; }
template <int N> Dual<N> operator*(const Dual<N>& a, double b) {
    Dual<N> r(a.v * b);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
    return r;
}
template <int N> Dual<N> operator*(double a, const Dual<N>& b) { return b * a; }
template <int N> Dual<N> operator/(const Dual<N>& a, double b) { return a / Dual<N>(b); }
template <int N> Dual<N> operator/(double a, const Dual<N>& b) { return Dual<N>(a) / b; }
template <int N> Dual<N>& operator+=(Dual<N>& a, const Dual<N>& b) { a = a + b; return a; }
template <int N> Dual<N>& operator*=(Dual<N>& a, const Dual<N>& b) { a = a * b; return a; }
template <int N> bool operator<(const Dual<N>& a, const Dual<N>& b) { return a.v < b.v; }
template <int N> bool operator<(const Dual<N>& a, double b) { return a.v < b; }
template <int N> bool operator<=(const Dual<N>& a, double b) { return a.v <= b; }
template <int N> bool operator>(const Dual<N>& a, double b) { return a.v > b; }
template <int N> Dual<N> exp(const Dual<N>& a) {
    Dual<N> r(std::exp(a.v));
    for (int i = 0; i < N; ++i) r.d[i] = r.v * a.d[i];
    return r;
}
template <int N> Dual<N> log(const Dual<N>& a) {
    Dual<N> r(std::log(a.v));
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / a.v;
    return r;
}
template <int N> Dual<N> log10(const Dual<N>& a) {
    Dual<N> r(std::log10(a.v));
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / (a.v * LN10);
    return r;
}
template <int N> Dual<N> pow(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(std::pow(a.v, b.v));
    double la = a.v > 0 ? std::log(a.v) : 0.0;
    for (int i = 0; i < N; ++i)
        r.d[i] = r.v * (b.d[i] * la + (a.v != 0 ? b.v * a.d[i] / a.v : 0.0));
    return r;
}
template <int N> Dual<N> pow(const Dual<N>& a, double b) {
    Dual<N> r(std::pow(a.v, b));
    for (int i = 0; i < N; ++i)
        r.d[i] = (a.v != 0 ? b * std::pow(a.v, b - 1.0) : 0.0) * a.d[i];
    return r;
}
template <int N> Dual<N> fmin(const Dual<N>& a, const Dual<N>& b) { return a.v <= b.v ? a : b; }
template <int N> Dual<N> fmax(const Dual<N>& a, const Dual<N>& b) { return a.v >= b.v ? a : b; }
template <int N> Dual<N> fmin(double a, const Dual<N>& b) { return fmin(Dual<N>(a), b); }
template <int N> Dual<N> fmax(double a, const Dual<N>& b) { return fmax(Dual<N>(a), b); }

inline double value(double x) { return x; }
inline double value(long double x) { return (double)x; }
template <int N> double value(const Dual<N>& x) { return x.v; }

using std::exp;
using std::log;
using std::log10;
using std::pow;
using std::fmin;
using std::fmax;

// fast_powi: loopy_utils/preambles_and_manglers.py:123-174 (multiplication
// chain, negative powers through the reciprocal)
template <class R>
R fast_powi(R val, int p) {
    if (p < 0) {
        val = R(1.0) / val;
        p = -p;
    }
    switch (p) {
        case 0: return R(1.0);
        case 1: return val;
        case 2: return val * val;
        case 3: return val * val * val;
        case 4: return val * val * val * val;
        case 5: return val * val * val * val * val;
    }
    R r = val * val * val * val * val * val;
    for (int i = 6; i < p; ++i) r = r * val;
    return r;
}

template <class R>
R clampT(R T) { return fmin(T_MAX, fmax(T_MIN, T)); }

// Arrhenius forms (rate_subs.py:2803-2870)
template <class R>
R arrhenius(int form, double A, double b, double Ta, const R& Tval,
            const R& logT, const R& Tinv) {
    switch (form) {
        case PYJAC_ARR_A: return R(A);
        case PYJAC_ARR_AT_B: return R(A) * fast_powi(Tval, (int)b);
        case PYJAC_ARR_NO_TA: return exp(fmin(EXP_MAX, R(A) + logT * b));
        case PYJAC_ARR_NO_B: return exp(fmin(EXP_MAX, R(A) - Ta * Tinv));
        default: return exp(fmin(EXP_MAX, R(A) + logT * b + -1.0 * Ta * Tinv));
    }
}

// NASA-7 polynomials (rate_subs.py:3335-3354), coefficient set by T_mid
inline const double* nasa(const pyjac_mech_tables& m, int k, double T) {
    return (T < m.T_mid[k]) ? m.a_lo + 7 * k : m.a_hi + 7 * k;
}
template <class R> R eval_cp(const double* a, const R& T) {
    return RU * (T * (T * (T * (T * a[4] + a[3]) + a[2]) + a[1]) + a[0]);
}
template <class R> R eval_cv(const double* a, const R& T) {
    return RU * (T * (T * (T * (T * a[4] + a[3]) + a[2]) + a[1]) + a[0] - 1.0);
}
template <class R> R eval_h(const double* a, const R& T) {
    return RU * (T * (T * (T * (T * (T * a[4] / 5.0 + a[3] / 4.0) + a[2] / 3.0) +
                           a[1] / 2.0) + a[0]) + a[5]);
}
template <class R> R eval_u(const double* a, const R& T) {
    return RU * (T * (T * (T * (T * (T * a[4] / 5.0 + a[3] / 4.0) + a[2] / 3.0) +
                           a[1] / 2.0) + a[0]) - T + a[5]);
}
template <class R> R eval_b(const double* a, const R& T, const R& logT, const R& Tinv) {
    return T * (T * (T * (T * a[4] / 20.0 + a[3] / 12.0) + a[2] / 6.0) + a[1] / 2.0) +
           (a[0] + -1.0) * logT + -1.0 * a[0] + a[6] + -1.0 * a[5] * Tinv;
}
inline double eval_dcp(const double* a, double T) {
    return RU * (T * (T * (4.0 * T * a[4] + 3.0 * a[3]) + 2.0 * a[2]) + a[1]);
}
inline double eval_db(const double* a, double T, double Tinv) {
    return T * (T * (T * a[4] / 5.0 + a[3] / 4.0) + a[2] / 3.0) + a[1] / 2.0 +
           Tinv * (a[0] + -1.0 + a[5] * Tinv);
}

// Work arrays of one state (the reference's rwk, kernel_gen.py:46-62)
template <class R>
struct Work {
    int Ns, Nr;
    std::vector<R> conc, kf, kf_fall, thd_conc, Pr, Fi, Fcent, Atroe, Btroe,
        X_sri, b, Kc, kr, pres_mod, rop_fwd, rop_rev, rop_net, wdot, energy,
        spec_heat;
    std::vector<R> ppoly, tpoly;
    R T_raw, Tc, P, V;
    void init(const pyjac_mech_tables& m) {
        Ns = m.n_species;
        Nr = m.n_reactions;
        conc.assign(Ns, R(0.0));
        kf.assign(Nr, R(0.0));
        kf_fall.assign(std::max(1, m.n_fall), R(0.0));
        thd_conc.assign(std::max(1, m.n_thd), R(0.0));
        pres_mod.assign(std::max(1, m.n_thd), R(0.0));
        Pr.assign(std::max(1, m.n_fall), R(0.0));
        Fi.assign(std::max(1, m.n_fall), R(0.0));
        Fcent.assign(std::max(1, m.n_troe), R(0.0));
        Atroe.assign(std::max(1, m.n_troe), R(0.0));
        Btroe.assign(std::max(1, m.n_troe), R(0.0));
        X_sri.assign(std::max(1, m.n_sri), R(0.0));
        b.assign(Ns, R(0.0));
        Kc.assign(std::max(1, m.n_rev), R(0.0));
        kr.assign(std::max(1, m.n_rev), R(0.0));
        rop_fwd.assign(Nr, R(0.0));
        rop_rev.assign(std::max(1, m.n_rev), R(0.0));
        rop_net.assign(Nr, R(0.0));
        wdot.assign(Ns, R(0.0));
        energy.assign(Ns, R(0.0));
        spec_heat.assign(Ns, R(0.0));
        int pm = std::max(2, std::max(m.cheb_max_p, m.cheb_max_t) + 1);
        ppoly.assign(pm, R(0.0));
        tpoly.assign(pm, R(0.0));
    }
};

// ---------------------------------------------------------------- rates
// get_concentrations, rate_subs.py:597-706
template <class R>
void get_concentrations(const pyjac_mech_tables& m, const R* phi, Work<R>& w) {
    const int Ns = m.n_species;
    R T_val = clampT(w.T_raw);
    R V_inv = 1.0 / fmax(V_MIN, w.V);
    R n_sum(0.0);
    R cns = w.P / (RU * T_val);
    for (int i = 0; i < Ns - 1; ++i) {
        R n = fmax(SMALL, phi[2 + i]);
        n_sum = n_sum + n;
        w.conc[i] = n * V_inv;
    }
    w.conc[Ns - 1] = cns + -1.0 * fmax(SMALL, n_sum) * V_inv;
}

// get_simple_arrhenius_rates, rate_subs.py:2642-3023 (and falloff=True)
template <class R>
void simple_rates(const pyjac_mech_tables& m, Work<R>& w) {
    R Tval = clampT(w.T_raw);
    R logT = log(Tval);
    R Tinv = 1.0 / Tval;
    for (int r = 0; r < m.n_reactions; ++r) {
        int s = m.simple_mask[r];
        if (s < 0) continue;
        w.kf[r] = arrhenius(m.simple_form[s], m.simple_A[s], m.simple_b[s],
                            m.simple_Ta[s], Tval, logT, Tinv);
    }
    for (int f = 0; f < m.n_fall; ++f)
        w.kf_fall[f] = arrhenius(m.fall_form[f], m.fall_A[f], m.fall_b[f],
                                 m.fall_Ta[f], Tval, logT, Tinv);
}

// get_plog_arrhenius_rates, rate_subs.py:2084-2248
template <class R>
void plog_rates(const pyjac_mech_tables& m, Work<R>& w) {
    if (!m.n_plog) return;
    R Tc = clampT(w.T_raw);
    R Tinv = 1.0 / Tc;
    R logT = log(Tc);
    R logP = log(fmax(SMALL, w.P));
    const int mp = m.plog_max_p, np_ = m.n_plog;
    for (int p = 0; p < np_; ++p) {
        const double* P0 = m.plog_params + (0 * np_ + p) * mp;
        int numP = m.plog_np[p] - 1;
        bool lower = logP <= P0[0];
        bool upper = logP > P0[numP];
        int lo = 0, hi = 0;
        if (lower) { lo = 0; hi = 0; }
        if (upper) { lo = numP; hi = numP; }
        for (int k = 0; k < mp - 1; ++k)
            if (k < numP && logP > P0[k] && logP <= P0[k + 1]) { lo = k; hi = k + 1; }
        double low[4], high[4];
        for (int q = 0; q < 4; ++q) {
            low[q] = m.plog_params[(q * np_ + p) * mp + lo];
            high[q] = m.plog_params[(q * np_ + p) * mp + hi];
        }
        R logk1 = low[1] + logT * low[2] - low[3] * Tinv;
        R logk2 = high[1] + logT * high[2] - high[3] * Tinv;
        R kf_temp = logk1;
        if (!(lower || upper))
            kf_temp = (-logk1 + logk2) * (logP - low[0]) / (high[0] - low[0]) + kf_temp;
        w.kf[m.plog_map[p]] = exp(fmin(EXP_MAX, kf_temp));
    }
}

// get_cheb_arrhenius_rates, rate_subs.py:1904-2081 (C form of exp10)
template <class R>
void cheb_rates(const pyjac_mech_tables& m, Work<R>& w) {
    if (!m.n_cheb) return;
    R Tinv = 1.0 / clampT(w.T_raw);
    R logP = log(fmax(SMALL, w.P));
    const int maxT = m.cheb_max_t, maxP = m.cheb_max_p;
    const int poly_max = std::max(maxP, maxT);
    for (int c = 0; c < m.n_cheb; ++c) {
        double Pmin = m.cheb_plim[2 * c], Pmax = m.cheb_plim[2 * c + 1];
        double Tmin = m.cheb_tlim[2 * c], Tmax = m.cheb_tlim[2 * c + 1];
        R Tred = (-Tmax - Tmin + 2.0 * Tinv) / (Tmax - Tmin);
        R Pred = (-Pmax - Pmin + 2.0 * logP) / (Pmax - Pmin);
        int numP = m.cheb_np[c], numT = m.cheb_nt[c];
        w.ppoly[0] = R(1.0); w.ppoly[1] = Pred;
        w.tpoly[0] = R(1.0); w.tpoly[1] = Tred;
        for (int p = 2; p < poly_max; ++p) {
            if (p < numP) w.ppoly[p] = 2.0 * Pred * w.ppoly[p - 1] - w.ppoly[p - 2];
            if (p < numT) w.tpoly[p] = 2.0 * Tred * w.tpoly[p - 1] - w.tpoly[p - 2];
        }
        const double* par = m.cheb_params + (size_t)c * maxT * maxP;
        R kf_temp(0.0);
        for (int t = 0; t < maxT; ++t) {
            R temp(0.0);
            for (int k = 0; k < maxP; ++k)
                if (k < numP) temp = temp + w.ppoly[k] * par[t * maxP + k];
            if (t < numT) kf_temp = kf_temp + w.tpoly[t] * temp;
        }
        w.kf[m.cheb_map[c]] = exp(LN10 * fmin(EXP_MAX, kf_temp));
    }
}

// get_thd_body_concs, rate_subs.py:1788-1901 (unclamped T)
template <class R>
void thd_body_concs(const pyjac_mech_tables& m, Work<R>& w) {
    for (int t = 0; t < m.n_thd; ++t) {
        int not_spec = m.thd_type[t] != 2;
        R thd = w.P * (double)not_spec / (RU * w.T_raw);
        for (int s = m.thd_off[t]; s < m.thd_off[t + 1]; ++s)
            thd = thd + (m.thd_eff[s] + -1.0 * not_spec) * w.conc[m.thd_spec[s]];
        w.thd_conc[t] = thd;
    }
}

// get_reduced_pressure_kernel, rate_subs.py:2251-2347
template <class R>
void reduced_pressure(const pyjac_mech_tables& m, Work<R>& w) {
    for (int f = 0; f < m.n_fall; ++f) {
        int r = m.fall_map[f];
        R kinf = m.fall_type[f] ? w.kf_fall[f] : w.kf[r];
        R k0 = m.fall_type[f] ? w.kf[r] : w.kf_fall[f];
        w.Pr[f] = w.thd_conc[m.fall_to_thd[f]] * k0 / kinf;
    }
}

// get_lind_kernel / get_troe_kernel / get_sri_kernel, rate_subs.py:2350-2639
template <class R>
void falloff_blend(const pyjac_mech_tables& m, Work<R>& w) {
    R T = clampT(w.T_raw);
    R Tinv = 1.0 / T;
    for (int f = 0; f < m.n_fall; ++f)
        if (m.fall_blend[f] == 1) w.Fi[f] = R(1.0);
    for (int f = 0; f < m.n_fall; ++f) {
        if (m.fall_blend[f] != 2) continue;
        int i = m.fall_param[f];
        R logPr = log10(fmax(SMALL, w.Pr[f]));
        R Fc = m.troe_a[i] * exp(fmin(EXP_MAX, -1.0 * T * m.troe_T1inv[i])) +
               (1.0 + -1.0 * m.troe_a[i]) * exp(fmin(EXP_MAX, -1.0 * T * m.troe_T3inv[i]));
        if (m.troe_T2[i] != 0.0)
            Fc = Fc + exp(fmin(EXP_MAX, -1.0 * m.troe_T2[i] / T));
        w.Fcent[i] = Fc;
        R logFc = log10(fmax(SMALL, Fc));
        R B = -1.1762 * logFc + -1.0 * 0.14 * logPr + 0.806;
        w.Btroe[i] = B;
        R A = -0.67 * logFc + logPr + -0.4;
        w.Atroe[i] = A;
        w.Fi[f] = pow(fmax(SMALL, Fc), 1.0 / (A * A / (B * B) + 1.0));
    }
    for (int f = 0; f < m.n_fall; ++f) {
        if (m.fall_blend[f] != 3) continue;
        int i = m.fall_param[f];
        R logPr = log10(fmax(SMALL, w.Pr[f]));
        R X = 1.0 / (logPr * logPr + 1.0);
        R base = m.sri_a[i] * exp(fmin(EXP_MAX, -m.sri_b[i] * Tinv)) +
                 exp(fmin(EXP_MAX, -T / m.sri_c[i]));
        R Fi = pow(base, X);
        if (m.sri_d[i] != 1.0) Fi = Fi * m.sri_d[i];
        if (m.sri_e[i] != 0.0) Fi = Fi * pow(T, m.sri_e[i]);
        w.Fi[f] = Fi;
        w.X_sri[i] = X;
    }
}

// eval_b + get_rev_rates, rate_subs.py:1627-1785 and :3281-3388
template <class R>
void rev_rates(const pyjac_mech_tables& m, Work<R>& w) {
    if (!m.n_rev) return;
    R T = clampT(w.T_raw);
    R logT = log(T);
    R Tinv = 1.0 / T;
    for (int k = 0; k < m.n_species; ++k)
        w.b[k] = eval_b(nasa(m, k, value(T)), T, logT, Tinv);
    for (int r = 0; r < m.n_reactions; ++r) {
        int ri = m.rev_mask[r];
        if (ri < 0) continue;
        double P_val = !(m.nu_sum[r] > 0) ? RU / PA : PA / RU;
        double P_sum = fast_powi(P_val, std::abs(m.nu_sum[r]));
        R B_sum(0.0);
        for (int s = m.rxn_off[r]; s < m.rxn_off[r + 1]; ++s) {
            int net_nu = m.nu_prod[s] + -1 * m.nu_reac[s];
            if (net_nu != 0) B_sum = B_sum + (double)net_nu * w.b[m.rxn_spec[s]];
        }
        B_sum = exp(fmin(EXP_MAX, B_sum));
        R Kc = P_sum * B_sum;
        w.kr[ri] = w.kf[r] / Kc;
        w.Kc[ri] = Kc;
    }
}

// get_rxn_pres_mod, rate_subs.py:1510-1624
template <class R>
void pres_mod(const pyjac_mech_tables& m, Work<R>& w) {
    for (int t = 0; t < m.n_thd; ++t) {
        if (m.fall_mask[m.thd_map[t]] >= 0) continue;
        w.pres_mod[t] = w.thd_conc[t];
    }
    for (int f = 0; f < m.n_fall; ++f) {
        R ci = w.Fi[f] / (1.0 + w.Pr[f]);
        if (!m.fall_type[f]) ci = ci * w.Pr[f];
        w.pres_mod[m.fall_to_thd[f]] = ci;
    }
}

// get_rop, get_rop_net, get_spec_rates: rate_subs.py:1079-1507
template <class R>
void rops(const pyjac_mech_tables& m, Work<R>& w) {
    for (int r = 0; r < m.n_reactions; ++r) {
        R rf = w.kf[r];
        for (int s = m.rxn_off[r]; s < m.rxn_off[r + 1]; ++s)
            rf = rf * fast_powi(w.conc[m.rxn_spec[s]], m.nu_reac[s]);
        w.rop_fwd[r] = rf;
    }
    for (int r = 0; r < m.n_reactions; ++r) {
        int ri = m.rev_mask[r];
        if (ri < 0) continue;
        R rr = w.kr[ri];
        for (int s = m.rxn_off[r]; s < m.rxn_off[r + 1]; ++s)
            rr = rr * fast_powi(w.conc[m.rxn_spec[s]], m.nu_prod[s]);
        w.rop_rev[ri] = rr;
    }
    for (int r = 0; r < m.n_reactions; ++r) {
        R net = w.rop_fwd[r];
        int ri = m.rev_mask[r];
        if (ri >= 0) net = net + -1.0 * w.rop_rev[ri];
        int t = m.thd_mask[r];
        if (t >= 0) net = net * w.pres_mod[t];
        w.rop_net[r] = net;
    }
    for (int k = 0; k < m.n_species; ++k) w.wdot[k] = R(0.0);
    for (int r = 0; r < m.n_reactions; ++r)
        for (int s = m.rxn_off[r]; s < m.rxn_off[r + 1]; ++s) {
            int nu = m.nu_prod[s] + -1 * m.nu_reac[s];
            int k = m.rxn_spec[s];
            w.wdot[k] = w.wdot[k] + (double)nu * w.rop_net[r];
        }
}

// molar rates, temperature rate, extra-variable rate:
// rate_subs.py:709-1076
template <class R>
void state_rates(const pyjac_mech_tables& m, int conp, Work<R>& w, R* dphi) {
    const int Ns = m.n_species;
    for (int i = 0; i < Ns + 1; ++i) dphi[i] = R(0.0);
    R V_val = fmax(V_MIN, w.V);
    for (int k = 0; k < Ns - 1; ++k) dphi[2 + k] = V_val * w.wdot[k];
    R T = clampT(w.T_raw);
    for (int k = 0; k < Ns; ++k) {
        const double* a = nasa(m, k, value(T));
        if (conp) {
            w.energy[k] = eval_h(a, T);
            w.spec_heat[k] = eval_cp(a, T);
        } else {
            w.energy[k] = eval_u(a, T);
            w.spec_heat[k] = eval_cv(a, T);
        }
    }
    R upper(0.0), lower(0.0);
    for (int k = 0; k < Ns; ++k) {
        lower = lower + w.conc[k] * w.spec_heat[k];
        upper = upper + w.energy[k] * w.wdot[k];
    }
    dphi[0] = dphi[0] + -1.0 * upper / lower;
    R dE(0.0);
    if (conp) {
        dphi[1] = V_val * dphi[0] / T;
        for (int k = 0; k < Ns - 1; ++k) dE = dE + (1.0 + -1.0 * m.mw_post[k]) * w.wdot[k];
        dphi[1] = dphi[1] + V_val * dE * T * RU / w.P;
    } else {
        dphi[1] = w.P * dphi[0] / T;
        for (int k = 0; k < Ns - 1; ++k) dE = dE + (1.0 + -1.0 * m.mw_post[k]) * w.wdot[k];
        dphi[1] = dphi[1] + T * RU * dE;
    }
}

template <class R>
void species_rates_state(const pyjac_mech_tables& m, int conp, R param,
                         const R* phi, R* dphi, Work<R>& w) {
    w.T_raw = phi[0];
    if (conp) { w.P = param; w.V = phi[1]; }
    else { w.P = phi[1]; w.V = param; }
    get_concentrations(m, phi, w);
    simple_rates(m, w);
    plog_rates(m, w);
    cheb_rates(m, w);
    thd_body_concs(m, w);
    reduced_pressure(m, w);
    falloff_blend(m, w);
    rev_rates(m, w);
    pres_mod(m, w);
    rops(m, w);
    state_rates(m, conp, w, dphi);
}

// ---------------------------------------------------------------- Jacobian
struct JacCtx {
    const pyjac_mech_tables& m;
    int conp, Ns, NN, ns;
    double* J;  // dense, row-major [row][col]
    Work<double>& w;
    double T, Tinv, P, V, dTdt;
    std::vector<double> cp_like, energy, dspec, db;
    double spec_heat_tot;
    double& at(int row, int col) { return J[row * NN + col]; }
    // stored entries of a sparse pattern (null: every entry): the T and
    // extra-variable rows sum only the stored species entries, as the
    // reference's conditional Jacobian reads do (jac_create)
    const unsigned char* mask = nullptr;
    bool stored(int row, int col) const { return !mask || mask[row * NN + col]; }
};

inline double conc_g(const JacCtx& c, int k) { return fmax(SMALL, c.w.conc[k]); }

// dRopidnj / dRopidnj_ns: create_jacobian.py:4169-4520
void dRopi_dnj(JacCtx& c) {
    const pyjac_mech_tables& m = c.m;
    for (int r = 0; r < m.n_reactions; ++r) {
        int ri = m.rev_mask[r], ti = m.thd_mask[r];
        double kr_i = ri >= 0 ? c.w.kr[ri] : 0.0;
        double kf_i = c.w.kf[r];
        double ci = ti >= 0 ? c.w.pres_mod[ti] : 1.0;
        for (int sk = m.rxn_off[r]; sk < m.rxn_off[r + 1]; ++sk) {
            int k = m.rxn_spec[sk];
            if (k == c.ns) continue;
            int nu_k = m.nu_prod[sk] + -1 * m.nu_reac[sk];
            for (int sj = m.rxn_off[r]; sj < m.rxn_off[r + 1]; ++sj) {
                int j = m.rxn_spec[sj];
                if (j == c.ns) continue;
                double Sr = m.nu_prod[sj], Sf = m.nu_reac[sj];
                for (int si = m.rxn_off[r]; si < m.rxn_off[r + 1]; ++si) {
                    int nr = m.nu_prod[si], nf = m.nu_reac[si];
                    int sp = m.rxn_spec[si];
                    if (sp == j) { nr = nr + -1; nf = nf + -1; }
                    Sr = Sr * fast_powi(conc_g(c, sp), nr);
                    Sf = Sf * fast_powi(conc_g(c, sp), nf);
                }
                c.at(k + 2, j + 2) += (kf_i * Sf + -1.0 * kr_i * Sr) * ci * nu_k;
            }
        }
    }
    for (int h = 0; h < m.n_rxn_has_ns; ++h) {
        int r = m.rxn_has_ns[h];
        int ri = m.rev_mask[r], ti = m.thd_mask[r];
        double kr_i = ri >= 0 ? c.w.kr[ri] : 0.0;
        double kf_i = c.w.kf[r];
        double ci = ti >= 0 ? c.w.pres_mod[ti] : 1.0;
        for (int sk = m.rxn_off[r]; sk < m.rxn_off[r + 1]; ++sk) {
            int k = m.rxn_spec[sk];
            if (k == c.ns) continue;
            int nu_k = m.nu_prod[sk] + -1 * m.nu_reac[sk];
            double Sr = 1.0, Sf = 1.0;
            for (int si = m.rxn_off[r]; si < m.rxn_off[r + 1]; ++si) {
                int nr = m.nu_prod[si], nf = m.nu_reac[si];
                int sp = m.rxn_spec[si];
                if (sp == c.ns) { Sr = Sr * nr; nr = nr + -1; }
                Sr = Sr * fast_powi(conc_g(c, sp), nr);
                if (sp == c.ns) { Sf = Sf * nf; nf = nf + -1; }
                Sf = Sf * fast_powi(conc_g(c, sp), nf);
            }
            double upd = (kr_i * Sr + -1.0 * kf_i * Sf) * ci * nu_k;
            for (int j = 0; j < c.Ns - 1; ++j) c.at(k + 2, j + 2) += upd;
        }
    }
}

inline double rop_diff(const JacCtx& c, int r) {
    double ropi = c.w.rop_fwd[r];
    int ri = c.m.rev_mask[r];
    if (ri >= 0) ropi = ropi + -1.0 * c.w.rop_rev[ri];
    return ropi;
}

// falloff dFi/dPr factor folded with the dPr/dn term, create_jacobian.py
// :3780-3890 (Lindemann, SRI, Troe)
double fall_dnj_factor(JacCtx& c, int f) {
    const pyjac_mech_tables& m = c.m;
    int r = m.fall_map[f];
    double Pr = c.w.Pr[f];
    double dFi = 0.0;
    if (m.fall_blend[f] == 2) {
        int i = m.fall_param[f];
        double A = c.w.Atroe[i], B = c.w.Btroe[i];
        dFi = A * A + B * B;
        dFi = -2.0 * A * B * (0.14 * A + B) * log(fmax(SMALL, c.w.Fcent[i])) /
              (fmax(SMALL, Pr) * dFi * dFi * LN10);
    } else if (m.fall_blend[f] == 3) {
        int i = m.fall_param[f];
        double X = c.w.X_sri[i];
        double T = c.T;
        double lv = log(fmax(SMALL, m.sri_a[i] * exp(fmin(EXP_MAX, -m.sri_b[i] / T)) +
                                         exp(fmin(EXP_MAX, -T / m.sri_c[i]))));
        dFi = -2.0 * X * X * lv * log(fmax(SMALL, Pr)) / (fmax(SMALL, Pr) * LN10SQ);
    }
    double Fi_fac = dFi;
    double k0, kinf;
    if (m.fall_type[f]) {
        k0 = c.w.kf[r];
        kinf = c.w.kf_fall[f];
    } else {
        kinf = c.w.kf[r];
        k0 = c.w.kf_fall[f];
        Fi_fac = Pr * Fi_fac + 1.0;
    }
    return k0 * (c.w.Fi[f] * Fi_fac + -1.0 * c.w.pres_mod[m.fall_to_thd[f]]) /
           (kinf * (Pr + 1.0));
}

inline bool thd_has_ns(const pyjac_mech_tables& m, int t, int ns) {
    int e = m.thd_off[t + 1];
    return e > m.thd_off[t] && m.thd_spec[e - 1] == ns;
}

// dci_thd_dnj / dci_{lind,sri,troe}_dnj (+ _ns): create_jacobian.py
// :3558-4166.  blend 0 = third-body only.
void dci_dnj(JacCtx& c, int blend) {
    const pyjac_mech_tables& m = c.m;
    std::vector<int> items;  // thd index (blend 0) or falloff index
    if (blend == 0) {
        for (int t = 0; t < m.n_thd; ++t)
            if (m.fall_mask[m.thd_map[t]] < 0) items.push_back(t);
    } else {
        for (int f = 0; f < m.n_fall; ++f)
            if (m.fall_blend[f] == blend) items.push_back(f);
    }
    for (int pass = 0; pass < 2; ++pass) {
        for (int it : items) {
            int t = blend == 0 ? it : m.fall_to_thd[it];
            int r = m.thd_map[t];
            if (pass == 1 && !thd_has_ns(m, t, c.ns)) continue;
            double ropi = rop_diff(c, r);
            double fac = blend == 0 ? 1.0 : fall_dnj_factor(c, it);
            if (pass == 0) {
                for (int sj = m.thd_off[t]; sj < m.thd_off[t + 1]; ++sj) {
                    int j = m.thd_spec[sj];
                    if (j == c.ns) continue;
                    double dci = 0.0;
                    if (m.thd_type[t] == 2) dci = 1.0;
                    if (m.thd_type[t] == 1) dci = m.thd_eff[sj] + -1.0;
                    for (int sk = m.rxn_off[r]; sk < m.rxn_off[r + 1]; ++sk) {
                        int k = m.rxn_spec[sk];
                        if (k == c.ns) continue;
                        int nu_k = m.nu_prod[sk] + -1 * m.nu_reac[sk];
                        if (blend == 0) c.at(k + 2, j + 2) += nu_k * dci * ropi;
                        else c.at(k + 2, j + 2) += nu_k * dci * ropi * fac;
                    }
                }
            } else {
                double ns_eff = m.thd_eff[m.thd_off[t + 1] - 1];
                double dci = 0.0;
                if (m.thd_type[t] == 2) dci = -1.0;
                if (m.thd_type[t] == 1) dci = 1.0 + -1.0 * ns_eff;
                for (int sk = m.rxn_off[r]; sk < m.rxn_off[r + 1]; ++sk) {
                    int k = m.rxn_spec[sk];
                    int nu_k = m.nu_prod[sk] + -1 * m.nu_reac[sk];
                    if (k == c.ns) continue;
                    for (int j = 0; j < c.Ns - 1; ++j) {
                        if (blend == 0) c.at(k + 2, j + 2) += nu_k * dci * ropi;
                        else c.at(k + 2, j + 2) += nu_k * dci * ropi * fac;
                    }
                }
            }
        }
        if (blend != 0 && pass == 0) {
            // the _ns kernel of this falloff form follows its plain kernel
        }
    }
}

// total_specific_energy + dcp/dcv + db: create_jacobian.py:3234-3555
void thermo_derivs(JacCtx& c) {
    const pyjac_mech_tables& m = c.m;
    double spec_tot = 0.0;
    for (int k = 0; k < c.Ns; ++k) spec_tot = spec_tot + c.w.spec_heat[k] * c.w.conc[k];
    c.spec_heat_tot = 0.0 + spec_tot;
    double T = clampT(c.T);
    double Tinv = 1.0 / T;
    c.dspec.assign(c.Ns, 0.0);
    c.db.assign(c.Ns, 0.0);
    for (int k = 0; k < c.Ns; ++k) {
        const double* a = nasa(m, k, T);
        c.dspec[k] = eval_dcp(a, T);
        if (m.n_rev) c.db[k] = eval_db(a, T, Tinv);
    }
}

// dTdot_dnj and dEdot_dnj: create_jacobian.py:3266-3484
void rows_TE_dnj(JacCtx& c) {
    const pyjac_mech_tables& m = c.m;
    const int ns = c.ns;
    const double fixed = c.conp ? c.P : c.V;
    const double extra = c.conp ? c.V : c.P;
    for (int j = 0; j < c.Ns - 1; ++j) {
        double sum = 0.0;
        for (int k = 0; k < c.Ns - 1; ++k)
            if (c.stored(k + 2, j + 2))
                sum = sum + (c.w.energy[k] + -1.0 * c.w.energy[ns] * m.mw_post[k]) *
                                c.at(k + 2, j + 2);
        c.at(0, j + 2) = -1.0 * (sum + c.dTdt * (c.w.spec_heat[j] + -1.0 * c.w.spec_heat[ns])) /
                         (c.V * c.spec_heat_tot);
    }
    for (int j = 0; j < c.Ns - 1; ++j) {
        double sum = 0.0;
        for (int k = 0; k < c.Ns - 1; ++k)
            if (c.stored(k + 2, j + 2)) sum = sum + (1.0 + -1.0 * m.mw_post[k]) * c.at(k + 2, j + 2);
        c.at(1, j + 2) = c.at(1, j + 2) + c.T * RU * sum / fixed + extra * c.at(0, j + 2) * (1.0 / c.T);
    }
}

inline void scatter_col(JacCtx& c, int r, int col, double val) {
    const pyjac_mech_tables& m = c.m;
    for (int s = m.rxn_off[r]; s < m.rxn_off[r + 1]; ++s)
        if (m.rxn_spec[s] != c.ns)
            c.at(m.rxn_spec[s] + 2, col) += (m.nu_prod[s] + -1.0 * m.nu_reac[s]) * val;
}

// d(rop)/dT contribution for one reaction given dkf = dln kf/dT
// (create_jacobian.py:2978-2992)
double dropi_dT_core(JacCtx& c, int r, double dkf) {
    const pyjac_mech_tables& m = c.m;
    double d = c.w.rop_fwd[r] * dkf;
    double ci = 1.0;
    int ri = m.rev_mask[r];
    if (ri >= 0) {
        double dB = 0.0;
        for (int s = m.rxn_off[r]; s < m.rxn_off[r + 1]; ++s)
            dB = dB + (m.nu_prod[s] + -1.0 * m.nu_reac[s]) * c.db[m.rxn_spec[s]];
        d = d + -1.0 * c.w.rop_rev[ri] * (dkf + -1.0 * dB);
    }
    int ti = m.thd_mask[r];
    if (ti >= 0) ci = c.w.pres_mod[ti];
    return d * ci * c.V;
}

// S'_ns / S''_ns products (create_jacobian.py:3027-3060)
void ns_products(JacCtx& c, int r, double& Sf, double& Sr) {
    const pyjac_mech_tables& m = c.m;
    int last = m.rxn_off[r + 1] - 1;
    Sf = m.nu_reac[last];
    Sr = m.nu_prod[last];
    for (int s = m.rxn_off[r]; s < m.rxn_off[r + 1]; ++s) {
        int nf = m.nu_reac[s], nr = m.nu_prod[s];
        int sp = m.rxn_spec[s];
        if (sp == c.ns) { nf = nf - 1; nr = nr - 1; }
        Sf = Sf * fast_powi(conc_g(c, sp), nf);
        Sr = Sr * fast_powi(conc_g(c, sp), nr);
    }
}

// PLOG interval search shared by the dT / dE kernels
// (create_jacobian.py:2823-2850)
void plog_interval(const pyjac_mech_tables& m, int p, double logP, int& lo, int& hi, int& numP) {
    const int mp = m.plog_max_p, np_ = m.n_plog;
    const double* P0 = m.plog_params + (0 * np_ + p) * mp;
    numP = m.plog_np[p] - 1;
    lo = 0;
    hi = numP;
    for (int k = 0; k < mp - 1; ++k)
        if (k <= numP && logP > P0[k] && logP <= P0[k + 1]) { lo = k; hi = k + 1; }
}
inline double plog_par(const pyjac_mech_tables& m, int q, int p, int k) {
    return m.plog_params[(q * m.n_plog + p) * m.plog_max_p + k];
}

// temperature column: dRopidT (+_ns), PLOG, Chebyshev, dci_*_dT,
// dTdotdT, dEdotdT (create_jacobian.py:1847-3230)
void column_T(JacCtx& c) {
    const pyjac_mech_tables& m = c.m;
    const double T = c.T, Tinv = 1.0 / T;
    for (int r = 0; r < m.n_reactions; ++r) {
        int s = m.simple_mask[r];
        if (s < 0) continue;
        double dkf = (m.simple_b[s] + m.simple_Ta[s] * Tinv) * Tinv;
        scatter_col(c, r, 0, dropi_dT_core(c, r, dkf));
    }
    for (int h = 0; h < m.n_rxn_has_ns; ++h) {
        int r = m.rxn_has_ns[h];
        int ri = m.rev_mask[r], ti = m.thd_mask[r];
        double kr_i = ri >= 0 ? c.w.kr[ri] : 0.0;
        double ci = ti >= 0 ? c.w.pres_mod[ti] : 1.0;
        double Sf, Sr;
        ns_products(c, r, Sf, Sr);
        double d = (Sr * kr_i + -1.0 * Sf * c.w.kf[r]) * c.V * ci * c.P / (RU * T * T);
        scatter_col(c, r, 0, d);
    }
    double logP = log(fmax(SMALL, c.P));
    for (int p = 0; p < m.n_plog; ++p) {
        int lo, hi, numP;
        plog_interval(m, p, logP, lo, hi, numP);
        double dkf;
        if (logP > plog_par(m, 0, p, numP))
            dkf = (plog_par(m, 2, p, hi) + plog_par(m, 3, p, hi) * Tinv) * Tinv;
        else
            dkf = (plog_par(m, 2, p, lo) + plog_par(m, 3, p, lo) * Tinv) * Tinv;
        if (logP > plog_par(m, 0, p, 0) && logP <= plog_par(m, 0, p, numP))
            dkf = dkf + Tinv * (logP - plog_par(m, 0, p, lo)) *
                            (plog_par(m, 2, p, hi) - plog_par(m, 2, p, lo) +
                             (plog_par(m, 3, p, hi) - plog_par(m, 3, p, lo)) * Tinv) /
                            (plog_par(m, 0, p, hi) - plog_par(m, 0, p, lo));
        int r = m.plog_map[p];
        scatter_col(c, r, 0, dropi_dT_core(c, r, dkf));
    }
    if (m.n_cheb) {
        const int maxT = m.cheb_max_t, maxP = m.cheb_max_p;
        const int poly_max = std::max(maxP, maxT - 1);
        std::vector<double> pp(std::max(2, poly_max + 1)), tp(std::max(2, poly_max + 1));
        for (int q = 0; q < m.n_cheb; ++q) {
            double Pmin = m.cheb_plim[2 * q], Pmax = m.cheb_plim[2 * q + 1];
            double Tmin = m.cheb_tlim[2 * q], Tmax = m.cheb_tlim[2 * q + 1];
            int numP = m.cheb_np[q], numT = m.cheb_nt[q] - 1;
            double Tred = (2 * Tinv - Tmax - Tmin) / (Tmax - Tmin);
            double Pred = (2 * logP - Pmax - Pmin) / (Pmax - Pmin);
            pp[0] = 1; pp[1] = Pred; tp[0] = 1; tp[1] = 2 * Tred;
            for (int p = 2; p < poly_max; ++p) {
                if (p < numP) pp[p] = 2 * Pred * pp[p - 1] - pp[p - 2];
                if (p < numT) tp[p] = 2 * Tred * tp[p - 1] - tp[p - 2];
            }
            const double* par = m.cheb_params + (size_t)q * maxT * maxP;
            double dkf = 0;
            for (int t = 0; t < maxT - 1; ++t) {
                double temp = 0;
                for (int k = 0; k < maxP; ++k)
                    if (k < numP) temp = temp + pp[k] * par[(t + 1) * maxP + k];
                if (t < numT) dkf = dkf + (t + 1) * tp[t] * temp;
            }
            dkf = -dkf * 2 * LN10 * Tinv * Tinv / (Tmax - Tmin);
            int r = m.cheb_map[q];
            scatter_col(c, r, 0, dropi_dT_core(c, r, dkf));
        }
    }
    // dci_thd_dT and dci_{lind,sri,troe}_dT (create_jacobian.py:2117-2588)
    for (int blend = 0; blend <= 3; ++blend) {
        int bl = blend == 0 ? 0 : (blend == 1 ? 1 : (blend == 2 ? 3 : 2));
        int ncount = bl == 0 ? m.n_thd : m.n_fall;
        for (int it = 0; it < ncount; ++it) {
            int t, f = -1;
            if (bl == 0) {
                t = it;
                if (m.fall_mask[m.thd_map[t]] >= 0) continue;
            } else {
                f = it;
                if (m.fall_blend[f] != bl) continue;
                t = m.fall_to_thd[f];
            }
            int r = m.thd_map[t];
            int e = m.thd_off[t + 1] - 1;
            double mod = 1.0;
            if (m.thd_type[t] == 2) mod = m.thd_spec[e] == c.ns;
            if (m.thd_type[t] == 1 && m.thd_spec[e] == c.ns) mod = m.thd_eff[e];
            double rop_net = rop_diff(c, r);
            if (bl == 0) {
                double fac = -1.0 * c.P * mod * RU_INV * Tinv * Tinv * c.V * rop_net;
                scatter_col(c, r, 0, fac);
                continue;
            }
            double thd_fac = -1.0 * c.P * mod * RU_INV * Tinv * Tinv;
            int s = m.simple_mask[r];
            double kf_0, kf_inf, beta_0, beta_inf, Ta_0, Ta_inf;
            if (m.fall_type[f]) {
                kf_0 = c.w.kf[r]; beta_0 = m.simple_b[s]; Ta_0 = m.simple_Ta[s];
                kf_inf = c.w.kf_fall[f]; beta_inf = m.fall_b[f]; Ta_inf = m.fall_Ta[f];
            } else {
                kf_0 = c.w.kf_fall[f]; beta_0 = m.fall_b[f]; Ta_0 = m.fall_Ta[f];
                kf_inf = c.w.kf[r]; beta_inf = m.simple_b[s]; Ta_inf = m.simple_Ta[s];
            }
            double pmod = c.w.pres_mod[t];
            double Pr = c.w.Pr[f];
            double theta_Pr = Tinv * (beta_0 - beta_inf + (Ta_0 - Ta_inf) * Tinv);
            double theta_no_Pr = thd_fac * kf_0 / kf_inf;
            double dFi = 0.0;
            if (bl == 2) {
                int i = m.fall_param[f];
                double a = m.troe_a[i], T1 = m.troe_T1inv[i], T2 = m.troe_T2[i], T3 = m.troe_T3inv[i];
                double A = c.w.Atroe[i], B = c.w.Btroe[i], Fc = c.w.Fcent[i];
                double dFcent = -1.0 * a * T1 * exp(fmin(EXP_MAX, -1.0 * T * T1)) +
                                (a + -1.0) * T3 * exp(fmin(EXP_MAX, -1.0 * T * T3)) +
                                T2 * Tinv * Tinv * exp(fmin(EXP_MAX, -1.0 * T2 * Tinv));
                double logFcent = log(fmax(SMALL, Fc));
                double absq = A * A + B * B;
                double absqsq = absq * absq;
                dFi = -1.0 * B *
                      (2.0 * A * Fc * (0.14 * A + B) * (Pr * theta_Pr + theta_no_Pr) * logFcent +
                       Pr * dFcent * (2.0 * A * (1.1762 * A + -1.0 * 0.67 * B) * logFcent +
                                      -1.0 * B * absq * LN10)) /
                      (Fc * fmax(SMALL, Pr) * absqsq * LN10);
            } else if (bl == 3) {
                int i = m.fall_param[f];
                double a = m.sri_a[i], b = m.sri_b[i], cc = m.sri_c[i], e_ = m.sri_e[i];
                double X = c.w.X_sri[i];
                double cinv = 1.0 / cc;
                double ec = exp(fmin(EXP_MAX, -T * cinv));
                double eb = exp(fmin(EXP_MAX, -b * Tinv));
                double lv = log(fmax(SMALL, a * eb + ec));
                dFi = -X * (ec * cinv - a * b * Tinv * Tinv * eb) / (a * eb + ec) + e_ * Tinv -
                      2 * X * X * lv * (Pr * theta_Pr + theta_no_Pr) * log(fmax(SMALL, Pr)) /
                          (fmax(SMALL, Pr) * LN10SQ);
            }
            double d = pmod * (-1.0 * (Pr * theta_Pr + theta_no_Pr) / (Pr + 1.0) + dFi);
            if (!m.fall_type[f])
                d = d + theta_Pr * pmod + c.w.Fi[f] * theta_no_Pr / (Pr + 1.0);
            d = d * c.V * rop_net;
            scatter_col(c, r, 0, d);
        }
    }
    // dTdot_dT (create_jacobian.py:1847-1998)
    const int ns = c.ns;
    double rate_sum = 0.0;
    double dTsum = (c.w.spec_heat[ns] * Tinv + -1.0 * c.dspec[ns]) * c.w.conc[ns];
    double Vinv = 1.0 / c.V;
    for (int i = 0; i < c.Ns - 1; ++i) {
        rate_sum = rate_sum + Vinv * c.at(i + 2, 0) *
                                  (-1.0 * c.w.energy[i] + c.w.energy[ns] * m.mw_post[i]);
        rate_sum = rate_sum + c.w.wdot[i] * (-1.0 * c.w.spec_heat[i] + m.mw_post[i] * c.w.spec_heat[ns]);
        dTsum = dTsum + (c.w.spec_heat[ns] * Tinv + -1.0 * c.dspec[i]) * c.w.conc[i];
    }
    c.at(0, 0) = c.at(0, 0) + (c.dTdt * dTsum + rate_sum) / c.spec_heat_tot;
    // dEdotdT (create_jacobian.py:2001-2114)
    double sum = 0.0;
    if (c.conp) {
        for (int i = 0; i < c.Ns - 1; ++i)
            sum = sum + (1.0 + -1.0 * m.mw_post[i]) * (Vinv * c.at(i + 2, 0) + Tinv * c.w.wdot[i]);
        c.at(1, 0) = c.at(1, 0) + c.V * Tinv * (c.at(0, 0) + -1.0 * Tinv * c.dTdt);
        c.at(1, 0) = c.at(1, 0) + RU * T * c.V * sum / c.P;
    } else {
        double fac = T / c.V;
        for (int i = 0; i < c.Ns - 1; ++i)
            sum = sum + (1.0 - m.mw_post[i]) * (c.at(i + 2, 0) * fac + c.w.wdot[i]);
        c.at(1, 0) = c.at(1, 0) + c.P * (c.at(0, 0) - c.dTdt * Tinv) * Tinv;
        c.at(1, 0) = c.at(1, 0) + RU * sum;
    }
}

// extra-variable column: dRopidE (+_ns), PLOG, Chebyshev, dci_*_dE,
// dTdotdE, dEdotdE (create_jacobian.py:362-1844)
void column_E(JacCtx& c) {
    const pyjac_mech_tables& m = c.m;
    const double T = c.T, Tinv = 1.0 / T;
    const int ns = c.ns;
    if (c.conp) {
        // simple then PLOG then Chebyshev, CONP form: d(V w)/dV
        auto conp_term = [&](int r, bool with_pmod) {
            int nu_f = -1, nu_r = -1;
            for (int s = m.rxn_off[r]; s < m.rxn_off[r + 1]; ++s) {
                nu_f = nu_f + m.nu_reac[s];
                nu_r = nu_r + m.nu_prod[s];
            }
            double d = -1.0 * nu_f * c.w.rop_fwd[r];
            int ri = m.rev_mask[r];
            if (ri >= 0) d = d + nu_r * c.w.rop_rev[ri];
            int ti = m.thd_mask[r];
            if (with_pmod && ti >= 0) d = d * c.w.pres_mod[ti];
            scatter_col(c, r, 1, d);
        };
        for (int r = 0; r < m.n_reactions; ++r)
            if (m.simple_mask[r] >= 0) conp_term(r, true);
        for (int h = 0; h < m.n_rxn_has_ns; ++h) {
            int r = m.rxn_has_ns[h];
            int ri = m.rev_mask[r], ti = m.thd_mask[r];
            double kr_i = ri >= 0 ? c.w.kr[ri] : 0.0;
            double ci = ti >= 0 ? c.w.pres_mod[ti] : 1.0;
            double Sf, Sr;
            ns_products(c, r, Sf, Sr);
            double fac = c.P / (RU * T);
            scatter_col(c, r, 1, (Sf * c.w.kf[r] - Sr * kr_i) * ci * fac);
        }
        for (int p = 0; p < m.n_plog; ++p) conp_term(m.plog_map[p], false);
        for (int q = 0; q < m.n_cheb; ++q) conp_term(m.cheb_map[q], false);
    } else {
        for (int h = 0; h < m.n_rxn_has_ns; ++h) {
            int r = m.rxn_has_ns[h];
            int ri = m.rev_mask[r], ti = m.thd_mask[r];
            double kr_i = ri >= 0 ? c.w.kr[ri] : 0.0;
            double ci = ti >= 0 ? c.w.pres_mod[ti] : 1.0;
            double Sf, Sr;
            ns_products(c, r, Sf, Sr);
            double fac = c.V / (RU * T);
            scatter_col(c, r, 1, (Sf * c.w.kf[r] - Sr * kr_i) * ci * fac);
        }
        double logP = log(fmax(SMALL, c.P));
        double logT = log(fmax(SMALL, T));
        for (int p = 0; p < m.n_plog; ++p) {
            int lo, hi, numP;
            plog_interval(m, p, logP, lo, hi, numP);
            double dkf = 0;
            if (logP > plog_par(m, 0, p, 0) && logP <= plog_par(m, 0, p, numP))
                dkf = (plog_par(m, 1, p, hi) - plog_par(m, 1, p, lo) +
                       logT * (plog_par(m, 2, p, hi) - plog_par(m, 2, p, lo)) -
                       (plog_par(m, 3, p, hi) - plog_par(m, 3, p, lo)) * Tinv) /
                      (c.P * (plog_par(m, 0, p, hi) - plog_par(m, 0, p, lo)));
            int r = m.plog_map[p];
            double d = c.w.rop_fwd[r];
            int ri = m.rev_mask[r];
            if (ri >= 0) d = d - c.w.rop_rev[ri];
            d = d * dkf * c.V;
            scatter_col(c, r, 1, d);
        }
        if (m.n_cheb) {
            const int maxT = m.cheb_max_t, maxP = m.cheb_max_p;
            const int poly_max = std::max(maxP - 1, maxT);
            std::vector<double> pp(std::max(2, poly_max + 1)), tp(std::max(2, poly_max + 1));
            for (int q = 0; q < m.n_cheb; ++q) {
                double Pmin = m.cheb_plim[2 * q], Pmax = m.cheb_plim[2 * q + 1];
                double Tmin = m.cheb_tlim[2 * q], Tmax = m.cheb_tlim[2 * q + 1];
                int numP = m.cheb_np[q] - 1, numT = m.cheb_nt[q];
                double Tred = (2 * Tinv - Tmax - Tmin) / (Tmax - Tmin);
                double Pred = (2 * logP - Pmax - Pmin) / (Pmax - Pmin);
                pp[0] = 1; pp[1] = 2 * Pred; tp[0] = 1; tp[1] = Tred;
                for (int p = 2; p < poly_max; ++p) {
                    if (p < numP) pp[p] = 2 * Pred * pp[p - 1] - pp[p - 2];
                    if (p < numT) tp[p] = 2 * Tred * tp[p - 1] - tp[p - 2];
                }
                const double* par = m.cheb_params + (size_t)q * maxT * maxP;
                double dkf = 0;
                for (int t = 0; t < maxT; ++t) {
                    double temp = 0;
                    for (int k = 0; k < maxP - 1; ++k)
                        if (k < numP) temp = temp + (k + 1) * pp[k] * par[t * maxP + k + 1];
                    if (t < numT) dkf = dkf + tp[t] * temp;
                }
                dkf = dkf * 2 * LN10 / (c.P * (Pmax - Pmin));
                int r = m.cheb_map[q];
                double d = c.w.rop_fwd[r];
                int ri = m.rev_mask[r];
                if (ri >= 0) d = d - c.w.rop_rev[ri];
                d = d * dkf * c.V;
                scatter_col(c, r, 1, d);
            }
        }
    }
    // dci_thd_dE and dci_{lind,sri,troe}_dE (create_jacobian.py:362-882)
    const double rt_inv = 1.0 / (RU * T);
    for (int blend = 0; blend <= 3; ++blend) {
        int bl = blend == 0 ? 0 : (blend == 1 ? 1 : (blend == 2 ? 3 : 2));
        int ncount = bl == 0 ? m.n_thd : m.n_fall;
        for (int it = 0; it < ncount; ++it) {
            int t, f = -1;
            if (bl == 0) {
                t = it;
                if (m.fall_mask[m.thd_map[t]] >= 0) continue;
            } else {
                f = it;
                if (m.fall_blend[f] != bl) continue;
                t = m.fall_to_thd[f];
            }
            int r = m.thd_map[t];
            int e = m.thd_off[t + 1] - 1;
            int ttype = m.thd_type[t];
            double rop_net = rop_diff(c, r);
            double mod = c.conp ? (double)(ttype == 1) : 1.0;
            if (ttype == 1 && m.thd_spec[e] == ns) mod = m.thd_eff[e];
            if (ttype == 2) mod = m.thd_spec[e] == ns;
            if (bl == 0) {
                if (c.conp) {
                    if (ttype != 3) mod = mod * c.P * rt_inv - c.w.pres_mod[t];
                    scatter_col(c, r, 1, mod * rop_net);
                } else {
                    scatter_col(c, r, 1, mod * rt_inv * c.V * rop_net);
                }
                continue;
            }
            int not_unity = ttype != 3;
            double kf_0, kf_inf;
            if (m.fall_type[f]) { kf_0 = c.w.kf[r]; kf_inf = c.w.kf_fall[f]; }
            else { kf_0 = c.w.kf_fall[f]; kf_inf = c.w.kf[r]; }
            if (c.conp) mod = mod * c.P * rt_inv * kf_0 / kf_inf;
            else mod = mod * rt_inv * kf_0 / kf_inf;
            double Pr = c.w.Pr[f];
            double theta = c.conp ? (mod + -1.0 * Pr * not_unity) : mod;
            double dFi = 0.0;
            if (bl == 2) {
                int i = m.fall_param[f];
                double A = c.w.Atroe[i], B = c.w.Btroe[i];
                double absqsq = A * A + B * B;
                absqsq = absqsq * absqsq;
                dFi = -2.0 * A * B * log(fmax(SMALL, c.w.Fcent[i])) * (0.14 * A + B) * theta /
                      (fmax(SMALL, Pr) * absqsq * LN10);
            } else if (bl == 3) {
                int i = m.fall_param[f];
                double a = m.sri_a[i], b = m.sri_b[i], cinv = 1.0 / m.sri_c[i];
                double X = c.w.X_sri[i];
                double inner;
                if (c.conp) {
                    double exp_cinv = exp(fmin(EXP_MAX, T * cinv));
                    double exp_b = exp(fmin(EXP_MAX, b * Tinv));
                    double exp_bc = exp(fmin(EXP_MAX, -b * Tinv - T * cinv));
                    inner = log(fmax(SMALL, (a * exp_cinv + exp_b) * exp_bc));
                } else {
                    double exp_b = exp(fmin(EXP_MAX, -b * Tinv));
                    double exp_cinv = exp(fmin(EXP_MAX, -T * cinv));
                    inner = log(fmax(SMALL, a * exp_b + exp_cinv));
                }
                dFi = -2 * X * X * theta * inner * log(fmax(SMALL, Pr)) / (fmax(SMALL, Pr) * LN10SQ);
            }
            double pm = c.w.pres_mod[t];
            double num = c.conp ? (-mod + Pr * not_unity) : -mod;
            double d = pm * (num / (Pr + 1) + dFi);
            if (!m.fall_type[f]) {
                if (c.conp) d = d + c.w.Fi[f] * mod / (Pr + 1) + -1.0 * not_unity * pm;
                else d = d + c.w.Fi[f] * mod / (Pr + 1);
            }
            d = c.conp ? d * rop_net : d * c.V * rop_net;
            scatter_col(c, r, 1, d);
        }
    }
    // dTdotdE (create_jacobian.py:1595-1737)
    if (c.conp) {
        double dTsum = 0.0, specsum = 0.0;
        double spec_inv = 1.0 / (c.spec_heat_tot * c.V);
        for (int i = 0; i < c.Ns - 1; ++i) {
            dTsum = dTsum + (c.w.spec_heat[i] + -1.0 * c.w.spec_heat[ns]) * c.w.conc[i];
            specsum = specsum + (c.w.energy[i] + -1.0 * c.w.energy[ns] * m.mw_post[i]) *
                                    (c.at(i + 2, 1) + -1.0 * c.w.wdot[i]);
        }
        c.at(0, 1) = c.at(0, 1) + (c.dTdt * dTsum + -1.0 * specsum) * spec_inv;
    } else {
        double sum = 0.0;
        double Vinv = 1.0 / c.V;
        for (int i = 0; i < c.Ns - 1; ++i)
            sum = sum + (c.w.energy[i] - c.w.energy[ns] * m.mw_post[i]) * c.at(i + 2, 1);
        double spec_inv = 1.0 / c.spec_heat_tot;
        c.at(0, 1) = c.at(0, 1) - (c.dTdt * c.w.spec_heat[ns] / (RU * T)) * spec_inv;
        c.at(0, 1) = c.at(0, 1) - (sum * Vinv) * spec_inv;
    }
    // dEdotdE (create_jacobian.py:1740-1844)
    {
        double sum = 0.0;
        for (int i = 0; i < c.Ns - 1; ++i) sum = sum + (1 - m.mw_post[i]) * c.at(i + 2, 1);
        double var = c.conp ? c.V : c.P;
        double par = c.conp ? c.P : c.V;
        c.at(1, 1) = c.at(1, 1) + (var * c.at(0, 1) + c.dTdt) / T;
        c.at(1, 1) = c.at(1, 1) + RU * T * sum / par;
    }
}

void jacobian_state(const pyjac_mech_tables& m, int conp, double param,
                    const double* phi, double* jac, Work<double>& w,
                    std::vector<double>& dphi, const unsigned char* mask = nullptr) {
    const int Ns = m.n_species, NN = Ns + 1;
    species_rates_state<double>(m, conp, param, phi, dphi.data(), w);
    std::fill(jac, jac + (size_t)NN * NN, 0.0);
    JacCtx c{m, conp, Ns, NN, Ns - 1, jac, w};
    c.mask = mask;
    c.T = phi[0];
    c.P = conp ? param : phi[1];
    c.V = conp ? phi[1] : param;
    c.dTdt = dphi[0];
    dRopi_dnj(c);
    dci_dnj(c, 0);
    dci_dnj(c, 1);
    dci_dnj(c, 3);
    dci_dnj(c, 2);
    thermo_derivs(c);
    rows_TE_dnj(c);
    column_T(c);
    column_E(c);
}

template <class F>
void parallel_states(long n, int nthreads, F&& f) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
        auto st = f.make_state();
#pragma omp for schedule(static)
        for (long s = 0; s < n; ++s) f(s, st);
    }
#else
    auto st = f.make_state();
    for (long s = 0; s < n; ++s) f(s, st);
#endif
}

}  // namespace

extern "C" {

// dphi[n][NN] from phi[n][NN], param[n] (C order per state)
int oracle_species_rates(const pyjac_mech_tables* m, int conp, long n,
                         const double* param, const double* phi, double* dphi,
                         int nthreads) {
    if (!m || !param || !phi || !dphi) return PYJAC_ERR_ARG;
    const int NN = m->n_species + 1;
    struct Fn {
        const pyjac_mech_tables* m; int conp, NN;
        const double *param, *phi; double* dphi;
        Work<double> make_state() const { Work<double> w; w.init(*m); return w; }
        void operator()(long s, Work<double>& w) const {
            species_rates_state<double>(*m, conp, param[s], phi + s * NN, dphi + s * NN, w);
        }
    } fn{m, conp, NN, param, phi, dphi};
    parallel_states(n, nthreads, fn);
    return PYJAC_OK;
}

// rates plus the validation outputs of the reference's for_validation
// build (kernel_argument_ordering, utils.py:252-256): rop_fwd[n][Nr],
// rop_rev[n][Nr] (0 for irreversible), pres_mod[n][Nr] (1 without third
// body), rop_net[n][Nr]
int oracle_species_rates_full(const pyjac_mech_tables* m, int conp, long n,
                              const double* param, const double* phi,
                              double* dphi, double* rop_fwd, double* rop_rev,
                              double* pmod, double* rop_net) {
    if (!m) return PYJAC_ERR_ARG;
    const int NN = m->n_species + 1, Nr = m->n_reactions;
    struct Fn {
        const pyjac_mech_tables* m; int conp, NN, Nr;
        const double *param, *phi; double *dphi, *rop_fwd, *rop_rev, *pmod, *rop_net;
        Work<double> make_state() const { Work<double> w; w.init(*m); return w; }
        void operator()(long s, Work<double>& w) const {
            species_rates_state<double>(*m, conp, param[s], phi + s * NN, dphi + s * NN, w);
            for (int r = 0; r < Nr; ++r) {
                rop_fwd[s * Nr + r] = w.rop_fwd[r];
                int ri = m->rev_mask[r], ti = m->thd_mask[r];
                rop_rev[s * Nr + r] = ri >= 0 ? w.rop_rev[ri] : 0.0;
                pmod[s * Nr + r] = ti >= 0 ? w.pres_mod[ti] : 1.0;
                rop_net[s * Nr + r] = w.rop_net[r];
            }
        }
    } fn{m, conp, NN, Nr, param, phi, dphi, rop_fwd, rop_rev, pmod, rop_net};
    parallel_states(n, 0, fn);
    return PYJAC_OK;
}

// The same species-rate expressions evaluated in x87 extended precision
// (long double, 64-bit significand: 11 bits beyond binary64), inputs and
// tables as given in binary64.  The result is returned as an unevaluated
// sum hi + lo of two doubles.  A reference of higher precision for the
// parity tests: how far a binary64 evaluation (this checker's, or the
// GPU's) lies from the exact value of the same formulas.
int oracle_species_rates_ld(const pyjac_mech_tables* m, int conp, long n, const double* param,
                            const double* phi, double* dphi_hi, double* dphi_lo,
                            int nthreads) {
    if (!m || !param || !phi || !dphi_hi || !dphi_lo) return PYJAC_ERR_ARG;
    const int NN = m->n_species + 1;
    struct St { Work<long double> w; std::vector<long double> x, y; };
    struct Fn {
        const pyjac_mech_tables* m; int conp, NN;
        const double *param, *phi; double *hi, *lo;
        St make_state() const {
            St st; st.w.init(*m); st.x.assign(NN, 0.0L); st.y.assign(NN, 0.0L); return st;
        }
        void operator()(long s, St& st) const {
            for (int i = 0; i < NN; ++i) st.x[i] = phi[s * NN + i];
            species_rates_state<long double>(*m, conp, (long double)param[s], st.x.data(),
                                             st.y.data(), st.w);
            for (int i = 0; i < NN; ++i) {
                const double h = (double)st.y[i];
                hi[s * NN + i] = h;
                lo[s * NN + i] = (double)(st.y[i] - (long double)h);
            }
        }
    } fn{m, conp, NN, param, phi, dphi_hi, dphi_lo};
    parallel_states(n, nthreads, fn);
    return PYJAC_OK;
}

// chem_utils (rate_subs.py:3391-3452): energy (h|u), heat capacity (cp|cv)
// and b, each [n][Ns]
int oracle_chem_utils(const pyjac_mech_tables* m, int conp, long n,
                      const double* phi, double* energy, double* heat,
                      double* bout) {
    if (!m) return PYJAC_ERR_ARG;
    const int Ns = m->n_species, NN = Ns + 1;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < n; ++s) {
        double T = clampT(phi[s * NN]);
        double logT = log(T), Tinv = 1.0 / T;
        for (int k = 0; k < Ns; ++k) {
            const double* a = nasa(*m, k, T);
            energy[s * Ns + k] = conp ? eval_h(a, T) : eval_u(a, T);
            heat[s * Ns + k] = conp ? eval_cp(a, T) : eval_cv(a, T);
            bout[s * Ns + k] = eval_b(a, T, logT, Tinv);
        }
    }
    return PYJAC_OK;
}

// analytical Jacobian, dense jac[n][NN][NN] (row, col)
int oracle_jacobian(const pyjac_mech_tables* m, int conp, long n,
                    const double* param, const double* phi, double* jac,
                    int nthreads) {
    if (!m || !param || !phi || !jac) return PYJAC_ERR_ARG;
    const int NN = m->n_species + 1;
    struct St { Work<double> w; std::vector<double> dphi; };
    struct Fn {
        const pyjac_mech_tables* m; int conp, NN;
        const double *param, *phi; double* jac;
        St make_state() const {
            St st; st.w.init(*m); st.dphi.assign(NN, 0.0); return st;
        }
        void operator()(long s, St& st) const {
            jacobian_state(*m, conp, param[s], phi + s * NN, jac + s * (size_t)NN * NN, st.w, st.dphi);
        }
    } fn{m, conp, NN, param, phi, jac};
    parallel_states(n, nthreads, fn);
    return PYJAC_OK;
}

// analytical Jacobian on a sparse pattern (mask[NN][NN], 1 = stored): the
// approximate Jacobian's pattern (determine_jac_inds, create_jacobian.py:187)
// keeps the exact values of the stored entries, and the T and extra-variable
// rows sum only those
int oracle_jacobian_masked(const pyjac_mech_tables* m, int conp, long n, const double* param,
                           const double* phi, double* jac, const unsigned char* mask) {
    if (!m || !param || !phi || !jac || !mask) return PYJAC_ERR_ARG;
    const int NN = m->n_species + 1;
    Work<double> w;
    w.init(*m);
    std::vector<double> dphi(NN, 0.0);
    for (long s = 0; s < n; ++s)
        jacobian_state(*m, conp, param[s], phi + s * NN, jac + s * (size_t)NN * NN, w, dphi, mask);
    return PYJAC_OK;
}

// forward-mode AD Jacobian of the species rates, dense jac[n][NN][NN]
int oracle_jacobian_ad(const pyjac_mech_tables* m, int conp, long n,
                       const double* param, const double* phi, double* jac) {
    if (!m) return PYJAC_ERR_ARG;
    const int NN = m->n_species + 1;
    const int CH = 16;
    typedef Dual<CH> D;
    Work<D> w;
    w.init(*m);
    std::vector<D> x(NN), y(NN);
    for (long s = 0; s < n; ++s) {
        for (int c0 = 0; c0 < NN; c0 += CH) {
            for (int i = 0; i < NN; ++i) {
                x[i] = D(phi[s * NN + i]);
                if (i >= c0 && i < c0 + CH) x[i].d[i - c0] = 1.0;
            }
            species_rates_state<D>(*m, conp, D(param[s]), x.data(), y.data(), w);
            for (int row = 0; row < NN; ++row)
                for (int q = 0; q < CH && c0 + q < NN; ++q)
                    jac[s * (size_t)NN * NN + row * NN + c0 + q] = y[row].d[q];
        }
    }
    return PYJAC_OK;
}

// finite-difference Jacobian (create_jacobian.py:4523-4896), dense
// jac[n][NN][NN]; mode 0 forward, 1 central, 2 backward.  The coefficients
// are the reference's tables (:4596-4617) with the textbook signs of forward
// orders 3 and 4 (the reference has +3/2 and +1/4 there, which makes the
// coefficients not sum to zero); the steps are CVODE's difference-quotient
// increments with inverse error weights w_i = 1 / (rtol |y_i| + atol), which
// the reference's formula follows with w_i = rtol |y_i| + atol (its own test
// marks that Jacobian broken, tests/test_jacobian.py:2683)
static bool fd_table(int mode, int order, std::vector<int>& x, std::vector<double>& y) {
    static const double fwd[7][7] = {
        {0},
        {-1, 1},
        {-3.0 / 2, 2, -1.0 / 2},
        {-11.0 / 6, 3, -3.0 / 2, 1.0 / 3},
        {-25.0 / 12, 4, -3, 4.0 / 3, -1.0 / 4},
        {-137.0 / 60, 5, -5, 10.0 / 3, -5.0 / 4, 1.0 / 5},
        {-49.0 / 20, 6, -15.0 / 2, 20.0 / 3, -15.0 / 4, 6.0 / 5, -1.0 / 6}};
    static const double cen[5][8] = {
        {0},
        {-0.5, 0.5},
        {1.0 / 12, -2.0 / 3, 2.0 / 3, -1.0 / 12},
        {-1.0 / 60, 3.0 / 20, -3.0 / 4, 3.0 / 4, -3.0 / 20, 1.0 / 60},
        {1.0 / 280, -4.0 / 105, 1.0 / 5, -4.0 / 5, 4.0 / 5, -1.0 / 5, 4.0 / 105, -1.0 / 280}};
    x.clear();
    y.clear();
    if (mode == 1) {
        if (order < 2 || order > 8 || order % 2) return false;
        for (int k = -order / 2; k <= order / 2; ++k)
            if (k) x.push_back(k);
        for (int k = 0; k < order; ++k) y.push_back(cen[order / 2][k]);
        return true;
    }
    if ((mode != 0 && mode != 2) || order < 1 || order > 6) return false;
    const double sgn = mode == 2 ? -1.0 : 1.0;
    for (int k = 0; k <= order; ++k) {
        x.push_back(mode == 2 ? -k : k);
        y.push_back(sgn * fwd[order][k]);
    }
    return true;
}

int oracle_fd_jacobian(const pyjac_mech_tables* m, int conp, long n, const double* param,
                       const double* phi, double* jac, int mode, int order, double rtol,
                       double atol) {
    if (!m) return PYJAC_ERR_ARG;
    std::vector<int> xk;
    std::vector<double> yk;
    if (!fd_table(mode, order, xk, yk)) return PYJAC_ERR_ARG;
    const int NN = m->n_species + 1;
    const double eps = DBL_EPSILON;
    Work<double> w;
    w.init(*m);
    std::vector<double> x(NN), f0(NN), f(NN), ewt(NN);
    for (long s = 0; s < n; ++s) {
        const double* p = phi + s * NN;
        double* J = jac + s * (size_t)NN * NN;
        for (int i = 0; i < NN; ++i) x[i] = p[i];
        species_rates_state<double>(*m, conp, param[s], x.data(), f0.data(), w);
        double sum = 0.0;
        for (int i = 0; i < NN; ++i) {
            ewt[i] = 1.0 / (atol + rtol * fabs(x[i]));
            sum = sum + (ewt[i] * fabs(f0[i])) * (ewt[i] * fabs(f0[i]));
        }
        const double fac = sqrt(sum / (double)NN);
        const double r0 = 1000.0 * rtol * eps * (double)NN * fac;
        const double srur = sqrt(eps);
        for (int i = 0; i < NN * NN; ++i) J[i] = 0.0;
        for (int i = 0; i < NN; ++i) {
            const double orig = x[i];
            const double r = fmax(srur * fabs(orig), r0 / ewt[i]);
            for (size_t k = 0; k < xk.size(); ++k) {
                x[i] = orig + xk[k] * r;
                species_rates_state<double>(*m, conp, param[s], x.data(), f.data(), w);
                for (int j = 0; j < NN; ++j) J[j * NN + i] = J[j * NN + i] + yk[k] * f[j];
            }
            x[i] = orig;
            for (int j = 0; j < NN; ++j) J[j * NN + i] = J[j * NN + i] / r;
        }
    }
    return PYJAC_OK;
}

// ROS2 integration of the species-rate ODE (the scheme of
// include/pyjac_integrator.h: two-stage L-stable Rosenbrock, gamma =
// 1 + 1/sqrt(2), Verwer et al. 1999), per state, with the analytical
// Jacobian and a dense LU with partial pivoting.  phi [n][NN] in place;
// stats[4]: accepted, rejected, largest attempts of a state, failed states.
static void lu_factor(std::vector<double>& A, std::vector<int>& piv, int N) {
    for (int k = 0; k < N; ++k) {
        int p = k;
        double best = -1.0;
        for (int i = k; i < N; ++i)
            if (fabs(A[i * N + k]) > best) { best = fabs(A[i * N + k]); p = i; }
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < N; ++j) std::swap(A[k * N + j], A[p * N + j]);
        for (int i = k + 1; i < N; ++i) A[i * N + k] /= A[k * N + k];
        for (int i = k + 1; i < N; ++i)
            for (int j = k + 1; j < N; ++j) A[i * N + j] -= A[i * N + k] * A[k * N + j];
    }
}

static void lu_solve(const std::vector<double>& A, const std::vector<int>& piv, int N,
                     std::vector<double>& b) {
    for (int k = 0; k < N; ++k)
        if (piv[k] != k) std::swap(b[k], b[piv[k]]);
    for (int k = 0; k < N; ++k)
        for (int i = k + 1; i < N; ++i) b[i] -= A[i * N + k] * b[k];
    for (int k = N - 1; k >= 0; --k) {
        b[k] /= A[k * N + k];
        for (int i = 0; i < k; ++i) b[i] -= A[i * N + k] * b[k];
    }
}

int oracle_integrate_ros2(const pyjac_mech_tables* m, int conp, long n, const double* param,
                          double* phi, double t_end, double rtol, double atol, double h0,
                          int fixed, long max_steps, long* stats) {
    if (!m || !param || !phi || !stats) return PYJAC_ERR_ARG;
    const int N = m->n_species + 1;
    const double gamma = 1.7071067811865475244;
    Work<double> w;
    w.init(*m);
    std::vector<double> y(N), f0(N), f1(N), k1(N), k2(N), y1(N), yn(N), J((size_t)N * N),
        A((size_t)N * N), dphi(N);
    std::vector<int> piv(N);
    for (int i = 0; i < 4; ++i) stats[i] = 0;
    for (long s = 0; s < n; ++s) {
        double* ys = phi + s * N;
        for (int i = 0; i < N; ++i) y[i] = ys[i];
        double t = 0.0, h = h0 > 0.0 ? fmin(h0, t_end) : 0.0;
        bool rejlast = false, done = !(t_end > 0.0);
        long att = 0;
        while (!done && att < max_steps) {
            species_rates_state<double>(*m, conp, param[s], y.data(), f0.data(), w);
            jacobian_state(*m, conp, param[s], y.data(), J.data(), w, dphi);
            if (h <= 0.0) {
                double r = 0.0;
                for (int i = 0; i < N; ++i) r = fmax(r, fabs(f0[i]) / (atol + rtol * fabs(y[i])));
                h = r > 0.0 ? 1e-2 / r : t_end;
                h = fmin(h, t_end - t);
            }
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < N; ++j)
                    A[i * N + j] = -gamma * h * J[i * N + j] + (i == j ? 1.0 : 0.0);
            lu_factor(A, piv, N);
            k1 = f0;
            lu_solve(A, piv, N, k1);
            for (int i = 0; i < N; ++i) y1[i] = y[i] + h * k1[i];
            species_rates_state<double>(*m, conp, param[s], y1.data(), f1.data(), w);
            for (int i = 0; i < N; ++i) k2[i] = f1[i] - 2.0 * k1[i];
            lu_solve(A, piv, N, k2);
            double sq = 0.0;
            for (int i = 0; i < N; ++i) {
                yn[i] = y[i] + 1.5 * h * k1[i] + 0.5 * h * k2[i];
                const double e = 0.5 * h * (k1[i] + k2[i]);
                const double wt = atol + rtol * fmax(fabs(y[i]), fabs(yn[i]));
                sq += (e / wt) * (e / wt);
            }
            const double err = sqrt(sq / N);
            const bool finite = err == err && err < 1e300;
            const bool accept = fixed ? finite : (finite && err <= 1.0);
            ++att;
            double hn = h;
            if (accept) {
                y = yn;
                t += h;
                ++stats[0];
                if (!fixed) hn = h * fmin(rejlast ? 1.0 : 5.0, fmax(0.2, 0.9 / sqrt(fmax(err, 1e-10))));
                rejlast = false;
            } else {
                ++stats[1];
                if (!fixed) hn = h * (finite ? fmax(0.2, 0.9 / sqrt(err)) : 0.2);
                rejlast = true;
            }
            if (t_end - t <= 1e-12 * t_end) {
                done = true;
                t = t_end;
            } else {
                hn = fmin(hn, t_end - t);
            }
            h = hn;
        }
        stats[2] = std::max(stats[2], att);
        if (!done) ++stats[3];
        for (int i = 0; i < N; ++i) ys[i] = y[i];
    }
    return PYJAC_OK;
}

int oracle_abi_version(void) { return PYJAC_MECH_ABI_VERSION; }

}  // extern "C"
