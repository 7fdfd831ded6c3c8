// pyjac_wave.hip -- wave-per-state species-rate and Jacobian kernels (gfx950).
//
// One wavefront evaluates one thermochemical state; its 64 lanes share the
// work inside the state: lanes over species for the thermodynamics and the
// concentration/production-rate bookkeeping, lanes over reactions for the
// rate constants, equilibrium constants and rates of progress, lanes over
// stored Jacobian entries for the write-out.  Everything a state needs at
// random access -- concentrations, Gibbs terms, production rates, third-body
// concentrations, falloff intermediates and the Jacobian accumulator --
// sits in the wave's slice of LDS (a few KB for H2/O2, ~19 KB for the GRI
// sparse Jacobian), so the only HBM traffic is the state in and the result
// out.  Reaction contributions are scattered into the LDS accumulators with
// ds_add_f64; one wave owns a state, so the accumulation order is fixed by
// program order and the result is deterministic.
//
// Workgroups are one wave.  The persistent grid walks the states with an
// XCD-aware remap: consecutive states are handled by workgroups on the same
// XCD at the same time, so their state-fastest (F order) writes combine in
// that XCD's L2 into full lines.
//
// The arithmetic follows the reference's generated code term by term
// (core/rate_subs.py, core/create_jacobian.py); citations per block.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_mech.h"
#include "kernels.h"

namespace pjw {

constexpr double RU = 8314.4621;
constexpr double PA = 101325.0;
constexpr double EXP_MAX = 690.775527898;
constexpr double SMALL = 1e-300;
constexpr double T_MIN = 100.0, T_MAX = 10000.0;
constexpr double V_MIN = 1e-30;
constexpr double LN10 = 2.302585092994046;
constexpr double LN10SQ = LN10 * LN10;
constexpr double RU_INV = 1.0 / RU;

template <class T>
__device__ __forceinline__ T tb(const T* p, int i) {
    return ((const __attribute__((address_space(4))) T*)p)[i];
}

__device__ __forceinline__ double fast_powi(double val, int p) {
    // loopy_utils/preambles_and_manglers.py:123-174
    if (p < 0) {
        val = 1.0 / val;
        p = -p;
    }
    switch (p) {
        case 0: return 1.0;
        case 1: return val;
        case 2: return val * val;
        case 3: return val * val * val;
        case 4: return val * val * val * val;
        case 5: return val * val * val * val * val;
    }
    double r = val * val * val * val * val * val;
    for (int i = 6; i < p; ++i) r *= val;
    return r;
}

__device__ __forceinline__ double clampT(double T) { return fmin(T_MAX, fmax(T_MIN, T)); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
    return v;
}

__device__ __forceinline__ void lds_fence() { __syncthreads(); }

// value of lane l (wave-uniform result)
__device__ __forceinline__ double readlane_d(double v, int l) {
    int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// s + v[lane 0] + v[lane 1] + ... + v[lane cnt-1], added one after the
// other: the summation order of the reference's sequential loops, which
// matters where the sum cancels (e.g. the bath-gas concentration by
// difference, the enthalpy-weighted production rate)
__device__ __forceinline__ double seq_sum(double s, double v, int cnt) {
    for (int l = 0; l < cnt; ++l) s = s + readlane_d(v, l);
    return s;
}

// Arrhenius forms, rate_subs.py:2803-2870
__device__ __forceinline__ double arrhenius(int form, double A, double b, double Ta,
                                            double Tval, double logT, double Tinv) {
    switch (form) {
        case 0: return A;
        case 1: return A * fast_powi(Tval, (int)b);
        case 3: return exp(fmin(EXP_MAX, A + logT * b));
        case 4: return exp(fmin(EXP_MAX, A - Ta * Tinv));
        default: return exp(fmin(EXP_MAX, A + logT * b + -1.0 * Ta * Tinv));
    }
}

// NASA-7 polynomials (rate_subs.py:3335-3354)
struct Nasa {
    double a0, a1, a2, a3, a4, a5, a6;
};
__device__ __forceinline__ Nasa nasa(const DevMech& m, int k, double T) {
    const double* a = (T < m.T_mid[k]) ? m.a_lo + 7 * k : m.a_hi + 7 * k;
    return Nasa{a[0], a[1], a[2], a[3], a[4], a[5], a[6]};
}

// Per-reaction record built by the runtime, padded to MAXSP species so the
// per-species loops unroll and all loads of a reaction issue together:
//   rx_i.x = kind | (nu_sum + 128) << 8   kind: 0-4 Arrhenius form,
//                                          5 PLOG, 6 Chebyshev
//   rx_i.y = reverse index or -1,  rx_i.z = third-body index or -1,
//   rx_i.w = number of species of the reaction (<= MAXSP)
//   rx_d   = A, b, Ta, index of the PLOG/Chebyshev rate
//   rx_s[8 r + q] = species | nu_reac << 16 | nu_prod << 24 for q < MAXSP
constexpr int MAXSP = 6;
__device__ __forceinline__ int sp_spec(int p) { return p & 0xffff; }
__device__ __forceinline__ int sp_nur(int p) { return (p >> 16) & 0xff; }
__device__ __forceinline__ int sp_nup(int p) { return (p >> 24) & 0xff; }

struct Rxn {
    int kind, nsum, rev, thd, cnt;
    double A, b, Ta;
    int kidx;
    int spec[MAXSP], nur[MAXSP], nup[MAXSP];
};

__device__ __forceinline__ Rxn load_rxn(const int4* rx_i, const double4* rx_d, const int4* rx_s,
                                        int r) {
    Rxn x;
    int4 ri = rx_i[r];
    double4 d = rx_d[r];
    int4 s0 = rx_s[2 * r], s1 = rx_s[2 * r + 1];
    x.kind = ri.x & 0xff;
    x.nsum = (ri.x >> 8) - 128;
    x.rev = ri.y;
    x.thd = ri.z;
    x.cnt = ri.w;
    x.A = d.x;
    x.b = d.y;
    x.Ta = d.z;
    x.kidx = (int)d.w;
    int sp[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
#pragma unroll
    for (int q = 0; q < MAXSP; ++q) {
        x.spec[q] = sp_spec(sp[q]);
        x.nur[q] = sp_nur(sp[q]);
        x.nup[q] = sp_nup(sp[q]);
    }
    return x;
}

// LDS layout of one wave (in doubles)
struct WaveLds {
    int conc, b, wdot, energy, heat, pmod, kspec, fall, U, dheat, db, jac, total;
};

__host__ __device__ inline WaveLds wave_lds(const DevMech& m, bool jac, int jac_vals) {
    WaveLds L;
    int o = 0;
    L.conc = o; o += m.Ns + 1;  // + padding slot
    L.b = o; o += m.Ns + 1;
    L.wdot = o; o += m.Ns + 1;
    L.energy = o; o += m.Ns;
    L.heat = o; o += m.Ns;
    L.pmod = o; o += m.n_thd > 0 ? m.n_thd : 1;
    L.kspec = o; o += m.n_plog + m.n_cheb + 1;
    L.fall = o; o += jac ? 8 * (m.n_fall > 0 ? m.n_fall : 1) : 0;
    L.U = o; o += jac ? m.Ns : 0;
    L.dheat = o; o += jac ? m.Ns : 0;
    L.db = o; o += jac ? m.Ns : 0;
    L.jac = o; o += jac ? jac_vals : 0;  // dense NN x NN accumulator
    L.total = o;
    return L;
}

// falloff intermediates kept for the Jacobian, per falloff index f
enum { F_KFALL = 0, F_PR, F_FI, F_FC, F_A, F_B, F_X, F_KF };

struct Ctx {
    const DevMech& m;
    double* sm;  // this wave's LDS slice
    const WaveLds& L;
    int lane;
    double T, P, V, Tc, logT, Tinv, logP;
    __device__ double& conc(int k) { return sm[L.conc + k]; }
    __device__ double& b(int k) { return sm[L.b + k]; }
    __device__ double& wdot(int k) { return sm[L.wdot + k]; }
    __device__ double& energy(int k) { return sm[L.energy + k]; }
    __device__ double& heat(int k) { return sm[L.heat + k]; }
    __device__ double& pmod(int t) { return sm[L.pmod + t]; }
    __device__ double& kspec(int i) { return sm[L.kspec + i]; }
    __device__ double& fall(int f, int q) { return sm[L.fall + 8 * f + q]; }
    __device__ double& U(int k) { return sm[L.U + k]; }
    __device__ double& dheat(int k) { return sm[L.dheat + k]; }
    __device__ double& db(int k) { return sm[L.db + k]; }
};


// Chebyshev polynomials of the first kind (first = x) or second kind
// (first = 2x) at x, evaluated by the reference's recurrence
// T_p = 2 x T_{p-1} - T_{p-2}; the values equal the reference's stored
// polynomial arrays bit for bit
struct ChebPoly {
    double x, pm2, pm1;
    int p;
    __device__ ChebPoly(double x_, double first) : x(x_), pm2(1.0), pm1(first), p(0) {}
    __device__ double next() {  // value of degree p, then advance
        double v;
        if (p == 0) v = 1.0;
        else if (p == 1) v = pm1;
        else {
            v = 2.0 * x * pm1 - pm2;
            pm2 = pm1;
            pm1 = v;
        }
        ++p;
        return v;
    }
};

// get_cheb_arrhenius_rates, rate_subs.py:1904-2081
__device__ __noinline__ void cheb_rate(Ctx& c, int q) {
    const DevMech& m = c.m;
    const int maxT = m.cheb_max_t, maxP = m.cheb_max_p;
    double Pmin = m.cheb_plim[2 * q], Pmax = m.cheb_plim[2 * q + 1];
    double Tmin = m.cheb_tlim[2 * q], Tmax = m.cheb_tlim[2 * q + 1];
    double Tred = (-Tmax - Tmin + 2.0 * c.Tinv) / (Tmax - Tmin);
    double Pred = (-Pmax - Pmin + 2.0 * c.logP) / (Pmax - Pmin);
    int numP = m.cheb_np[q], numT = m.cheb_nt[q];
    const double* par = m.cheb_params + (size_t)q * maxT * maxP;
    double kt = 0.0;
    ChebPoly tp(Tred, Tred);
    for (int t = 0; t < maxT; ++t) {
        double tt = tp.next();
        double temp = 0.0;
        ChebPoly pp(Pred, Pred);
        for (int k = 0; k < maxP; ++k) {
            double pk = pp.next();
            if (k < numP) temp = temp + pk * par[t * maxP + k];
        }
        if (t < numT) kt = kt + tt * temp;
    }
    c.kspec(m.n_plog + q) = exp(LN10 * fmin(EXP_MAX, kt));
}

// Phase 1: concentrations, thermodynamics, third-body concentrations,
// PLOG/Chebyshev rate constants and falloff pressure modification.
template <bool JAC, class PhiF>
__device__ void state_setup(Ctx& c, int conp, double param, PhiF phi) {
    const DevMech& m = c.m;
    const int Ns = m.Ns;
    c.T = phi(0);
    if (conp) { c.P = param; c.V = phi(1); }
    else { c.P = phi(1); c.V = param; }
    c.Tc = clampT(c.T);
    c.logT = log(c.Tc);
    c.Tinv = 1.0 / c.Tc;
    c.logP = log(fmax(SMALL, c.P));
    // get_concentrations, rate_subs.py:597-706
    const double V_inv = 1.0 / fmax(V_MIN, c.V);
    double nsum = 0.0;
    for (int k0 = 0; k0 < Ns - 1; k0 += 64) {
        int k = k0 + c.lane;
        double n = 0.0;
        if (k < Ns - 1) {
            n = fmax(SMALL, phi(2 + k));
            c.conc(k) = n * V_inv;
        }
        nsum = seq_sum(nsum, n, min(64, Ns - 1 - k0));
    }
    if (c.lane == 0) c.conc(Ns - 1) = c.P / (RU * c.Tc) + -1.0 * fmax(SMALL, nsum) * V_inv;
    // b, energy (h|u), heat capacity (cp|cv): rate_subs.py:3281-3388
    for (int k = c.lane; k < Ns; k += 64) {
        Nasa a = nasa(m, k, c.Tc);
        const double T = c.Tc;
        c.b(k) = T * (T * (T * (T * a.a4 / 20.0 + a.a3 / 12.0) + a.a2 / 6.0) + a.a1 / 2.0) +
                 (a.a0 + -1.0) * c.logT + -1.0 * a.a0 + a.a6 + -1.0 * a.a5 * c.Tinv;
        double x = T * (T * (T * (T * (T * a.a4 / 5.0 + a.a3 / 4.0) + a.a2 / 3.0) + a.a1 / 2.0) + a.a0);
        double y = T * (T * (T * (T * a.a4 + a.a3) + a.a2) + a.a1) + a.a0;
        c.energy(k) = RU * (conp ? x + a.a5 : x - T + a.a5);
        c.heat(k) = RU * (conp ? y : y - 1.0);
        c.wdot(k) = 0.0;
        if (JAC) {
            c.U(k) = 0.0;
            c.dheat(k) = RU * (T * (T * (4.0 * T * a.a4 + 3.0 * a.a3) + 2.0 * a.a2) + a.a1);
            c.db(k) = T * (T * (T * a.a4 / 5.0 + a.a3 / 4.0) + a.a2 / 3.0) + a.a1 / 2.0 +
                      c.Tinv * (a.a0 + -1.0 + a.a5 * c.Tinv);
        }
    }
    lds_fence();
    // get_thd_body_concs, rate_subs.py:1788-1901 (unclamped T)
    for (int t = c.lane; t < m.n_thd; t += 64) {
        int not_spec = m.thd_type[t] != 2;
        double thd = c.P * (double)not_spec / (RU * c.T);
        int e = m.thd_off[t + 1];
        for (int s = m.thd_off[t]; s < e; ++s)
            thd = thd + (m.thd_eff[s] + -1.0 * not_spec) * c.conc(m.thd_spec[s]);
        c.pmod(t) = thd;
    }
    // PLOG rates, rate_subs.py:2084-2248
    for (int p = c.lane; p < m.n_plog; p += 64) {
        const int mp = m.plog_max_p, np_ = m.n_plog;
        const double* pl = m.plog_params + (size_t)p * mp;
        const size_t q = (size_t)np_ * mp;
        int numP = m.plog_np[p] - 1;
        bool lower = c.logP <= pl[0];
        bool upper = c.logP > pl[numP];
        int lo = 0, hi = 0;
        if (upper) { lo = numP; hi = numP; }
        for (int k = 0; k < mp - 1; ++k)
            if (k < numP && c.logP > pl[k] && c.logP <= pl[k + 1]) { lo = k; hi = k + 1; }
        double logk1 = pl[q + lo] + c.logT * pl[2 * q + lo] - pl[3 * q + lo] * c.Tinv;
        double logk2 = pl[q + hi] + c.logT * pl[2 * q + hi] - pl[3 * q + hi] * c.Tinv;
        double kt = logk1;
        if (!(lower || upper)) kt = (-logk1 + logk2) * (c.logP - pl[lo]) / (pl[hi] - pl[lo]) + kt;
        c.kspec(p) = exp(fmin(EXP_MAX, kt));
    }
    // Chebyshev rates, rate_subs.py:1904-2081
    for (int q = c.lane; q < m.n_cheb; q += 64) cheb_rate(c, q);
    lds_fence();
    // falloff: kf, kf_fall, Pr, blending, pressure modification
    // (rate_subs.py:2251-2639, :1510-1624)
    for (int f = c.lane; f < m.n_fall; f += 64) {
        int r = m.fall_map[f];
        int si = m.simple_mask[r];
        double kf = arrhenius(m.simple_form[si], m.simple_A[si], m.simple_b[si], m.simple_Ta[si],
                              c.Tc, c.logT, c.Tinv);
        double kff = arrhenius(m.fall_form[f], m.fall_A[f], m.fall_b[f], m.fall_Ta[f], c.Tc,
                               c.logT, c.Tinv);
        int ftype = m.fall_type[f];
        double kinf = ftype ? kff : kf;
        double k0 = ftype ? kf : kff;
        int t = m.fall_to_thd[f];
        double Pr = c.pmod(t) * k0 / kinf;
        int blend = m.fall_blend[f];
        double Fi = 1.0, Fc = 0.0, A = 0.0, B = 0.0, X = 0.0;
        if (blend == 2) {
            int i = m.fall_param[f];
            double a = m.troe_a[i];
            double logPr = log10(fmax(SMALL, Pr));
            Fc = a * exp(fmin(EXP_MAX, -1.0 * c.Tc * m.troe_T1inv[i])) +
                 (1.0 + -1.0 * a) * exp(fmin(EXP_MAX, -1.0 * c.Tc * m.troe_T3inv[i]));
            double T2 = m.troe_T2[i];
            if (T2 != 0.0) Fc = Fc + exp(fmin(EXP_MAX, -1.0 * T2 / c.Tc));
            double logFc = log10(fmax(SMALL, Fc));
            B = -1.1762 * logFc + -1.0 * 0.14 * logPr + 0.806;
            A = -0.67 * logFc + logPr + -0.4;
            Fi = pow(fmax(SMALL, Fc), 1.0 / (A * A / (B * B) + 1.0));
        } else if (blend == 3) {
            int i = m.fall_param[f];
            double logPr = log10(fmax(SMALL, Pr));
            X = 1.0 / (logPr * logPr + 1.0);
            double base = m.sri_a[i] * exp(fmin(EXP_MAX, -m.sri_b[i] * c.Tinv)) +
                          exp(fmin(EXP_MAX, -c.Tc / m.sri_c[i]));
            Fi = pow(base, X);
            double d = m.sri_d[i], e = m.sri_e[i];
            if (d != 1.0) Fi = Fi * d;
            if (e != 0.0) Fi = Fi * pow(c.Tc, e);
        }
        double ci = Fi / (1.0 + Pr);
        if (!ftype) ci = ci * Pr;
        if (JAC) {
            c.fall(f, F_KFALL) = kff; c.fall(f, F_PR) = Pr; c.fall(f, F_FI) = Fi;
            c.fall(f, F_FC) = Fc; c.fall(f, F_A) = A; c.fall(f, F_B) = B; c.fall(f, F_X) = X;
            c.fall(f, F_KF) = kf;
        }
        // the third-body concentration of t is read only by this lane
        c.pmod(t) = ci;
    }
    lds_fence();
}

// rate constants and rates of progress of one reaction (one lane)
struct RxnVals {
    double kf, kr, rf, rr, ci;
    double cv[MAXSP];  // concentrations of the reaction's species
};

__device__ __forceinline__ RxnVals reaction_rates(Ctx& c, const Rxn& x) {
    RxnVals v;
#pragma unroll
    for (int q = 0; q < MAXSP; ++q) v.cv[q] = c.conc(q < x.cnt ? x.spec[q] : c.m.Ns);
    if (x.kind <= 4) v.kf = arrhenius(x.kind, x.A, x.b, x.Ta, c.Tc, c.logT, c.Tinv);
    else v.kf = c.kspec(x.kidx);
    // get_rop, rate_subs.py:1371-1507
    v.rf = v.kf;
#pragma unroll
    for (int q = 0; q < MAXSP; ++q)
        if (q < x.cnt) v.rf = v.rf * fast_powi(v.cv[q], x.nur[q]);
    v.kr = 0.0;
    v.rr = 0.0;
    if (x.rev >= 0) {
        // get_rev_rates, rate_subs.py:1627-1785
        double P_val = !(x.nsum > 0) ? RU / PA : PA / RU;
        double P_sum = fast_powi(P_val, x.nsum < 0 ? -x.nsum : x.nsum);
        double B_sum = 0.0;
#pragma unroll
        for (int q = 0; q < MAXSP; ++q) {
            int net_nu = x.nup[q] + -1 * x.nur[q];
            if (q < x.cnt && net_nu != 0) B_sum = B_sum + (double)net_nu * c.b(x.spec[q]);
        }
        B_sum = exp(fmin(EXP_MAX, B_sum));
        v.kr = v.kf / (P_sum * B_sum);
        v.rr = v.kr;
#pragma unroll
        for (int q = 0; q < MAXSP; ++q)
            if (q < x.cnt) v.rr = v.rr * fast_powi(v.cv[q], x.nup[q]);
    }
    v.ci = x.thd >= 0 ? c.pmod(x.thd) : 1.0;
    return v;
}

// get_spec_rates, rate_subs.py:1079-1167: nu_k * rop_net into wdot
__device__ __forceinline__ void scatter_wdot(Ctx& c, const Rxn& x, double net) {
#pragma unroll
    for (int q = 0; q < MAXSP; ++q) {
        int nu = x.nup[q] + -1 * x.nur[q];
        if (q < x.cnt && nu != 0) atomicAdd(&c.wdot(x.spec[q]), (double)nu * net);
    }
}

// Phase 3: dphi from the production rates (rate_subs.py:709-1076)
template <class OutF>
__device__ void state_rates_out(Ctx& c, int conp, OutF out, double& Tdot_out) {
    const DevMech& m = c.m;
    const int Ns = m.Ns;
    const double V_val = fmax(V_MIN, c.V);
    double upper = 0.0, lower = 0.0, dE = 0.0;
    for (int k0 = 0; k0 < Ns; k0 += 64) {
        int k = k0 + c.lane;
        double lo_k = 0.0, up_k = 0.0, dE_k = 0.0;
        if (k < Ns) {
            double w = c.wdot(k);
            lo_k = c.conc(k) * c.heat(k);
            up_k = c.energy(k) * w;
            if (k < Ns - 1) {
                out(2 + k, V_val * w);
                dE_k = (1.0 + -1.0 * m.mw_post[k]) * w;
            }
        }
        lower = seq_sum(lower, lo_k, min(64, Ns - k0));
        upper = seq_sum(upper, up_k, min(64, Ns - k0));
        dE = seq_sum(dE, dE_k, min(64, Ns - 1 - k0));
    }
    double Tdot = 0.0 + -1.0 * upper / lower;
    double Edot;
    if (conp) {
        Edot = V_val * Tdot / c.Tc;
        Edot = Edot + V_val * dE * c.Tc * RU / c.P;
    } else {
        Edot = c.P * Tdot / c.Tc;
        Edot = Edot + c.Tc * RU * dE;
    }
    if (c.lane == 0) {
        out(0, Tdot);
        out(1, Edot);
    }
    Tdot_out = Tdot;
}

struct Layout {
    size_t sN, sI;
};

// XCD-aware state order: block g on XCD g%8 takes a contiguous range
__device__ __forceinline__ long remap_block(long g, long G) {
    if (G % 8) return g;
    return (g % 8) * (G / 8) + g / 8;
}

__global__ void __launch_bounds__(64)
rates_wave_kernel(DevMech m, int conp, long n, const double* __restrict__ param,
                  const double* __restrict__ phi, double* __restrict__ dphi, Layout L,
                  const int4* __restrict__ rx_i, const double4* __restrict__ rx_d,
                  const int4* __restrict__ rx_s, WaveLds WL) {
    extern __shared__ double sm[];
    const int lane = threadIdx.x;
    const long G = gridDim.x;
    const long g = remap_block(blockIdx.x, G);
    Ctx c{m, sm, WL, lane};
    if (lane == 0) { sm[WL.conc + m.Ns] = 1.0; sm[WL.b + m.Ns] = 0.0; }
    for (long s = g; s < n; s += G) {
        auto ph = [&](int i) { return phi[s * L.sN + i * L.sI]; };
        state_setup<false>(c, conp, param[s], ph);
        for (int r = lane; r < m.Nr; r += 64) {
            Rxn x = load_rxn(rx_i, rx_d, rx_s, r);
            RxnVals v = reaction_rates(c, x);
            double net = v.rf;
            if (x.rev >= 0) net = net + -1.0 * v.rr;
            if (x.thd >= 0) net = net * v.ci;
            scatter_wdot(c, x, net);
        }
        lds_fence();
        double Tdot;
        auto out = [&](int i, double v) { dphi[s * L.sN + i * L.sI] = v; };
        state_rates_out(c, conp, out, Tdot);
        lds_fence();
    }
}

// ------------------------------------------------------------------ Jacobian
// dense accumulator of the Jacobian: entry (row, col) at row * NN + col
__device__ __forceinline__ void jadd(Ctx& c, int row, int col, double v) {
    atomicAdd(&c.sm[c.L.jac + row * c.m.NN + col], v);
}

__device__ __noinline__ double fall_dnj_factor(Ctx& c, int f) {
    // create_jacobian.py:3780-3890
    const DevMech& m = c.m;
    double Pr = c.fall(f, F_PR);
    double dFi = 0.0;
    int blend = m.fall_blend[f];
    if (blend == 2) {
        double A = c.fall(f, F_A), B = c.fall(f, F_B);
        dFi = A * A + B * B;
        dFi = -2.0 * A * B * (0.14 * A + B) * log(fmax(SMALL, c.fall(f, F_FC))) /
              (fmax(SMALL, Pr) * dFi * dFi * LN10);
    } else if (blend == 3) {
        int i = m.fall_param[f];
        double X = c.fall(f, F_X);
        double lv = log(fmax(SMALL, m.sri_a[i] * exp(fmin(EXP_MAX, -m.sri_b[i] / c.T)) +
                                        exp(fmin(EXP_MAX, -c.T / m.sri_c[i]))));
        dFi = -2.0 * X * X * lv * log(fmax(SMALL, Pr)) / (fmax(SMALL, Pr) * LN10SQ);
    }
    double Fi_fac = dFi, k0, kinf;
    double kf = c.fall(f, F_KF), kff = c.fall(f, F_KFALL);
    if (m.fall_type[f]) {
        k0 = kf;
        kinf = kff;
    } else {
        kinf = kf;
        k0 = kff;
        Fi_fac = Pr * Fi_fac + 1.0;
    }
    return k0 * (c.fall(f, F_FI) * Fi_fac + -1.0 * c.pmod(m.fall_to_thd[f])) / (kinf * (Pr + 1.0));
}

// d(pressure modification)/dT and d/dE of third-body / falloff reaction r
// (create_jacobian.py:362-882 and :2117-2588); returns the two factors the
// rows of the reaction's species are scaled by
__device__ __noinline__ void dci_dT_dE(Ctx& c, int conp, int t, int f, double rop_net, double& dT,
                          double& dE) {
    const DevMech& m = c.m;
    const int ns = m.ns;
    const double T = c.T, Tinv = 1.0 / T;
    int e = m.thd_off[t + 1] - 1;
    int tt = m.thd_type[t];
    // temperature
    double mod = 1.0;
    if (tt == 2) mod = m.thd_spec[e] == ns;
    if (tt == 1 && m.thd_spec[e] == ns) mod = m.thd_eff[e];
    const int r = m.thd_map[t];
    if (f < 0) {
        dT = -1.0 * c.P * mod * RU_INV * Tinv * Tinv * c.V * rop_net;
    } else {
        double thd_fac = -1.0 * c.P * mod * RU_INV * Tinv * Tinv;
        int s = m.simple_mask[r];
        double kf = c.fall(f, F_KF), kff = c.fall(f, F_KFALL);
        double kf_0, kf_inf, beta_0, beta_inf, Ta_0, Ta_inf;
        if (m.fall_type[f]) {
            kf_0 = kf; beta_0 = m.simple_b[s]; Ta_0 = m.simple_Ta[s];
            kf_inf = kff; beta_inf = m.fall_b[f]; Ta_inf = m.fall_Ta[f];
        } else {
            kf_0 = kff; beta_0 = m.fall_b[f]; Ta_0 = m.fall_Ta[f];
            kf_inf = kf; beta_inf = m.simple_b[s]; Ta_inf = m.simple_Ta[s];
        }
        double pmod = c.pmod(t), Pr = c.fall(f, F_PR);
        double theta_Pr = Tinv * (beta_0 - beta_inf + (Ta_0 - Ta_inf) * Tinv);
        double theta_no_Pr = thd_fac * kf_0 / kf_inf;
        double dFi = 0.0;
        int blend = m.fall_blend[f];
        if (blend == 2) {
            int i = m.fall_param[f];
            double a = m.troe_a[i], T1 = m.troe_T1inv[i], T2 = m.troe_T2[i], T3 = m.troe_T3inv[i];
            double A = c.fall(f, F_A), B = c.fall(f, F_B), Fc = c.fall(f, F_FC);
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
        } else if (blend == 3) {
            int i = m.fall_param[f];
            double a = m.sri_a[i], b = m.sri_b[i], e_ = m.sri_e[i];
            double X = c.fall(f, F_X);
            double cinv = 1.0 / m.sri_c[i];
            double ec = exp(fmin(EXP_MAX, -T * cinv));
            double eb = exp(fmin(EXP_MAX, -b * Tinv));
            double lv = log(fmax(SMALL, a * eb + ec));
            dFi = -X * (ec * cinv - a * b * Tinv * Tinv * eb) / (a * eb + ec) + e_ * Tinv -
                  2 * X * X * lv * (Pr * theta_Pr + theta_no_Pr) * log(fmax(SMALL, Pr)) /
                      (fmax(SMALL, Pr) * LN10SQ);
        }
        double d = pmod * (-1.0 * (Pr * theta_Pr + theta_no_Pr) / (Pr + 1.0) + dFi);
        if (!m.fall_type[f]) d = d + theta_Pr * pmod + c.fall(f, F_FI) * theta_no_Pr / (Pr + 1.0);
        dT = d * c.V * rop_net;
    }
    // extra variable
    const double rt_inv = 1.0 / (RU * T);
    mod = conp ? (double)(tt == 1) : 1.0;
    if (tt == 1 && m.thd_spec[e] == ns) mod = m.thd_eff[e];
    if (tt == 2) mod = m.thd_spec[e] == ns;
    if (f < 0) {
        if (conp) {
            if (tt != 3) mod = mod * c.P * rt_inv - c.pmod(t);
            dE = mod * rop_net;
        } else {
            dE = mod * rt_inv * c.V * rop_net;
        }
        return;
    }
    int not_unity = tt != 3;
    double kf = c.fall(f, F_KF), kff = c.fall(f, F_KFALL);
    double kf_0, kf_inf;
    if (m.fall_type[f]) { kf_0 = kf; kf_inf = kff; }
    else { kf_0 = kff; kf_inf = kf; }
    if (conp) mod = mod * c.P * rt_inv * kf_0 / kf_inf;
    else mod = mod * rt_inv * kf_0 / kf_inf;
    double Pr = c.fall(f, F_PR);
    double theta = conp ? (mod + -1.0 * Pr * not_unity) : mod;
    double dFi = 0.0;
    int blend = m.fall_blend[f];
    if (blend == 2) {
        double A = c.fall(f, F_A), B = c.fall(f, F_B);
        double absqsq = A * A + B * B;
        absqsq = absqsq * absqsq;
        dFi = -2.0 * A * B * log(fmax(SMALL, c.fall(f, F_FC))) * (0.14 * A + B) * theta /
              (fmax(SMALL, Pr) * absqsq * LN10);
    } else if (blend == 3) {
        int i = m.fall_param[f];
        double a = m.sri_a[i], b = m.sri_b[i], cinv = 1.0 / m.sri_c[i];
        double X = c.fall(f, F_X);
        double inner;
        if (conp) {
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
    double pm = c.pmod(t);
    double num = conp ? (-mod + Pr * not_unity) : -mod;
    double d = pm * (num / (Pr + 1) + dFi);
    if (!m.fall_type[f]) {
        if (conp) d = d + c.fall(f, F_FI) * mod / (Pr + 1) + -1.0 * not_unity * pm;
        else d = d + c.fall(f, F_FI) * mod / (Pr + 1);
    }
    dE = conp ? d * rop_net : d * c.V * rop_net;
}

__device__ __forceinline__ double plog_par(const DevMech& m, int q, int p, int k) {
    return m.plog_params[((size_t)q * m.n_plog + p) * m.plog_max_p + k];
}

// d ln kf / dT of a PLOG / Chebyshev reaction (create_jacobian.py:2823-2976)
// and, for CONV, d ln kf / dP (create_jacobian.py:1158-1300)
__device__ __noinline__ void special_dlnk(Ctx& c, int conp, int kind, int idx, double& dkT,
                                          double& dkP) {
    const DevMech& m = c.m;
    const double Tinv = 1.0 / c.T;
    dkT = 0.0;
    dkP = 0.0;
    if (kind == 5) {
        int p = idx;
        int numP = m.plog_np[p] - 1;
        int lo = 0, hi = numP;
        for (int k = 0; k < m.plog_max_p - 1; ++k)
            if (k <= numP && c.logP > plog_par(m, 0, p, k) && c.logP <= plog_par(m, 0, p, k + 1)) {
                lo = k;
                hi = k + 1;
            }
        if (c.logP > plog_par(m, 0, p, numP))
            dkT = (plog_par(m, 2, p, hi) + plog_par(m, 3, p, hi) * Tinv) * Tinv;
        else
            dkT = (plog_par(m, 2, p, lo) + plog_par(m, 3, p, lo) * Tinv) * Tinv;
        bool inr = c.logP > plog_par(m, 0, p, 0) && c.logP <= plog_par(m, 0, p, numP);
        if (inr)
            dkT = dkT + Tinv * (c.logP - plog_par(m, 0, p, lo)) *
                            (plog_par(m, 2, p, hi) - plog_par(m, 2, p, lo) +
                             (plog_par(m, 3, p, hi) - plog_par(m, 3, p, lo)) * Tinv) /
                            (plog_par(m, 0, p, hi) - plog_par(m, 0, p, lo));
        if (!conp && inr) {
            const double logT = log(fmax(SMALL, c.T));
            dkP = (plog_par(m, 1, p, hi) - plog_par(m, 1, p, lo) +
                   logT * (plog_par(m, 2, p, hi) - plog_par(m, 2, p, lo)) -
                   (plog_par(m, 3, p, hi) - plog_par(m, 3, p, lo)) * Tinv) /
                  (c.P * (plog_par(m, 0, p, hi) - plog_par(m, 0, p, lo)));
        }
        return;
    }
    int q = idx;
    const int maxT = m.cheb_max_t, maxP = m.cheb_max_p;
    double Pmin = m.cheb_plim[2 * q], Pmax = m.cheb_plim[2 * q + 1];
    double Tmin = m.cheb_tlim[2 * q], Tmax = m.cheb_tlim[2 * q + 1];
    double Tred = (2 * Tinv - Tmax - Tmin) / (Tmax - Tmin);
    double Pred = (2 * c.logP - Pmax - Pmin) / (Pmax - Pmin);
    const double* par = m.cheb_params + (size_t)q * maxT * maxP;
    {
        int numP = m.cheb_np[q], numT = m.cheb_nt[q] - 1;
        double dkf = 0;
        ChebPoly tp(Tred, 2 * Tred);
        for (int t = 0; t < maxT - 1; ++t) {
            double tt = tp.next();
            double temp = 0;
            ChebPoly pp(Pred, Pred);
            for (int k = 0; k < maxP; ++k) {
                double pk = pp.next();
                if (k < numP) temp = temp + pk * par[(t + 1) * maxP + k];
            }
            if (t < numT) dkf = dkf + (t + 1) * tt * temp;
        }
        dkT = -dkf * 2 * LN10 * Tinv * Tinv / (Tmax - Tmin);
    }
    if (!conp) {
        int numP = m.cheb_np[q] - 1, numT = m.cheb_nt[q];
        double dkf = 0;
        ChebPoly tp(Tred, Tred);
        for (int t = 0; t < maxT; ++t) {
            double tt = tp.next();
            double temp = 0;
            ChebPoly pp(Pred, 2 * Pred);
            for (int k = 0; k < maxP - 1; ++k) {
                double pk = pp.next();
                if (k < numP) temp = temp + (k + 1) * pk * par[t * maxP + k + 1];
            }
            if (t < numT) dkf = dkf + tt * temp;
        }
        dkP = dkf * 2 * LN10 / (c.P * (Pmax - Pmin));
    }
}

// all per-reaction contributions of one reaction to the Jacobian
// (create_jacobian.py kernels dRopidnj(+_ns), dci_*_dnj(+_ns), dRopi*_dT
// (+_ns), dci_*_dT, dRopi*_dE(+_ns), dci_*_dE), summed per reaction and
// scattered into the LDS accumulator
__device__ __forceinline__ void reaction_jacobian(Ctx& c, int conp, const Rxn& x,
                                                  const RxnVals& v, double ropi) {
    const DevMech& m = c.m;
    const int ns = m.ns, NN = m.NN;
    const double T = c.T, Tinv = 1.0 / T;
    const double kr_i = x.rev >= 0 ? v.kr : 0.0;
    double cg[MAXSP];
#pragma unroll
    for (int q = 0; q < MAXSP; ++q) cg[q] = fmax(SMALL, v.cv[q]);
    // dRopidnj (create_jacobian.py:4169-4484)
#pragma unroll
    for (int j = 0; j < MAXSP; ++j) {
        if (j >= x.cnt || x.spec[j] == ns) continue;
        double Sr = x.nup[j], Sf = x.nur[j];
#pragma unroll
        for (int i = 0; i < MAXSP; ++i) {
            if (i >= x.cnt) continue;
            int nr = x.nup[i] - (i == j), nf = x.nur[i] - (i == j);
            Sr = Sr * fast_powi(cg[i], nr);
            Sf = Sf * fast_powi(cg[i], nf);
        }
        double d = (v.kf * Sf + -1.0 * kr_i * Sr) * v.ci;
        const int col = x.spec[j] + 2;
#pragma unroll
        for (int k = 0; k < MAXSP; ++k)
            if (k < x.cnt && x.spec[k] != ns)
                jadd(c, x.spec[k] + 2, col, d * (x.nup[k] - x.nur[k]));
    }
    // temperature / extra-variable derivative of the rate of progress
    double dkT, dkP = 0.0;
    if (x.kind <= 4) dkT = (x.b + x.Ta * Tinv) * Tinv;
    else special_dlnk(c, conp, x.kind, x.kind == 6 ? x.kidx - m.n_plog : x.kidx, dkT, dkP);
    double dT, dE = 0.0;
    {
        // create_jacobian.py:2978-2992
        double d = v.rf * dkT;
        if (x.rev >= 0) {
            double dB = 0.0;
#pragma unroll
            for (int q = 0; q < MAXSP; ++q)
                if (q < x.cnt) dB = dB + (x.nup[q] + -1.0 * x.nur[q]) * c.db(x.spec[q]);
            d = d + -1.0 * v.rr * (dkT + -1.0 * dB);
        }
        dT = d * v.ci * c.V;
    }
    if (conp) {
        // d(V w)/dV, create_jacobian.py:1040-1080
        int nu_f = -1, nu_r = -1;
#pragma unroll
        for (int q = 0; q < MAXSP; ++q)
            if (q < x.cnt) { nu_f += x.nur[q]; nu_r += x.nup[q]; }
        double d = -1.0 * nu_f * v.rf;
        if (x.rev >= 0) d = d + nu_r * v.rr;
        if (x.thd >= 0 && x.kind <= 4) d = d * v.ci;
        dE = d;
    } else if (x.kind > 4) {
        double d = v.rf;
        if (x.rev >= 0) d = d - v.rr;
        dE = d * dkP * c.V;
    }
    // bath-gas terms (the _ns kernels)
    int last = 0, nur_l = 0, nup_l = 0;
#pragma unroll
    for (int q = 0; q < MAXSP; ++q)
        if (q == x.cnt - 1) { last = x.spec[q]; nur_l = x.nur[q]; nup_l = x.nup[q]; }
    if (last == ns) {
        double Sf = nur_l, Sr = nup_l;
#pragma unroll
        for (int q = 0; q < MAXSP; ++q) {
            if (q >= x.cnt) continue;
            int is_ns = x.spec[q] == ns;
            Sf = Sf * fast_powi(cg[q], x.nur[q] - is_ns);
            Sr = Sr * fast_powi(cg[q], x.nup[q] - is_ns);
        }
        dT += (Sr * kr_i + -1.0 * Sf * v.kf) * c.V * v.ci * c.P / (RU * T * T);
        double fac = conp ? c.P / (RU * T) : c.V / (RU * T);
        dE += (Sf * v.kf - Sr * kr_i) * v.ci * fac;
        double u = (kr_i * Sr + -1.0 * v.kf * Sf) * v.ci;
#pragma unroll
        for (int q = 0; q < MAXSP; ++q)
            if (q < x.cnt && x.spec[q] != ns) atomicAdd(&c.U(x.spec[q]), u * (x.nup[q] - x.nur[q]));
    }
    // third-body / falloff terms
    if (x.thd >= 0) {
        const int t = x.thd;
        const int r_f = m.thd_map[t];
        const int f = m.fall_mask[r_f];
        const double fac = f >= 0 ? fall_dnj_factor(c, f) : 1.0;
        const int tt = m.thd_type[t];
        const int e0 = m.thd_off[t], e1 = m.thd_off[t + 1];
        for (int sj = e0; sj < e1; ++sj) {
            int j = m.thd_spec[sj];
            if (j == ns) continue;
            double dci = 0.0;
            if (tt == 2) dci = 1.0;
            if (tt == 1) dci = m.thd_eff[sj] + -1.0;
#pragma unroll
            for (int q = 0; q < MAXSP; ++q) {
                if (q >= x.cnt || x.spec[q] == ns) continue;
                int nu_k = x.nup[q] - x.nur[q];
                jadd(c, x.spec[q] + 2, j + 2, f < 0 ? nu_k * dci * ropi : nu_k * dci * ropi * fac);
            }
        }
        if (e1 > e0 && m.thd_spec[e1 - 1] == ns) {
            double ns_eff = m.thd_eff[e1 - 1];
            double dci = 0.0;
            if (tt == 2) dci = -1.0;
            if (tt == 1) dci = 1.0 + -1.0 * ns_eff;
#pragma unroll
            for (int q = 0; q < MAXSP; ++q) {
                if (q >= x.cnt || x.spec[q] == ns) continue;
                int nu_k = x.nup[q] - x.nur[q];
                atomicAdd(&c.U(x.spec[q]), f < 0 ? nu_k * dci * ropi : nu_k * dci * ropi * fac);
            }
        }
        double dTc, dEc;
        dci_dT_dE(c, conp, t, f, ropi, dTc, dEc);
        dT += dTc;
        dE += dEc;
    }
#pragma unroll
    for (int q = 0; q < MAXSP; ++q) {
        if (q >= x.cnt || x.spec[q] == ns) continue;
        double nu_k = x.nup[q] + -1.0 * x.nur[q];
        const int row = x.spec[q] + 2;
        atomicAdd(&c.sm[c.L.jac + row * NN], nu_k * dT);
        atomicAdd(&c.sm[c.L.jac + row * NN + 1], nu_k * dE);
    }
}

__global__ void __launch_bounds__(64)
jacobian_wave_kernel(DevMech m, int conp, long n, const double* __restrict__ param,
                     const double* __restrict__ phi, double* __restrict__ jac, Layout L,
                     int out_C, const int4* __restrict__ rx_i, const double4* __restrict__ rx_d,
                     const int4* __restrict__ rx_s, WaveLds WL) {
    extern __shared__ double sm[];
    const int lane = threadIdx.x;
    const long G = gridDim.x;
    const long g = remap_block(blockIdx.x, G);
    const int NN = m.NN, ns = m.ns, nout = m.jac_nnz;
    Ctx c{m, sm, WL, lane};
    double* J = sm + WL.jac;
    if (lane == 0) { sm[WL.conc + m.Ns] = 1.0; sm[WL.b + m.Ns] = 0.0; }
    for (long s = g; s < n; s += G) {
        auto ph = [&](int i) { return phi[s * L.sN + i * L.sI]; };
        for (int e = lane; e < NN * NN; e += 64) J[e] = 0.0;
        state_setup<true>(c, conp, param[s], ph);
        // reaction pass: production rates and per-reaction contributions
        for (int r = lane; r < m.Nr; r += 64) {
            Rxn x = load_rxn(rx_i, rx_d, rx_s, r);
            RxnVals v = reaction_rates(c, x);
            double ropi = v.rf;
            if (x.rev >= 0) ropi = ropi + -1.0 * v.rr;
            scatter_wdot(c, x, x.thd >= 0 ? ropi * v.ci : ropi);
            reaction_jacobian(c, conp, x, v, ropi);
        }
        lds_fence();
        double Tdot;
        auto no_out = [&](int, double) {};
        state_rates_out(c, conp, no_out, Tdot);
        // bath-gas sums into every species column of their rows
        for (int k = lane; k < m.Ns - 1; k += 64) {
            double u = c.U(k);
            if (u != 0.0)
                for (int j = 0; j < m.Ns - 1; ++j) J[(k + 2) * NN + j + 2] += u;
        }
        lds_fence();
        // heat capacity total (create_jacobian.py:3487-3555)
        double heat_tot = 0.0;
        for (int k0 = 0; k0 < m.Ns; k0 += 64) {
            int k = k0 + lane;
            double hv = k < m.Ns ? c.heat(k) * c.conc(k) : 0.0;
            heat_tot = seq_sum(heat_tot, hv, min(64, m.Ns - k0));
        }
        heat_tot = 0.0 + heat_tot;
        const double T = c.T, Tinv = 1.0 / T;
        // dTdot_dnj / dEdot_dnj: lanes over species columns
        // (create_jacobian.py:3266-3484)
        const double fixed = conp ? c.P : c.V;
        const double extra = conp ? c.V : c.P;
        const double e_ns = c.energy(ns), h_ns = c.heat(ns);
        for (int j = lane; j < m.Ns - 1; j += 64) {
            double sum = 0.0, sum2 = 0.0;
            for (int k = 0; k < m.Ns - 1; ++k) {
                double jv = J[(k + 2) * NN + j + 2];
                sum = sum + (c.energy(k) + -1.0 * e_ns * m.mw_post[k]) * jv;
                sum2 = sum2 + (1.0 + -1.0 * m.mw_post[k]) * jv;
            }
            double j0 = -1.0 * (sum + Tdot * (c.heat(j) + -1.0 * h_ns)) / (c.V * heat_tot);
            J[j + 2] = j0;
            J[NN + j + 2] = 0.0 + T * RU * sum2 / fixed + extra * j0 * Tinv;
        }
        {
            // columns 0 and 1 of rows 0 and 1, sequential sums in the
            // reference's order (create_jacobian.py:1595-2114)
            const double Vinv = 1.0 / c.V;
            double rate_sum = 0.0, sumT = 0.0, dTsumE = 0.0, specsum = 0.0, sumE = 0.0,
                   sumEE = 0.0;
            double dTsum = (h_ns * Tinv + -1.0 * c.dheat(ns)) * c.conc(ns);
            for (int i0 = 0; i0 < m.Ns - 1; i0 += 64) {
                int i = i0 + lane;
                double ra = 0, rb = 0, dt = 0, st = 0, de = 0, sp = 0, se = 0, see = 0;
                if (i < m.Ns - 1) {
                    double j0 = J[(i + 2) * NN];
                    double j1 = J[(i + 2) * NN + 1];
                    double mw = m.mw_post[i];
                    double w = c.wdot(i);
                    ra = Vinv * j0 * (-1.0 * c.energy(i) + e_ns * mw);
                    rb = w * (-1.0 * c.heat(i) + mw * h_ns);
                    dt = (h_ns * Tinv + -1.0 * c.dheat(i)) * c.conc(i);
                    if (conp) {
                        st = (1.0 + -1.0 * mw) * (Vinv * j0 + Tinv * w);
                        de = (c.heat(i) + -1.0 * h_ns) * c.conc(i);
                        sp = (c.energy(i) + -1.0 * e_ns * mw) * (j1 + -1.0 * w);
                    } else {
                        st = (1.0 - mw) * (j0 * (T / c.V) + w);
                        se = (c.energy(i) - e_ns * mw) * j1;
                    }
                    see = (1 - mw) * j1;
                }
                int cnt = min(64, m.Ns - 1 - i0);
                for (int l = 0; l < cnt; ++l) {
                    rate_sum = rate_sum + readlane_d(ra, l);
                    rate_sum = rate_sum + readlane_d(rb, l);
                }
                dTsum = seq_sum(dTsum, dt, cnt);
                sumT = seq_sum(sumT, st, cnt);
                dTsumE = seq_sum(dTsumE, de, cnt);
                specsum = seq_sum(specsum, sp, cnt);
                sumE = seq_sum(sumE, se, cnt);
                sumEE = seq_sum(sumEE, see, cnt);
            }
            if (lane == 0) {
                double j00 = J[0] + (Tdot * dTsum + rate_sum) / heat_tot;
                J[0] = j00;
                double j10;
                if (conp) {
                    j10 = J[NN] + c.V * Tinv * (j00 + -1.0 * Tinv * Tdot);
                    j10 = j10 + RU * T * c.V * sumT / c.P;
                } else {
                    j10 = J[NN] + c.P * (j00 - Tdot * Tinv) * Tinv;
                    j10 = j10 + RU * sumT;
                }
                J[NN] = j10;
                double j01;
                if (conp) {
                    j01 = J[1] + (Tdot * dTsumE + -1.0 * specsum) * (1.0 / (heat_tot * c.V));
                } else {
                    double spec_inv = 1.0 / heat_tot;
                    j01 = J[1] - (Tdot * h_ns / (RU * T)) * spec_inv;
                    j01 = j01 - (sumE * Vinv) * spec_inv;
                }
                J[1] = j01;
                double var = conp ? c.V : c.P;
                double par = conp ? c.P : c.V;
                double j11 = J[NN + 1] + (var * j01 + Tdot) / T;
                J[NN + 1] = j11 + RU * T * sumEE / par;
            }
        }
        lds_fence();
        // write-out: stored value e (output order) of this state
        for (int e = lane; e < nout; e += 64) {
            double val = J[m.jac_pos[e]];
            if (out_C) jac[s * (size_t)nout + e] = val;
            else jac[(size_t)e * n + s] = val;
        }
        lds_fence();
    }
}

}  // namespace pjw

namespace pjk_host {

size_t wave_lds_bytes(const DevMech& m, bool jac, int jac_vals) {
    return (size_t)pjw::wave_lds(m, jac, jac_vals).total * sizeof(double);
}

hipError_t launch_rates_wave(const DevMech& m, int conp, char order, long n, const double* param,
                             const double* phi, double* dphi, const WaveTables& wt, int grid,
                             hipStream_t stream) {
    pjw::Layout L = order == 'F' ? pjw::Layout{1, (size_t)n} : pjw::Layout{(size_t)m.NN, 1};
    pjw::WaveLds WL = pjw::wave_lds(m, false, 0);
    size_t lds = (size_t)WL.total * sizeof(double);
    hipLaunchKernelGGL(pjw::rates_wave_kernel, dim3(grid), dim3(64), lds, stream, m, conp, n,
                       param, phi, dphi, L, (const int4*)wt.rx_i, (const double4*)wt.rx_d,
                       (const int4*)wt.rx_s, WL);
    return hipGetLastError();
}

hipError_t launch_jacobian_wave(const DevMech& m, int conp, char order, long n,
                                const double* param, const double* phi, double* jac,
                                const WaveTables& wt, int grid, hipStream_t stream) {
    pjw::Layout L = order == 'F' ? pjw::Layout{1, (size_t)n} : pjw::Layout{(size_t)m.NN, 1};
    pjw::WaveLds WL = pjw::wave_lds(m, true, m.NN * m.NN);
    size_t lds = (size_t)WL.total * sizeof(double);
    hipLaunchKernelGGL(pjw::jacobian_wave_kernel, dim3(grid), dim3(64), lds, stream, m, conp, n,
                       param, phi, jac, L, (int)(order == 'C'), (const int4*)wt.rx_i,
                       (const double4*)wt.rx_d, (const int4*)wt.rx_s, WL);
    return hipGetLastError();
}

}  // namespace pjk_host
