// pyjac_kernels.hip -- species-rate and analytical-Jacobian kernels for
// MI355X (gfx950, wave64).
//
// Execution model: one thermochemical state per lane.  Every loop runs over
// mechanism tables with a wave-uniform trip count and wave-uniform indices,
// so the tables are fetched with scalar loads (constant address space) and
// the control flow never diverges inside a wave.  The per-state arrays that
// need random access (concentrations, Gibbs terms, production rates, rate
// constants, the dense Jacobian accumulator) live in a per-wave device
// workspace laid out [element][64 lanes], so every access of a wave is one
// fully coalesced 512-byte transaction.  The kernels are persistent: the
// grid is sized to the device and each wave walks the state tiles with a
// grid stride, reusing its workspace slot, which keeps the workspace
// footprint (and its cache residency) independent of the batch size.
//
// The arithmetic restates the reference's generated code
// (core/rate_subs.py, core/create_jacobian.py; instances in
// examples/opencl/{h2o2,ch4}/*.ocl) expression by expression; function
// comments cite the sub-kernel each block follows.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_mech.h"
#include "kernels.h"

namespace pjk {

constexpr double RU = 8314.4621;
constexpr double PA = 101325.0;
constexpr double EXP_MAX = 690.775527898;
constexpr double SMALL = 1e-300;
constexpr double T_MIN = 100.0, T_MAX = 10000.0;
constexpr double V_MIN = 1e-30;
constexpr double LN10 = 2.302585092994046;
constexpr double LN10SQ = LN10 * LN10;
constexpr double RU_INV = 1.0 / RU;

// wave-uniform table read through the constant address space (scalar load)
template <class T>
__device__ __forceinline__ T tb(const T* p, int i) {
    return ((const __attribute__((address_space(4))) T*)p)[i];
}

// one lane's view of a per-wave SoA workspace array
struct Lane {
    double* p;
    __device__ __forceinline__ double& operator[](int i) const { return p[(size_t)i * 64]; }
};

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

// NASA-7 (rate_subs.py:3335-3354); coefficient pointer for species k
__device__ __forceinline__ const double* nasa(const DevMech& m, int k, double T) {
    return (T < tb(m.T_mid, k)) ? m.a_lo + 7 * k : m.a_hi + 7 * k;
}
__device__ __forceinline__ double nasa_cp(const double* a, double T, bool cv) {
    double x = T * (T * (T * (T * tb(a, 4) + tb(a, 3)) + tb(a, 2)) + tb(a, 1)) + tb(a, 0);
    return RU * (cv ? x - 1.0 : x);
}
__device__ __forceinline__ double nasa_h(const double* a, double T, bool u) {
    double x = T * (T * (T * (T * (T * tb(a, 4) / 5.0 + tb(a, 3) / 4.0) + tb(a, 2) / 3.0) +
                         tb(a, 1) / 2.0) + tb(a, 0));
    return RU * (u ? x - T + tb(a, 5) : x + tb(a, 5));
}
__device__ __forceinline__ double nasa_b(const double* a, double T, double logT, double Tinv) {
    return T * (T * (T * (T * tb(a, 4) / 20.0 + tb(a, 3) / 12.0) + tb(a, 2) / 6.0) + tb(a, 1) / 2.0) +
           (tb(a, 0) + -1.0) * logT + -1.0 * tb(a, 0) + tb(a, 6) + -1.0 * tb(a, 5) * Tinv;
}
__device__ __forceinline__ double nasa_dcp(const double* a, double T) {
    return RU * (T * (T * (4.0 * T * tb(a, 4) + 3.0 * tb(a, 3)) + 2.0 * tb(a, 2)) + tb(a, 1));
}
__device__ __forceinline__ double nasa_db(const double* a, double T, double Tinv) {
    return T * (T * (T * tb(a, 4) / 5.0 + tb(a, 3) / 4.0) + tb(a, 2) / 3.0) + tb(a, 1) / 2.0 +
           Tinv * (tb(a, 0) + -1.0 + tb(a, 5) * Tinv);
}

// workspace layout of one lane (offsets in doubles), see ws_layout()
struct WsLayout {
    int conc, b, wdot, kf, kr, pmod, rop_f, rop_r, kf_fall, Pr, Fi, Fc, At, Bt, X,
        energy, heat, dheat, db, jac, total;
};

__host__ __device__ inline WsLayout ws_layout(int Ns, int Nr, int n_fall, bool jac) {
    WsLayout L;
    int o = 0;
    L.conc = o; o += Ns;
    L.b = o; o += Ns;
    L.wdot = o; o += Ns;
    L.pmod = o; o += (Nr > 0 ? Nr : 1);  // indexed by third-body index (< Nr)
    L.kf = o; o += jac ? Nr : 0;
    L.kr = o; o += jac ? Nr : 0;
    L.rop_f = o; o += jac ? Nr : 0;
    L.rop_r = o; o += jac ? Nr : 0;
    int nf = n_fall > 0 ? n_fall : 1;
    L.kf_fall = o; o += nf;
    L.Pr = o; o += nf;
    L.Fi = o; o += nf;
    L.Fc = o; o += nf;
    L.At = o; o += nf;
    L.Bt = o; o += nf;
    L.X = o; o += nf;
    L.energy = o; o += Ns;
    L.heat = o; o += Ns;
    L.dheat = o; o += jac ? Ns : 0;
    L.db = o; o += jac ? Ns : 0;
    L.jac = o; o += jac ? (Ns + 1) * (Ns + 1) : 0;
    L.total = o;
    return L;
}

struct LaneWs {
    Lane conc, b, wdot, kf, kr, pmod, rop_f, rop_r, kf_fall, Pr, Fi, Fc, At, Bt, X, energy, heat,
        dheat, db, jac;
    __device__ LaneWs(double* base, const WsLayout& L)
        : conc{base + (size_t)L.conc * 64}, b{base + (size_t)L.b * 64},
          wdot{base + (size_t)L.wdot * 64}, kf{base + (size_t)L.kf * 64},
          kr{base + (size_t)L.kr * 64}, pmod{base + (size_t)L.pmod * 64},
          rop_f{base + (size_t)L.rop_f * 64}, rop_r{base + (size_t)L.rop_r * 64},
          kf_fall{base + (size_t)L.kf_fall * 64}, Pr{base + (size_t)L.Pr * 64},
          Fi{base + (size_t)L.Fi * 64}, Fc{base + (size_t)L.Fc * 64},
          At{base + (size_t)L.At * 64}, Bt{base + (size_t)L.Bt * 64},
          X{base + (size_t)L.X * 64}, energy{base + (size_t)L.energy * 64},
          heat{base + (size_t)L.heat * 64}, dheat{base + (size_t)L.dheat * 64},
          db{base + (size_t)L.db * 64}, jac{base + (size_t)L.jac * 64} {}
};

struct StateIn {
    double T, P, V;  // raw T, and P / V by CONP / CONV convention
};

// ------------------------------------------------------------------ rates
// Fused species-rate chain of rate_subs.py:3092-3200 for one lane's state.
// phi(i) reads state variable i.  Writes dphi via out(i).  With JAC the
// per-reaction intermediates the Jacobian needs are kept in the workspace.
template <bool JAC, class PhiF, class OutF>
__device__ void species_rates_lane(const DevMech& m, int conp, double param, PhiF phi,
                                   OutF out, LaneWs& w, StateIn& st) {
    const int Ns = m.Ns, Nr = m.Nr;
    st.T = phi(0);
    if (conp) { st.P = param; st.V = phi(1); }
    else { st.P = phi(1); st.V = param; }
    const double Tc = clampT(st.T);
    const double logT = log(Tc);
    const double Tinv = 1.0 / Tc;

    // get_concentrations, rate_subs.py:597-706
    {
        double V_inv = 1.0 / fmax(V_MIN, st.V);
        double n_sum = 0.0;
        double cns = st.P / (RU * Tc);
        for (int i = 0; i < Ns - 1; ++i) {
            double n = fmax(SMALL, phi(2 + i));
            n_sum = n_sum + n;
            w.conc[i] = n * V_inv;
        }
        w.conc[Ns - 1] = cns + -1.0 * fmax(SMALL, n_sum) * V_inv;
    }
    // eval_b, rate_subs.py:3281-3388
    if (m.n_rev)
        for (int k = 0; k < Ns; ++k) w.b[k] = nasa_b(nasa(m, k, Tc), Tc, logT, Tinv);
    // get_thd_body_concs, rate_subs.py:1788-1901 (unclamped T); stored in
    // the pressure-modification slot, replaced below for falloff
    for (int t = 0; t < m.n_thd; ++t) {
        int tt = tb(m.thd_type, t);
        int not_spec = tt != 2;
        double thd = st.P * (double)not_spec / (RU * st.T);
        int e = tb(m.thd_off, t + 1);
        for (int s = tb(m.thd_off, t); s < e; ++s)
            thd = thd + (tb(m.thd_eff, s) + -1.0 * not_spec) * w.conc[tb(m.thd_spec, s)];
        w.pmod[t] = thd;
    }
    for (int k = 0; k < Ns; ++k) w.wdot[k] = 0.0;

    const double logP = log(fmax(SMALL, st.P));
    for (int r = 0; r < Nr; ++r) {
        // forward rate constant
        double kf;
        int si = tb(m.simple_mask, r);
        if (si >= 0) {
            kf = arrhenius(tb(m.simple_form, si), tb(m.simple_A, si), tb(m.simple_b, si),
                           tb(m.simple_Ta, si), Tc, logT, Tinv);
        } else if (tb(m.plog_mask, r) >= 0) {
            // get_plog_arrhenius_rates, rate_subs.py:2084-2248
            int p = tb(m.plog_mask, r);
            const int mp = m.plog_max_p, np_ = m.n_plog;
            const double* P0 = m.plog_params + (size_t)p * mp;
            int numP = tb(m.plog_np, p) - 1;
            bool lower = logP <= tb(P0, 0);
            bool upper = logP > tb(P0, numP);
            int lo = 0, hi = 0;
            if (upper) { lo = numP; hi = numP; }
            for (int k = 0; k < mp - 1; ++k)
                if (k < numP && logP > tb(P0, k) && logP <= tb(P0, k + 1)) { lo = k; hi = k + 1; }
            const double* pl = m.plog_params + (size_t)p * mp;
            const size_t q = (size_t)np_ * mp;
            double logk1 = tb(pl, q + lo) + logT * tb(pl, 2 * q + lo) - tb(pl, 3 * q + lo) * Tinv;
            double logk2 = tb(pl, q + hi) + logT * tb(pl, 2 * q + hi) - tb(pl, 3 * q + hi) * Tinv;
            double kt = logk1;
            if (!(lower || upper))
                kt = (-logk1 + logk2) * (logP - tb(pl, lo)) / (tb(pl, hi) - tb(pl, lo)) + kt;
            kf = exp(fmin(EXP_MAX, kt));
        } else {
            // get_cheb_arrhenius_rates, rate_subs.py:1904-2081
            int c = tb(m.cheb_mask, r);
            const int maxT = m.cheb_max_t, maxP = m.cheb_max_p;
            const int poly_max = maxP > maxT ? maxP : maxT;
            double Pmin = tb(m.cheb_plim, 2 * c), Pmax = tb(m.cheb_plim, 2 * c + 1);
            double Tmin = tb(m.cheb_tlim, 2 * c), Tmax = tb(m.cheb_tlim, 2 * c + 1);
            double Tred = (-Tmax - Tmin + 2.0 * Tinv) / (Tmax - Tmin);
            double Pred = (-Pmax - Pmin + 2.0 * logP) / (Pmax - Pmin);
            int numP = tb(m.cheb_np, c), numT = tb(m.cheb_nt, c);
            double pp[16], tp[16];
            pp[0] = 1.0; pp[1] = Pred; tp[0] = 1.0; tp[1] = Tred;
            for (int p = 2; p < poly_max && p < 16; ++p) {
                if (p < numP) pp[p] = 2.0 * Pred * pp[p - 1] - pp[p - 2];
                if (p < numT) tp[p] = 2.0 * Tred * tp[p - 1] - tp[p - 2];
            }
            const double* par = m.cheb_params + (size_t)c * maxT * maxP;
            double kt = 0.0;
            for (int t = 0; t < maxT; ++t) {
                double temp = 0.0;
                for (int k = 0; k < maxP; ++k)
                    if (k < numP) temp = temp + pp[k] * tb(par, t * maxP + k);
                if (t < numT) kt = kt + tp[t] * temp;
            }
            kf = exp(LN10 * fmin(EXP_MAX, kt));
        }
        // falloff / chemically activated pressure modification,
        // rate_subs.py:2251-2639 and :1510-1624
        int f = tb(m.fall_mask, r);
        if (f >= 0) {
            double kff = arrhenius(tb(m.fall_form, f), tb(m.fall_A, f), tb(m.fall_b, f),
                                   tb(m.fall_Ta, f), Tc, logT, Tinv);
            int ftype = tb(m.fall_type, f);
            double kinf = ftype ? kff : kf;
            double k0 = ftype ? kf : kff;
            int t = tb(m.fall_to_thd, f);
            double Pr = w.pmod[t] * k0 / kinf;
            int blend = tb(m.fall_blend, f);
            double Fi = 1.0;
            if (blend == 2) {
                int i = tb(m.fall_param, f);
                double a = tb(m.troe_a, i);
                double logPr = log10(fmax(SMALL, Pr));
                double Fc = a * exp(fmin(EXP_MAX, -1.0 * Tc * tb(m.troe_T1inv, i))) +
                            (1.0 + -1.0 * a) * exp(fmin(EXP_MAX, -1.0 * Tc * tb(m.troe_T3inv, i)));
                double T2 = tb(m.troe_T2, i);
                if (T2 != 0.0) Fc = Fc + exp(fmin(EXP_MAX, -1.0 * T2 / Tc));
                double logFc = log10(fmax(SMALL, Fc));
                double B = -1.1762 * logFc + -1.0 * 0.14 * logPr + 0.806;
                double A = -0.67 * logFc + logPr + -0.4;
                Fi = pow(fmax(SMALL, Fc), 1.0 / (A * A / (B * B) + 1.0));
                if (JAC) { w.Fc[f] = Fc; w.At[f] = A; w.Bt[f] = B; }
            } else if (blend == 3) {
                int i = tb(m.fall_param, f);
                double logPr = log10(fmax(SMALL, Pr));
                double X = 1.0 / (logPr * logPr + 1.0);
                double base = tb(m.sri_a, i) * exp(fmin(EXP_MAX, -tb(m.sri_b, i) * Tinv)) +
                              exp(fmin(EXP_MAX, -Tc / tb(m.sri_c, i)));
                Fi = pow(base, X);
                double d = tb(m.sri_d, i), e = tb(m.sri_e, i);
                if (d != 1.0) Fi = Fi * d;
                if (e != 0.0) Fi = Fi * pow(Tc, e);
                if (JAC) w.X[f] = X;
            }
            double ci = Fi / (1.0 + Pr);
            if (!ftype) ci = ci * Pr;
            w.pmod[t] = ci;
            if (JAC) { w.kf_fall[f] = kff; w.Pr[f] = Pr; w.Fi[f] = Fi; }
        }
        // forward / reverse rates of progress, rate_subs.py:1371-1785
        const int o0 = tb(m.rxn_off, r), o1 = tb(m.rxn_off, r + 1);
        double rf = kf;
        for (int s = o0; s < o1; ++s) rf = rf * fast_powi(w.conc[tb(m.rxn_spec, s)], tb(m.nu_reac, s));
        double net = rf;
        int ri = tb(m.rev_mask, r);
        double rr = 0.0, kr = 0.0;
        if (ri >= 0) {
            int nsum = tb(m.nu_sum, r);
            double P_val = !(nsum > 0) ? RU / PA : PA / RU;
            double P_sum = fast_powi(P_val, nsum < 0 ? -nsum : nsum);
            double B_sum = 0.0;
            for (int s = o0; s < o1; ++s) {
                int net_nu = tb(m.nu_prod, s) + -1 * tb(m.nu_reac, s);
                if (net_nu != 0) B_sum = B_sum + (double)net_nu * w.b[tb(m.rxn_spec, s)];
            }
            B_sum = exp(fmin(EXP_MAX, B_sum));
            kr = kf / (P_sum * B_sum);
            rr = kr;
            for (int s = o0; s < o1; ++s) rr = rr * fast_powi(w.conc[tb(m.rxn_spec, s)], tb(m.nu_prod, s));
            net = net + -1.0 * rr;
        }
        int ti = tb(m.thd_mask, r);
        if (ti >= 0) net = net * w.pmod[ti];
        if (JAC) { w.kf[r] = kf; w.kr[r] = kr; w.rop_f[r] = rf; w.rop_r[r] = rr; }
        // get_spec_rates, rate_subs.py:1079-1167
        for (int s = o0; s < o1; ++s) {
            int k = tb(m.rxn_spec, s);
            int nu = tb(m.nu_prod, s) + -1 * tb(m.nu_reac, s);
            w.wdot[k] = w.wdot[k] + (double)nu * net;
        }
    }
    // get_molar_rates / temperature_rate / get_extra_var_rates,
    // rate_subs.py:709-1076
    double V_val = fmax(V_MIN, st.V);
    for (int k = 0; k < Ns - 1; ++k) out(2 + k, V_val * w.wdot[k]);
    double upper = 0.0, lower = 0.0;
    for (int k = 0; k < Ns; ++k) {
        const double* a = nasa(m, k, Tc);
        double e = nasa_h(a, Tc, !conp);
        double c = nasa_cp(a, Tc, !conp);
        if (JAC) { w.energy[k] = e; w.heat[k] = c; }
        lower = lower + w.conc[k] * c;
        upper = upper + e * w.wdot[k];
    }
    double Tdot = 0.0 + -1.0 * upper / lower;
    out(0, Tdot);
    double dE = 0.0;
    for (int k = 0; k < Ns - 1; ++k) dE = dE + (1.0 + -1.0 * tb(m.mw_post, k)) * w.wdot[k];
    double Edot;
    if (conp) {
        Edot = V_val * Tdot / Tc;
        Edot = Edot + V_val * dE * Tc * RU / st.P;
    } else {
        Edot = st.P * Tdot / Tc;
        Edot = Edot + Tc * RU * dE;
    }
    out(1, Edot);
}

// --------------------------------------------------------------- Jacobian
struct JacLane {
    const DevMech& m;
    LaneWs& w;
    int conp;
    double T, P, V, Tdot, heat_tot;
    __device__ __forceinline__ double& J(int row, int col) { return w.jac[row * m.NN + col]; }
    __device__ __forceinline__ double cg(int k) { return fmax(SMALL, w.conc[k]); }
    __device__ __forceinline__ double kr_of(int r) { return tb(m.rev_mask, r) >= 0 ? w.kr[r] : 0.0; }
    __device__ __forceinline__ double ci_of(int r) {
        int t = tb(m.thd_mask, r);
        return t >= 0 ? w.pmod[t] : 1.0;
    }
    __device__ __forceinline__ double ropi(int r) {
        double v = w.rop_f[r];
        if (tb(m.rev_mask, r) >= 0) v = v + -1.0 * w.rop_r[r];
        return v;
    }
    __device__ __forceinline__ void scatter(int r, int col, double val) {
        int o1 = tb(m.rxn_off, r + 1);
        for (int s = tb(m.rxn_off, r); s < o1; ++s) {
            int k = tb(m.rxn_spec, s);
            if (k != m.ns) J(k + 2, col) += (tb(m.nu_prod, s) + -1.0 * tb(m.nu_reac, s)) * val;
        }
    }
    __device__ void ns_products(int r, double& Sf, double& Sr) {
        int o0 = tb(m.rxn_off, r), o1 = tb(m.rxn_off, r + 1);
        Sf = tb(m.nu_reac, o1 - 1);
        Sr = tb(m.nu_prod, o1 - 1);
        for (int s = o0; s < o1; ++s) {
            int nf = tb(m.nu_reac, s), nr = tb(m.nu_prod, s);
            int sp = tb(m.rxn_spec, s);
            if (sp == m.ns) { nf -= 1; nr -= 1; }
            Sf = Sf * fast_powi(cg(sp), nf);
            Sr = Sr * fast_powi(cg(sp), nr);
        }
    }
};

// dRopidnj + dRopidnj_ns, create_jacobian.py:4169-4520
__device__ void jac_dropi_dnj(JacLane& c) {
    const DevMech& m = c.m;
    for (int r = 0; r < m.Nr; ++r) {
        double kr_i = c.kr_of(r), kf_i = c.w.kf[r], ci = c.ci_of(r);
        int o0 = tb(m.rxn_off, r), o1 = tb(m.rxn_off, r + 1);
        for (int sk = o0; sk < o1; ++sk) {
            int k = tb(m.rxn_spec, sk);
            if (k == m.ns) continue;
            int nu_k = tb(m.nu_prod, sk) + -1 * tb(m.nu_reac, sk);
            for (int sj = o0; sj < o1; ++sj) {
                int j = tb(m.rxn_spec, sj);
                if (j == m.ns) continue;
                double Sr = tb(m.nu_prod, sj), Sf = tb(m.nu_reac, sj);
                for (int si = o0; si < o1; ++si) {
                    int nr = tb(m.nu_prod, si), nf = tb(m.nu_reac, si);
                    int sp = tb(m.rxn_spec, si);
                    if (sp == j) { nr -= 1; nf -= 1; }
                    double cv = c.cg(sp);
                    Sr = Sr * fast_powi(cv, nr);
                    Sf = Sf * fast_powi(cv, nf);
                }
                c.J(k + 2, j + 2) += (kf_i * Sf + -1.0 * kr_i * Sr) * ci * nu_k;
            }
        }
    }
    for (int h = 0; h < m.n_rxn_has_ns; ++h) {
        int r = tb(m.rxn_has_ns, h);
        double kr_i = c.kr_of(r), kf_i = c.w.kf[r], ci = c.ci_of(r);
        int o0 = tb(m.rxn_off, r), o1 = tb(m.rxn_off, r + 1);
        for (int sk = o0; sk < o1; ++sk) {
            int k = tb(m.rxn_spec, sk);
            if (k == m.ns) continue;
            int nu_k = tb(m.nu_prod, sk) + -1 * tb(m.nu_reac, sk);
            double Sr = 1.0, Sf = 1.0;
            for (int si = o0; si < o1; ++si) {
                int nr = tb(m.nu_prod, si), nf = tb(m.nu_reac, si);
                int sp = tb(m.rxn_spec, si);
                if (sp == m.ns) { Sr = Sr * nr; nr -= 1; }
                Sr = Sr * fast_powi(c.cg(sp), nr);
                if (sp == m.ns) { Sf = Sf * nf; nf -= 1; }
                Sf = Sf * fast_powi(c.cg(sp), nf);
            }
            double upd = (kr_i * Sr + -1.0 * kf_i * Sf) * ci * nu_k;
            for (int j = 0; j < m.Ns - 1; ++j) c.J(k + 2, j + 2) += upd;
        }
    }
}

// dPr/dn_j factor of a falloff reaction, create_jacobian.py:3780-3890
__device__ double fall_dnj_factor(JacLane& c, int f) {
    const DevMech& m = c.m;
    int r = tb(m.fall_map, f);
    double Pr = c.w.Pr[f];
    double dFi = 0.0;
    int blend = tb(m.fall_blend, f);
    if (blend == 2) {
        double A = c.w.At[f], B = c.w.Bt[f];
        dFi = A * A + B * B;
        dFi = -2.0 * A * B * (0.14 * A + B) * log(fmax(SMALL, c.w.Fc[f])) /
              (fmax(SMALL, Pr) * dFi * dFi * LN10);
    } else if (blend == 3) {
        int i = tb(m.fall_param, f);
        double X = c.w.X[f];
        double T = c.T;
        double lv = log(fmax(SMALL, tb(m.sri_a, i) * exp(fmin(EXP_MAX, -tb(m.sri_b, i) / T)) +
                                        exp(fmin(EXP_MAX, -T / tb(m.sri_c, i)))));
        dFi = -2.0 * X * X * lv * log(fmax(SMALL, Pr)) / (fmax(SMALL, Pr) * LN10SQ);
    }
    double Fi_fac = dFi, k0, kinf;
    if (tb(m.fall_type, f)) {
        k0 = c.w.kf[r];
        kinf = c.w.kf_fall[f];
    } else {
        kinf = c.w.kf[r];
        k0 = c.w.kf_fall[f];
        Fi_fac = Pr * Fi_fac + 1.0;
    }
    return k0 * (c.w.Fi[f] * Fi_fac + -1.0 * c.w.pmod[tb(m.fall_to_thd, f)]) / (kinf * (Pr + 1.0));
}

__device__ __forceinline__ bool thd_has_ns(const DevMech& m, int t) {
    int e = tb(m.thd_off, t + 1);
    return e > tb(m.thd_off, t) && tb(m.thd_spec, e - 1) == m.ns;
}

// dci_thd_dnj, dci_{lind,sri,troe}_dnj and their _ns kernels,
// create_jacobian.py:3558-4166.  item = third-body index (thd-only) or
// falloff index.
__device__ void jac_dci_dnj_one(JacLane& c, int t, int f, bool ns_pass) {
    const DevMech& m = c.m;
    int r = tb(m.thd_map, t);
    double ropi = c.ropi(r);
    double fac = f >= 0 ? fall_dnj_factor(c, f) : 1.0;
    int tt = tb(m.thd_type, t);
    int o0 = tb(m.rxn_off, r), o1 = tb(m.rxn_off, r + 1);
    if (!ns_pass) {
        int e = tb(m.thd_off, t + 1);
        for (int sj = tb(m.thd_off, t); sj < e; ++sj) {
            int j = tb(m.thd_spec, sj);
            if (j == m.ns) continue;
            double dci = 0.0;
            if (tt == 2) dci = 1.0;
            if (tt == 1) dci = tb(m.thd_eff, sj) + -1.0;
            for (int sk = o0; sk < o1; ++sk) {
                int k = tb(m.rxn_spec, sk);
                if (k == m.ns) continue;
                int nu_k = tb(m.nu_prod, sk) + -1 * tb(m.nu_reac, sk);
                if (f < 0) c.J(k + 2, j + 2) += nu_k * dci * ropi;
                else c.J(k + 2, j + 2) += nu_k * dci * ropi * fac;
            }
        }
    } else {
        double ns_eff = tb(m.thd_eff, tb(m.thd_off, t + 1) - 1);
        double dci = 0.0;
        if (tt == 2) dci = -1.0;
        if (tt == 1) dci = 1.0 + -1.0 * ns_eff;
        for (int sk = o0; sk < o1; ++sk) {
            int k = tb(m.rxn_spec, sk);
            if (k == m.ns) continue;
            int nu_k = tb(m.nu_prod, sk) + -1 * tb(m.nu_reac, sk);
            double v = f < 0 ? nu_k * dci * ropi : nu_k * dci * ropi * fac;
            for (int j = 0; j < m.Ns - 1; ++j) c.J(k + 2, j + 2) += v;
        }
    }
}

__device__ void jac_dci_dnj(JacLane& c) {
    const DevMech& m = c.m;
    for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < m.n_thd_only; ++i) {
            int t = tb(m.thd_only, i);
            if (pass == 1 && !thd_has_ns(m, t)) continue;
            jac_dci_dnj_one(c, t, -1, pass == 1);
        }
    // Lindemann, SRI, Troe (fall_order), each followed by its _ns kernel
    int blend_prev = -1, start = 0;
    for (int i = 0; i <= m.n_fall; ++i) {
        int blend = i < m.n_fall ? tb(m.fall_blend, tb(m.fall_order, i)) : -2;
        if (blend != blend_prev && blend_prev != -1) {
            for (int pass = 0; pass < 2; ++pass)
                for (int q = start; q < i; ++q) {
                    int f = tb(m.fall_order, q);
                    int t = tb(m.fall_to_thd, f);
                    if (pass == 1 && !thd_has_ns(m, t)) continue;
                    jac_dci_dnj_one(c, t, f, pass == 1);
                }
            start = i;
        }
        blend_prev = blend;
    }
}

// dTdot_dnj and dEdot_dnj, create_jacobian.py:3266-3484
__device__ void jac_rows_TE_dnj(JacLane& c) {
    const DevMech& m = c.m;
    const int ns = m.ns;
    const double fixed = c.conp ? c.P : c.V;
    const double extra = c.conp ? c.V : c.P;
    const double Tinv = 1.0 / c.T;
    const double e_ns = c.w.energy[ns], h_ns = c.w.heat[ns];
    for (int j = 0; j < m.Ns - 1; ++j) {
        double sum = 0.0, sum2 = 0.0;
        for (int k = 0; k < m.Ns - 1; ++k) {
            double jv = c.J(k + 2, j + 2);
            sum = sum + (c.w.energy[k] + -1.0 * e_ns * tb(m.mw_post, k)) * jv;
            sum2 = sum2 + (1.0 + -1.0 * tb(m.mw_post, k)) * jv;
        }
        double j0 = -1.0 * (sum + c.Tdot * (c.w.heat[j] + -1.0 * h_ns)) / (c.V * c.heat_tot);
        c.J(0, j + 2) = j0;
        c.J(1, j + 2) = c.J(1, j + 2) + c.T * RU * sum2 / fixed + extra * j0 * Tinv;
    }
}

// d(rop)/dT core of one reaction, create_jacobian.py:2978-2992
__device__ double dropi_dT_core(JacLane& c, int r, double dkf) {
    const DevMech& m = c.m;
    double d = c.w.rop_f[r] * dkf;
    if (tb(m.rev_mask, r) >= 0) {
        double dB = 0.0;
        int o1 = tb(m.rxn_off, r + 1);
        for (int s = tb(m.rxn_off, r); s < o1; ++s)
            dB = dB + (tb(m.nu_prod, s) + -1.0 * tb(m.nu_reac, s)) * c.w.db[tb(m.rxn_spec, s)];
        d = d + -1.0 * c.w.rop_r[r] * (dkf + -1.0 * dB);
    }
    return d * c.ci_of(r) * c.V;
}

__device__ __forceinline__ double plog_par(const DevMech& m, int q, int p, int k) {
    return tb(m.plog_params, ((size_t)q * m.n_plog + p) * m.plog_max_p + k);
}
__device__ __forceinline__ void plog_interval(const DevMech& m, int p, double logP, int& lo,
                                              int& hi, int& numP) {
    numP = tb(m.plog_np, p) - 1;
    lo = 0;
    hi = numP;
    for (int k = 0; k < m.plog_max_p - 1; ++k)
        if (k <= numP && logP > plog_par(m, 0, p, k) && logP <= plog_par(m, 0, p, k + 1)) {
            lo = k;
            hi = k + 1;
        }
}

// falloff / third-body loop helper: visits thd-only reactions, then
// Lindemann, SRI, Troe falloff reactions (the reference's kernel order)
template <class F>
__device__ void for_each_pmod(const DevMech& m, F&& fn) {
    for (int i = 0; i < m.n_thd_only; ++i) fn(tb(m.thd_only, i), -1);
    for (int i = 0; i < m.n_fall; ++i) {
        int f = tb(m.fall_order, i);
        fn(tb(m.fall_to_thd, f), f);
    }
}

// temperature column, create_jacobian.py:1847-3230
__device__ void jac_column_T(JacLane& c) {
    const DevMech& m = c.m;
    const double T = c.T, Tinv = 1.0 / T;
    for (int r = 0; r < m.Nr; ++r) {
        int s = tb(m.simple_mask, r);
        if (s < 0) continue;
        double dkf = (tb(m.simple_b, s) + tb(m.simple_Ta, s) * Tinv) * Tinv;
        c.scatter(r, 0, dropi_dT_core(c, r, dkf));
    }
    for (int h = 0; h < m.n_rxn_has_ns; ++h) {
        int r = tb(m.rxn_has_ns, h);
        double Sf, Sr;
        c.ns_products(r, Sf, Sr);
        double d = (Sr * c.kr_of(r) + -1.0 * Sf * c.w.kf[r]) * c.V * c.ci_of(r) * c.P / (RU * T * T);
        c.scatter(r, 0, d);
    }
    const double logP = log(fmax(SMALL, c.P));
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
        int r = tb(m.plog_map, p);
        c.scatter(r, 0, dropi_dT_core(c, r, dkf));
    }
    if (m.n_cheb) {
        const int maxT = m.cheb_max_t, maxP = m.cheb_max_p;
        const int poly_max = maxP > maxT - 1 ? maxP : maxT - 1;
        for (int q = 0; q < m.n_cheb; ++q) {
            double Pmin = tb(m.cheb_plim, 2 * q), Pmax = tb(m.cheb_plim, 2 * q + 1);
            double Tmin = tb(m.cheb_tlim, 2 * q), Tmax = tb(m.cheb_tlim, 2 * q + 1);
            int numP = tb(m.cheb_np, q), numT = tb(m.cheb_nt, q) - 1;
            double Tred = (2 * Tinv - Tmax - Tmin) / (Tmax - Tmin);
            double Pred = (2 * logP - Pmax - Pmin) / (Pmax - Pmin);
            double pp[16], tp[16];
            pp[0] = 1; pp[1] = Pred; tp[0] = 1; tp[1] = 2 * Tred;
            for (int p = 2; p < poly_max && p < 16; ++p) {
                if (p < numP) pp[p] = 2 * Pred * pp[p - 1] - pp[p - 2];
                if (p < numT) tp[p] = 2 * Tred * tp[p - 1] - tp[p - 2];
            }
            const double* par = m.cheb_params + (size_t)q * maxT * maxP;
            double dkf = 0;
            for (int t = 0; t < maxT - 1; ++t) {
                double temp = 0;
                for (int k = 0; k < maxP; ++k)
                    if (k < numP) temp = temp + pp[k] * tb(par, (t + 1) * maxP + k);
                if (t < numT) dkf = dkf + (t + 1) * tp[t] * temp;
            }
            dkf = -dkf * 2 * LN10 * Tinv * Tinv / (Tmax - Tmin);
            int r = tb(m.cheb_map, q);
            c.scatter(r, 0, dropi_dT_core(c, r, dkf));
        }
    }
    // dci_thd_dT / dci_{lind,sri,troe}_dT, create_jacobian.py:2117-2588
    for_each_pmod(m, [&](int t, int f) {
        int r = tb(m.thd_map, t);
        int e = tb(m.thd_off, t + 1) - 1;
        int tt = tb(m.thd_type, t);
        double mod = 1.0;
        if (tt == 2) mod = tb(m.thd_spec, e) == m.ns;
        if (tt == 1 && tb(m.thd_spec, e) == m.ns) mod = tb(m.thd_eff, e);
        double rop_net = c.ropi(r);
        if (f < 0) {
            c.scatter(r, 0, -1.0 * c.P * mod * RU_INV * Tinv * Tinv * c.V * rop_net);
            return;
        }
        double thd_fac = -1.0 * c.P * mod * RU_INV * Tinv * Tinv;
        int s = tb(m.simple_mask, r);
        double kf_0, kf_inf, beta_0, beta_inf, Ta_0, Ta_inf;
        if (tb(m.fall_type, f)) {
            kf_0 = c.w.kf[r]; beta_0 = tb(m.simple_b, s); Ta_0 = tb(m.simple_Ta, s);
            kf_inf = c.w.kf_fall[f]; beta_inf = tb(m.fall_b, f); Ta_inf = tb(m.fall_Ta, f);
        } else {
            kf_0 = c.w.kf_fall[f]; beta_0 = tb(m.fall_b, f); Ta_0 = tb(m.fall_Ta, f);
            kf_inf = c.w.kf[r]; beta_inf = tb(m.simple_b, s); Ta_inf = tb(m.simple_Ta, s);
        }
        double pmod = c.w.pmod[t], Pr = c.w.Pr[f];
        double theta_Pr = Tinv * (beta_0 - beta_inf + (Ta_0 - Ta_inf) * Tinv);
        double theta_no_Pr = thd_fac * kf_0 / kf_inf;
        double dFi = 0.0;
        int blend = tb(m.fall_blend, f);
        if (blend == 2) {
            int i = tb(m.fall_param, f);
            double a = tb(m.troe_a, i), T1 = tb(m.troe_T1inv, i), T2 = tb(m.troe_T2, i),
                   T3 = tb(m.troe_T3inv, i);
            double A = c.w.At[f], B = c.w.Bt[f], Fc = c.w.Fc[f];
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
            int i = tb(m.fall_param, f);
            double a = tb(m.sri_a, i), b = tb(m.sri_b, i), e_ = tb(m.sri_e, i);
            double X = c.w.X[f];
            double cinv = 1.0 / tb(m.sri_c, i);
            double ec = exp(fmin(EXP_MAX, -T * cinv));
            double eb = exp(fmin(EXP_MAX, -b * Tinv));
            double lv = log(fmax(SMALL, a * eb + ec));
            dFi = -X * (ec * cinv - a * b * Tinv * Tinv * eb) / (a * eb + ec) + e_ * Tinv -
                  2 * X * X * lv * (Pr * theta_Pr + theta_no_Pr) * log(fmax(SMALL, Pr)) /
                      (fmax(SMALL, Pr) * LN10SQ);
        }
        double d = pmod * (-1.0 * (Pr * theta_Pr + theta_no_Pr) / (Pr + 1.0) + dFi);
        if (!tb(m.fall_type, f)) d = d + theta_Pr * pmod + c.w.Fi[f] * theta_no_Pr / (Pr + 1.0);
        d = d * c.V * rop_net;
        c.scatter(r, 0, d);
    });
    // dTdot_dT and dEdotdT, create_jacobian.py:1847-2114
    const int ns = m.ns;
    double rate_sum = 0.0;
    double dTsum = (c.w.heat[ns] * Tinv + -1.0 * c.w.dheat[ns]) * c.w.conc[ns];
    double Vinv = 1.0 / c.V;
    for (int i = 0; i < m.Ns - 1; ++i) {
        rate_sum = rate_sum + Vinv * c.J(i + 2, 0) *
                                  (-1.0 * c.w.energy[i] + c.w.energy[ns] * tb(m.mw_post, i));
        rate_sum = rate_sum + c.w.wdot[i] * (-1.0 * c.w.heat[i] + tb(m.mw_post, i) * c.w.heat[ns]);
        dTsum = dTsum + (c.w.heat[ns] * Tinv + -1.0 * c.w.dheat[i]) * c.w.conc[i];
    }
    double j00 = c.J(0, 0) + (c.Tdot * dTsum + rate_sum) / c.heat_tot;
    c.J(0, 0) = j00;
    double sum = 0.0;
    if (c.conp) {
        for (int i = 0; i < m.Ns - 1; ++i)
            sum = sum + (1.0 + -1.0 * tb(m.mw_post, i)) * (Vinv * c.J(i + 2, 0) + Tinv * c.w.wdot[i]);
        double j10 = c.J(1, 0) + c.V * Tinv * (j00 + -1.0 * Tinv * c.Tdot);
        c.J(1, 0) = j10 + RU * T * c.V * sum / c.P;
    } else {
        double fac = T / c.V;
        for (int i = 0; i < m.Ns - 1; ++i)
            sum = sum + (1.0 - tb(m.mw_post, i)) * (c.J(i + 2, 0) * fac + c.w.wdot[i]);
        double j10 = c.J(1, 0) + c.P * (j00 - c.Tdot * Tinv) * Tinv;
        c.J(1, 0) = j10 + RU * sum;
    }
}

// extra-variable column, create_jacobian.py:362-1844
__device__ void jac_column_E(JacLane& c) {
    const DevMech& m = c.m;
    const double T = c.T, Tinv = 1.0 / T;
    const int ns = m.ns;
    if (c.conp) {
        auto conp_term = [&](int r, bool with_pmod) {
            int nu_f = -1, nu_r = -1;
            int o1 = tb(m.rxn_off, r + 1);
            for (int s = tb(m.rxn_off, r); s < o1; ++s) {
                nu_f += tb(m.nu_reac, s);
                nu_r += tb(m.nu_prod, s);
            }
            double d = -1.0 * nu_f * c.w.rop_f[r];
            if (tb(m.rev_mask, r) >= 0) d = d + nu_r * c.w.rop_r[r];
            int ti = tb(m.thd_mask, r);
            if (with_pmod && ti >= 0) d = d * c.w.pmod[ti];
            c.scatter(r, 1, d);
        };
        for (int r = 0; r < m.Nr; ++r)
            if (tb(m.simple_mask, r) >= 0) conp_term(r, true);
        for (int h = 0; h < m.n_rxn_has_ns; ++h) {
            int r = tb(m.rxn_has_ns, h);
            double Sf, Sr;
            c.ns_products(r, Sf, Sr);
            double fac = c.P / (RU * T);
            c.scatter(r, 1, (Sf * c.w.kf[r] - Sr * c.kr_of(r)) * c.ci_of(r) * fac);
        }
        for (int p = 0; p < m.n_plog; ++p) conp_term(tb(m.plog_map, p), false);
        for (int q = 0; q < m.n_cheb; ++q) conp_term(tb(m.cheb_map, q), false);
    } else {
        for (int h = 0; h < m.n_rxn_has_ns; ++h) {
            int r = tb(m.rxn_has_ns, h);
            double Sf, Sr;
            c.ns_products(r, Sf, Sr);
            double fac = c.V / (RU * T);
            c.scatter(r, 1, (Sf * c.w.kf[r] - Sr * c.kr_of(r)) * c.ci_of(r) * fac);
        }
        const double logP = log(fmax(SMALL, c.P));
        const double logT = log(fmax(SMALL, T));
        for (int p = 0; p < m.n_plog; ++p) {
            int lo, hi, numP;
            plog_interval(m, p, logP, lo, hi, numP);
            double dkf = 0;
            if (logP > plog_par(m, 0, p, 0) && logP <= plog_par(m, 0, p, numP))
                dkf = (plog_par(m, 1, p, hi) - plog_par(m, 1, p, lo) +
                       logT * (plog_par(m, 2, p, hi) - plog_par(m, 2, p, lo)) -
                       (plog_par(m, 3, p, hi) - plog_par(m, 3, p, lo)) * Tinv) /
                      (c.P * (plog_par(m, 0, p, hi) - plog_par(m, 0, p, lo)));
            int r = tb(m.plog_map, p);
            double d = c.w.rop_f[r];
            if (tb(m.rev_mask, r) >= 0) d = d - c.w.rop_r[r];
            c.scatter(r, 1, d * dkf * c.V);
        }
        if (m.n_cheb) {
            const int maxT = m.cheb_max_t, maxP = m.cheb_max_p;
            const int poly_max = maxP - 1 > maxT ? maxP - 1 : maxT;
            for (int q = 0; q < m.n_cheb; ++q) {
                double Pmin = tb(m.cheb_plim, 2 * q), Pmax = tb(m.cheb_plim, 2 * q + 1);
                double Tmin = tb(m.cheb_tlim, 2 * q), Tmax = tb(m.cheb_tlim, 2 * q + 1);
                int numP = tb(m.cheb_np, q) - 1, numT = tb(m.cheb_nt, q);
                double Tred = (2 * Tinv - Tmax - Tmin) / (Tmax - Tmin);
                double Pred = (2 * logP - Pmax - Pmin) / (Pmax - Pmin);
                double pp[16], tp[16];
                pp[0] = 1; pp[1] = 2 * Pred; tp[0] = 1; tp[1] = Tred;
                for (int p = 2; p < poly_max && p < 16; ++p) {
                    if (p < numP) pp[p] = 2 * Pred * pp[p - 1] - pp[p - 2];
                    if (p < numT) tp[p] = 2 * Tred * tp[p - 1] - tp[p - 2];
                }
                const double* par = m.cheb_params + (size_t)q * maxT * maxP;
                double dkf = 0;
                for (int t = 0; t < maxT; ++t) {
                    double temp = 0;
                    for (int k = 0; k < maxP - 1; ++k)
                        if (k < numP) temp = temp + (k + 1) * pp[k] * tb(par, t * maxP + k + 1);
                    if (t < numT) dkf = dkf + tp[t] * temp;
                }
                dkf = dkf * 2 * LN10 / (c.P * (Pmax - Pmin));
                int r = tb(m.cheb_map, q);
                double d = c.w.rop_f[r];
                if (tb(m.rev_mask, r) >= 0) d = d - c.w.rop_r[r];
                c.scatter(r, 1, d * dkf * c.V);
            }
        }
    }
    // dci_thd_dE / dci_{lind,sri,troe}_dE, create_jacobian.py:362-882
    const double rt_inv = 1.0 / (RU * T);
    for_each_pmod(m, [&](int t, int f) {
        int r = tb(m.thd_map, t);
        int e = tb(m.thd_off, t + 1) - 1;
        int tt = tb(m.thd_type, t);
        double rop_net = c.ropi(r);
        double mod = c.conp ? (double)(tt == 1) : 1.0;
        if (tt == 1 && tb(m.thd_spec, e) == ns) mod = tb(m.thd_eff, e);
        if (tt == 2) mod = tb(m.thd_spec, e) == ns;
        if (f < 0) {
            if (c.conp) {
                if (tt != 3) mod = mod * c.P * rt_inv - c.w.pmod[t];
                c.scatter(r, 1, mod * rop_net);
            } else {
                c.scatter(r, 1, mod * rt_inv * c.V * rop_net);
            }
            return;
        }
        int not_unity = tt != 3;
        double kf_0, kf_inf;
        if (tb(m.fall_type, f)) { kf_0 = c.w.kf[r]; kf_inf = c.w.kf_fall[f]; }
        else { kf_0 = c.w.kf_fall[f]; kf_inf = c.w.kf[r]; }
        if (c.conp) mod = mod * c.P * rt_inv * kf_0 / kf_inf;
        else mod = mod * rt_inv * kf_0 / kf_inf;
        double Pr = c.w.Pr[f];
        double theta = c.conp ? (mod + -1.0 * Pr * not_unity) : mod;
        double dFi = 0.0;
        int blend = tb(m.fall_blend, f);
        if (blend == 2) {
            double A = c.w.At[f], B = c.w.Bt[f];
            double absqsq = A * A + B * B;
            absqsq = absqsq * absqsq;
            dFi = -2.0 * A * B * log(fmax(SMALL, c.w.Fc[f])) * (0.14 * A + B) * theta /
                  (fmax(SMALL, Pr) * absqsq * LN10);
        } else if (blend == 3) {
            int i = tb(m.fall_param, f);
            double a = tb(m.sri_a, i), b = tb(m.sri_b, i), cinv = 1.0 / tb(m.sri_c, i);
            double X = c.w.X[f];
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
        double pm = c.w.pmod[t];
        double num = c.conp ? (-mod + Pr * not_unity) : -mod;
        double d = pm * (num / (Pr + 1) + dFi);
        if (!tb(m.fall_type, f)) {
            if (c.conp) d = d + c.w.Fi[f] * mod / (Pr + 1) + -1.0 * not_unity * pm;
            else d = d + c.w.Fi[f] * mod / (Pr + 1);
        }
        d = c.conp ? d * rop_net : d * c.V * rop_net;
        c.scatter(r, 1, d);
    });
    // dTdotdE and dEdotdE, create_jacobian.py:1595-1844
    double j01;
    if (c.conp) {
        double dTsum = 0.0, specsum = 0.0;
        double spec_inv = 1.0 / (c.heat_tot * c.V);
        for (int i = 0; i < m.Ns - 1; ++i) {
            dTsum = dTsum + (c.w.heat[i] + -1.0 * c.w.heat[ns]) * c.w.conc[i];
            specsum = specsum + (c.w.energy[i] + -1.0 * c.w.energy[ns] * tb(m.mw_post, i)) *
                                    (c.J(i + 2, 1) + -1.0 * c.w.wdot[i]);
        }
        j01 = c.J(0, 1) + (c.Tdot * dTsum + -1.0 * specsum) * spec_inv;
    } else {
        double sum = 0.0;
        double Vinv = 1.0 / c.V;
        for (int i = 0; i < m.Ns - 1; ++i)
            sum = sum + (c.w.energy[i] - c.w.energy[ns] * tb(m.mw_post, i)) * c.J(i + 2, 1);
        double spec_inv = 1.0 / c.heat_tot;
        j01 = c.J(0, 1) - (c.Tdot * c.w.heat[ns] / (RU * T)) * spec_inv;
        j01 = j01 - (sum * Vinv) * spec_inv;
    }
    c.J(0, 1) = j01;
    double sum = 0.0;
    for (int i = 0; i < m.Ns - 1; ++i) sum = sum + (1 - tb(m.mw_post, i)) * c.J(i + 2, 1);
    double var = c.conp ? c.V : c.P;
    double par = c.conp ? c.P : c.V;
    double j11 = c.J(1, 1) + (var * j01 + c.Tdot) / T;
    c.J(1, 1) = j11 + RU * T * sum / par;
}

// ------------------------------------------------------------ kernels
// state index s of element i: phi[s * sN + i * sI]
struct Layout {
    size_t sN, sI;
};

__global__ void __launch_bounds__(256)
species_rates_kernel(DevMech m, int conp, long n, const double* __restrict__ param,
                     const double* __restrict__ phi, double* __restrict__ dphi, Layout L,
                     double* __restrict__ ws, WsLayout WL) {
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const long n_waves = (long)gridDim.x * (blockDim.x >> 6);
    double* base = ws + (size_t)wave * WL.total * 64 + lane;
    LaneWs w(base, WL);
    for (long tile = wave; tile * 64 < n; tile += n_waves) {
        const long s = tile * 64 + lane;
        if (s >= n) continue;
        auto ph = [&](int i) { return phi[s * L.sN + i * L.sI]; };
        auto out = [&](int i, double v) { dphi[s * L.sN + i * L.sI] = v; };
        StateIn st;
        species_rates_lane<false>(m, conp, param[s], ph, out, w, st);
    }
}

__global__ void __launch_bounds__(256)
jacobian_kernel(DevMech m, int conp, long n, const double* __restrict__ param,
                const double* __restrict__ phi, double* __restrict__ jac, Layout L, int out_C,
                double* __restrict__ ws, WsLayout WL) {
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const long n_waves = (long)gridDim.x * (blockDim.x >> 6);
    double* base = ws + (size_t)wave * WL.total * 64 + lane;
    LaneWs w(base, WL);
    const int NN = m.NN;
    const int nout = m.jac_nnz;  // number of stored values per state
    for (long tile = wave; tile * 64 < n; tile += n_waves) {
        const long s = tile * 64 + lane;
        if (s >= n) continue;
        auto ph = [&](int i) { return phi[s * L.sN + i * L.sI]; };
        double Tdot = 0.0;
        auto out = [&](int i, double v) { if (i == 0) Tdot = v; };
        StateIn st;
        species_rates_lane<true>(m, conp, param[s], ph, out, w, st);
        for (int e = 0; e < NN * NN; ++e) w.jac[e] = 0.0;
        JacLane c{m, w, conp, st.T, st.P, st.V, Tdot, 0.0};
        jac_dropi_dnj(c);
        jac_dci_dnj(c);
        // total_specific_energy and dcp/dcv, db, create_jacobian.py:3234-3555
        {
            double spec_tot = 0.0;
            for (int k = 0; k < m.Ns; ++k) spec_tot = spec_tot + w.heat[k] * w.conc[k];
            c.heat_tot = 0.0 + spec_tot;
            double Tc = clampT(c.T), Tinv = 1.0 / Tc;
            for (int k = 0; k < m.Ns; ++k) {
                const double* a = nasa(m, k, Tc);
                w.dheat[k] = nasa_dcp(a, Tc);
                if (m.n_rev) w.db[k] = nasa_db(a, Tc, Tinv);
            }
        }
        jac_rows_TE_dnj(c);
        jac_column_T(c);
        jac_column_E(c);
        // write-out: each stored value once, state-fastest (F) or
        // value-fastest (C) order
        for (int e = 0; e < nout; ++e) {
            int de = tb(m.jac_pos, e);  // dense index of stored value e
            double v = w.jac[de];
            if (out_C) jac[s * (size_t)nout + e] = v;
            else jac[(size_t)e * n + s] = v;
        }
    }
}

__global__ void __launch_bounds__(256)
chem_utils_kernel(DevMech m, int conp, long n, const double* __restrict__ phi, Layout L,
                  Layout Lo, double* __restrict__ energy, double* __restrict__ heat,
                  double* __restrict__ bout) {
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    double Tc = clampT(phi[s * L.sN]);
    double logT = log(Tc), Tinv = 1.0 / Tc;
    for (int k = 0; k < m.Ns; ++k) {
        const double* a = nasa(m, k, Tc);
        size_t o = s * Lo.sN + k * Lo.sI;
        energy[o] = nasa_h(a, Tc, !conp);
        heat[o] = nasa_cp(a, Tc, !conp);
        bout[o] = nasa_b(a, Tc, logT, Tinv);
    }
}

}  // namespace pjk

// ---------------------------------------------------------- host launchers
namespace pjk_host {

size_t ws_doubles_per_state(const DevMech& m, bool jac) {
    return (size_t)pjk::ws_layout(m.Ns, m.Nr, m.n_fall, jac).total;
}

hipError_t launch_species_rates(const DevMech& m, int conp, char order, long n,
                                const double* param, const double* phi, double* dphi,
                                double* ws, int grid, int block, hipStream_t stream) {
    pjk::Layout L = order == 'F' ? pjk::Layout{1, (size_t)n} : pjk::Layout{(size_t)m.NN, 1};
    pjk::WsLayout WL = pjk::ws_layout(m.Ns, m.Nr, m.n_fall, false);
    hipLaunchKernelGGL(pjk::species_rates_kernel, dim3(grid), dim3(block), 0, stream, m, conp, n,
                       param, phi, dphi, L, ws, WL);
    return hipGetLastError();
}

hipError_t launch_jacobian(const DevMech& m, int conp, char order, long n, const double* param,
                           const double* phi, double* jac, double* ws, int grid, int block,
                           hipStream_t stream) {
    pjk::Layout L = order == 'F' ? pjk::Layout{1, (size_t)n} : pjk::Layout{(size_t)m.NN, 1};
    pjk::WsLayout WL = pjk::ws_layout(m.Ns, m.Nr, m.n_fall, true);
    hipLaunchKernelGGL(pjk::jacobian_kernel, dim3(grid), dim3(block), 0, stream, m, conp, n,
                       param, phi, jac, L, (int)(order == 'C'), ws, WL);
    return hipGetLastError();
}

hipError_t launch_chem_utils(const DevMech& m, int conp, char order, long n, const double* phi,
                             double* energy, double* heat, double* b, hipStream_t stream) {
    pjk::Layout L = order == 'F' ? pjk::Layout{1, (size_t)n} : pjk::Layout{(size_t)m.NN, 1};
    pjk::Layout Lo = order == 'F' ? pjk::Layout{1, (size_t)n} : pjk::Layout{(size_t)m.Ns, 1};
    int block = 256;
    int grid = (int)((n + block - 1) / block);
    hipLaunchKernelGGL(pjk::chem_utils_kernel, dim3(grid), dim3(block), 0, stream, m, conp, n,
                       phi, L, Lo, energy, heat, b);
    return hipGetLastError();
}

}  // namespace pjk_host
