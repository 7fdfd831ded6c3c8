// pyjac_block.hip -- block-per-state species-rate and Jacobian kernels (gfx950).
//
// One workgroup of WPS wavefronts (WPS = 1, 2, 4 or 8) evaluates one
// thermochemical state (one launch, a workgroup per state).  All
// arrays a state needs at random access -- concentrations, Gibbs terms,
// production rates, third-body / falloff factors and, for the Jacobian, a
// dense (Ns+1)^2 accumulator -- live in the workgroup's LDS, so HBM sees only
// the state in and the result out.
//
// A state is processed in phases separated by workgroup barriers:
//   1  concentrations and NASA-7 thermodynamics (threads over species);
//      the bath-gas concentration by difference, summed in the reference's
//      sequential order (the difference cancels when the bath gas is rare)
//   2  third-body concentrations, falloff blending and every per-third-body
//      derivative factor; PLOG / Chebyshev rate constants and their
//      temperature / pressure derivatives (threads over those reactions,
//      third-body entries ordered by kind so a wave runs one code path)
//   3  reactions (threads over reactions, records sorted on the host by
//      structure so a wave's lanes follow the same branches): rate
//      constants, rates of progress, production rates and the per-reaction
//      Jacobian contributions, scattered into LDS with ds_add_f64
//   4  third-body columns (waves over third-body reactions, lanes over
//      species columns), temperature rate
//   5  temperature / extra-variable rows (lanes over columns, waves over
//      row blocks) and the 2x2 corner
//   6  write-out of the stored Jacobian entries, in C or F order
//
// Rate constants use the specialised Arrhenius forms of the reference
// (rate_subs.py:2803-2870) evaluated branch-free; every formula cites the
// reference kernel it follows.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "device_mech.h"
#include "kernels.h"

namespace pjb {

constexpr double RU = 8314.4621;
constexpr double PA = 101325.0;
constexpr double EXP_MAX = 690.775527898;
constexpr double SMALL = 1e-300;
constexpr double T_MIN = 100.0, T_MAX = 10000.0;
constexpr double V_MIN = 1e-30;
constexpr double LN10 = 2.302585092994046;
constexpr double LN10SQ = LN10 * LN10;
constexpr double RU_INV = 1.0 / RU;
constexpr int MAXSP = pjk_host::WAVE_MAXSP;

using pjk_host::BlockTables;
using pjk_host::BlockLds;
using pjk_host::BlockArgs;

// wave-uniform load through the scalar data path
template <class T>
__device__ __forceinline__ T sld(const T* p, long i) {
    return ((const __attribute__((address_space(4))) T*)p)[i];
}

// per-lane reads of the read-only tables go through the global address
// space: a flat load also counts as an LDS operation, so waiting for its data
// would wait for every LDS operation in flight, and the reverse
template <class T>
using gptr = const __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr<T> gp(const T* p) {
    return (gptr<T>)p;
}

// the same for the packed int / double records (through clang's vector
// types, which load from the global address space where HIP's classes do not)
typedef int gint2 __attribute__((ext_vector_type(2)));
typedef int gint4 __attribute__((ext_vector_type(4)));
typedef double gdouble4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ int2 ld_int2(const void* p, long i) {
    const gint2 v = gp((const gint2*)p)[i];
    return make_int2(v.x, v.y);
}
__device__ __forceinline__ int4 ld_int4(const void* p, long i) {
    const gint4 v = gp((const gint4*)p)[i];
    return make_int4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ double4 ld_double4(const void* p, long i) {
    const gdouble4 v = gp((const gdouble4*)p)[i];
    return make_double4(v.x, v.y, v.z, v.w);
}

// fast_powi of the reference (loopy_utils/preambles_and_manglers.py:123-174):
// a left-to-right product chain, here without data-dependent branches for
// exponents up to 3
__device__ __forceinline__ double powi(double val, int p) {
    double v = p < 0 ? 1.0 / val : val;
    int e = p < 0 ? -p : p;
    double r = e >= 1 ? v : 1.0;
    r = e >= 2 ? r * v : r;
    r = e >= 3 ? r * v : r;
    for (int i = 3; i < e; ++i) r *= v;
    return r;
}

__device__ __forceinline__ double clampT(double T) { return fmin(T_MAX, fmax(T_MIN, T)); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) v += __shfl_xor(v, s);
    return v;
}

// Arrhenius forms, rate_subs.py:2803-2870 (A holds log A for forms 2-4)
__device__ __forceinline__ double arrhenius(int form, double A, double b, double Ta, double Tc,
                                            double logT, double Tinv) {
    switch (form) {
        case 0: return A;
        case 1: return A * powi(Tc, (int)b);
        case 3: return exp(fmin(EXP_MAX, A + logT * b));
        case 4: return exp(fmin(EXP_MAX, A - Ta * Tinv));
        default: return exp(fmin(EXP_MAX, A + logT * b + -1.0 * Ta * Tinv));
    }
}

// Per-reaction record (runtime.cpp builds them, sorted by structure):
//   rec_i.x = kind | (nu_sum + 128) << 8   kind: 0-4 Arrhenius form,
//                                           5 PLOG, 6 Chebyshev
//   rec_i.y = 1 if reversible, rec_i.z = third-body index or -1,
//   rec_i.w = number of species slots (<= MAXSP)
//   rec_d   = A (log A for forms 2-4), b, Ta, PLOG/Chebyshev rate index
//   rec_s[8 r + q] = species | nu_reac << 16 | nu_prod << 24, padded with
//                    species Ns (concentration 1, Gibbs term 0); slot 6
//                    holds a species of the reaction with a Jacobian row
//                    (the target of the lane's zero-valued atomics)
// Records are processed 64 at a time; grp_flags[r / 64] describes the group
// (largest species count, and which optional parts any of its reactions
// has) so that a wave skips whole blocks with uniform branches.
struct Rec {
    int kind, nsum, rev, thd, cnt, kidx, home;
    double A, b, Ta;
    int sp[MAXSP];  // packed species slots
    __device__ int spec(int q) const { return sp[q] & 0xffff; }
    __device__ int nur(int q) const { return (sp[q] >> 16) & 0xff; }
    __device__ int nup(int q) const { return (sp[q] >> 24) & 0xff; }
    __device__ int nu(int q) const { return nup(q) - nur(q); }
};

__device__ __forceinline__ Rec load_rec(const BlockTables& bt, int r) {
    Rec x;
    const int4 ri = ld_int4(bt.rec_i, r);
    const double4 d = ld_double4(bt.rec_d, r);
    const int4 s0 = ld_int4(bt.rec_s, 2 * r), s1 = ld_int4(bt.rec_s, 2 * r + 1);
    x.kind = ri.x & 0xff;
    x.nsum = (ri.x >> 8) - 128;
    x.rev = ri.y;
    x.thd = ri.z;
    x.cnt = ri.w;
    x.A = d.x;
    x.b = d.y;
    x.Ta = d.z;
    x.kidx = (int)d.w;
    x.sp[0] = s0.x; x.sp[1] = s0.y; x.sp[2] = s0.z; x.sp[3] = s0.w;
    x.sp[4] = s1.x; x.sp[5] = s1.y;
    x.home = s1.z;
    return x;
}

// per-state values shared by all threads of the workgroup
struct State {
    double T, P, V, Tc, logT, Tinv, logP;
    double Traw_inv, rt_inv, V_inv;
};

typedef __attribute__((address_space(3))) double lds_double;

__device__ __forceinline__ void lds_add(lds_double* p, double v) {
    __builtin_amdgcn_ds_atomic_fadd_f64(p, v);
}

// Everything a phase needs besides the state: table pointers (read-only
// device memory), the LDS layout and the sizes, the latter two held by
// value so that no LDS store forces them to be fetched again
struct Ctx {
    const DevMech& m;
    const BlockTables& bt;
    const BlockLds L;
    lds_double* sm;
    int conp;
    int Ns, NN, ns;
    __device__ lds_double& conc(int k) { return sm[L.conc + k]; }
    __device__ lds_double& nmol(int k) { return sm[L.nmol + k]; }
    __device__ lds_double& b(int k) { return sm[L.b + k]; }
    __device__ lds_double& db(int k) { return sm[L.db + k]; }
    __device__ lds_double& wdot(int k) { return sm[L.wdot + k]; }
    __device__ lds_double& energy(int k) { return sm[L.energy + k]; }
    __device__ lds_double& heat(int k) { return sm[L.heat + k]; }
    __device__ lds_double& dheat(int k) { return sm[L.dheat + k]; }
    __device__ lds_double& U(int k) { return sm[L.U + k]; }
    __device__ lds_double& ci(int t) { return sm[L.ci + t]; }
    __device__ lds_double& nfac(int t) { return sm[L.nfac + t]; }
    __device__ lds_double& dTf(int t) { return sm[L.dTf + t]; }
    __device__ lds_double& dEf(int t) { return sm[L.dEf + t]; }
    __device__ lds_double& Cp(int p) { return sm[L.Cp + p]; }
    __device__ lds_double& mw(int k) { return sm[L.mw + k]; }
    __device__ lds_double& ksp(int i) { return sm[L.ksp + i]; }
    __device__ lds_double& dkT(int i) { return sm[L.dkT + i]; }
    __device__ lds_double& dkP(int i) { return sm[L.dkP + i]; }
    __device__ lds_double* J() { return sm + L.jac; }
};

__device__ __forceinline__ Ctx make_ctx(const BlockArgs* A, double* smem, int conp) {
    return Ctx{A->m, A->bt, A->L, (lds_double*)smem, conp, A->m.Ns, A->m.NN, A->m.ns};
}

// ------------------------------------------------------------- phase 1
// get_concentrations (rate_subs.py:597-706) and the NASA-7 properties
// (rate_subs.py:3281-3388); `phi(i)` reads entry i of this state
// species k's energy (h or u), heat capacity (cp or cv) and its T
// derivative from the NASA-7 polynomials (rate_subs.py:3281-3388)
struct Thermo {
    double energy, heat, dheat;
};

__device__ __forceinline__ gptr<double> nasa_coeffs(const DevMech& m, double T, int k) {
    return gp((T < gp(m.T_mid)[k]) ? m.a_lo + 7 * k : m.a_hi + 7 * k);
}

__device__ __forceinline__ Thermo nasa_thermo(gptr<double> a, double T, int conp) {
    const double a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3], a4 = a[4], a5 = a[5];
    const double x = T * (T * (T * (T * (T * a4 / 5.0 + a3 / 4.0) + a2 / 3.0) + a1 / 2.0) + a0);
    const double y = T * (T * (T * (T * a4 + a3) + a2) + a1) + a0;
    Thermo h;
    h.energy = RU * (conp ? x + a5 : x - T + a5);
    h.heat = RU * (conp ? y : y - 1.0);
    h.dheat = RU * (T * (T * (4.0 * T * a4 + 3.0 * a3) + 2.0 * a2) + a1);
    return h;
}

// With LATE, the dense Jacobian kernel's layout: no moles array (the bath
// gas sums phi itself) and the energies, heat capacities and their
// derivatives are recomputed after the reaction pass (they share LDS with
// arrays that only the reaction pass reads).
template <bool JAC, int NT, bool LATE = false, class PhiF>
__device__ __forceinline__ void phase_species(Ctx& c, const State& st, int tid, PhiF phi) {
    const DevMech& m = c.m;
    const int Ns = m.Ns;
    for (int k = tid; k < Ns; k += NT) {
        if (k < Ns - 1) {
            double n = fmax(SMALL, phi(2 + k));
            if (!LATE) c.nmol(k) = n;
            c.conc(k) = n * st.V_inv;
        }
        const double T = st.Tc;
        const gptr<double> a = nasa_coeffs(m, T, k);
        const double a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3], a4 = a[4], a5 = a[5], a6 = a[6];
        c.b(k) = T * (T * (T * (T * a4 / 20.0 + a3 / 12.0) + a2 / 6.0) + a1 / 2.0) +
                 (a0 + -1.0) * st.logT + -1.0 * a0 + a6 + -1.0 * a5 * st.Tinv;
        if (!LATE) {
            const Thermo h = nasa_thermo(a, T, c.conp);
            c.energy(k) = h.energy;
            c.heat(k) = h.heat;
            if (JAC) c.dheat(k) = h.dheat;
        }
        c.wdot(k) = 0.0;
        if (JAC) {
            c.U(k) = 0.0;
            c.db(k) = T * (T * (T * a4 / 5.0 + a3 / 4.0) + a2 / 3.0) + a1 / 2.0 +
                      st.Tinv * (a0 + -1.0 + a5 * st.Tinv);
        }
    }
}

// bath-gas concentration by difference, summed in the reference's
// sequential order (rate_subs.py:665-706); every wave computes it and its
// lane 0 stores it (all waves store the same value)
__device__ __forceinline__ void bath_gas(Ctx& c, const State& st, int lane) {
    const int Ns = c.m.Ns;
    double nsum = 0.0;
    for (int k = 0; k < Ns - 1; ++k) nsum = nsum + c.nmol(k);
    if (lane == 0)
        c.conc(Ns - 1) = __dadd_rn(st.P / (RU * st.Tc), __dmul_rn(-1.0 * fmax(SMALL, nsum), st.V_inv));
}

// the same sum over the floored moles read from the state itself
template <class PhiF>
__device__ __forceinline__ void bath_gas_phi(Ctx& c, const State& st, int lane, PhiF phi) {
    const int Ns = c.m.Ns;
    double nsum = 0.0;
#pragma unroll 8
    for (int k = 0; k < Ns - 1; ++k) nsum = nsum + fmax(SMALL, phi(2 + k));
    if (lane == 0)
        c.conc(Ns - 1) = __dadd_rn(st.P / (RU * st.Tc), __dmul_rn(-1.0 * fmax(SMALL, nsum), st.V_inv));
}

// ------------------------------------------------------------- phase 2
// third-body concentration, falloff blending and the derivative factors of
// third-body entry t:
//   ci   pressure modification (get_thd_body_concs, rate_subs.py:1788-1901;
//        falloff: rate_subs.py:2251-2639, :1510-1624)
//   nfac multiplier of the third-body columns d(ci)/dn_j
//        (create_jacobian.py:3780-3890; 1 for plain third bodies)
//   dTf, dEf  d(ci)/dT, d(ci)/dE per unit rate of progress
//        (create_jacobian.py:362-882 and :2117-2588)
// The parts of a falloff entry that depend on the temperature only: the
// rate constants of both limits, the Troe centre Fcent with its logarithms
// and (Jacobian) its T derivative.  The dense Jacobian kernel evaluates them
// in phase 1, on threads the species leave idle, and the rest of the entry
// after the barrier.
struct ThdPre {
    double kf, kff, Fc, logFc, lnFc, dFcent;
    int r, f;  // the entry's reaction and falloff index
};

template <bool JAC>
__device__ __forceinline__ ThdPre thd_pre(const Ctx& c, const State& st, int t) {
    const DevMech& m = c.m;
    ThdPre p{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0, -1};
    const int r = gp(m.thd_map)[t];
    const int f = gp(m.fall_mask)[r];
    p.r = r;
    p.f = f;
    if (f < 0) return p;
    const int si = gp(m.simple_mask)[r];
    p.kf = arrhenius(gp(m.simple_form)[si], gp(m.simple_A)[si], gp(m.simple_b)[si], gp(m.simple_Ta)[si], st.Tc,
                     st.logT, st.Tinv);
    p.kff = arrhenius(gp(m.fall_form)[f], gp(m.fall_A)[f], gp(m.fall_b)[f], gp(m.fall_Ta)[f], st.Tc, st.logT,
                      st.Tinv);
    // the species-rate kernel evaluates the Troe centre where it is used
    if (JAC && gp(m.fall_blend)[f] == 2) {
        const int i = gp(m.fall_param)[f];
        const double a = gp(m.troe_a)[i];
        double Fc = a * exp(fmin(EXP_MAX, -1.0 * st.Tc * gp(m.troe_T1inv)[i])) +
                    (1.0 + -1.0 * a) * exp(fmin(EXP_MAX, -1.0 * st.Tc * gp(m.troe_T3inv)[i]));
        const double T2 = gp(m.troe_T2)[i];
        if (T2 != 0.0) Fc = Fc + exp(fmin(EXP_MAX, -1.0 * T2 / st.Tc));
        p.Fc = Fc;
        p.logFc = log10(fmax(SMALL, Fc));
        if (JAC) {
            p.lnFc = log(fmax(SMALL, Fc));
            const double T = st.T, Tinv = st.Traw_inv;
            const double T1 = gp(m.troe_T1inv)[i], T3 = gp(m.troe_T3inv)[i];
            p.dFcent = -1.0 * a * T1 * exp(fmin(EXP_MAX, -1.0 * T * T1)) +
                       (a + -1.0) * T3 * exp(fmin(EXP_MAX, -1.0 * T * T3)) +
                       T2 * Tinv * Tinv * exp(fmin(EXP_MAX, -1.0 * T2 * Tinv));
        }
    }
    return p;
}

template <bool JAC>
__device__ void thd_task(Ctx& c, const State& st, int t);

// get_pre() yields the entry's ThdPre, evaluated on the spot or kept from
// phase 1; it is called after the third-body sum, where the rate constants
// are first needed
template <bool JAC, bool EARLY = false, class PreF>
__device__ void thd_post(Ctx& c, const State& st, int t, PreF get_pre) {
    const DevMech& m = c.m;
    const int conp = c.conp;
    const int tt = gp(m.thd_type)[t];
    const int not_spec = tt != 2;
    double thd = st.P * (double)not_spec / (RU * st.T);
    const int e1 = gp(m.thd_off)[t + 1];
    int s = gp(m.thd_off)[t];
    // four efficiencies per step with their loads issued together (the
    // Jacobian kernels; the species-rate kernel has no registers to spare);
    // the sum keeps its sequential order
    for (; JAC && s + 4 <= e1; s += 4) {
        const double f0 = gp(m.thd_eff)[s], f1 = gp(m.thd_eff)[s + 1], f2 = gp(m.thd_eff)[s + 2],
                     f3 = gp(m.thd_eff)[s + 3];
        const int k0 = gp(m.thd_spec)[s], k1 = gp(m.thd_spec)[s + 1], k2 = gp(m.thd_spec)[s + 2],
                  k3 = gp(m.thd_spec)[s + 3];
        const double c0 = c.conc(k0), c1 = c.conc(k1), c2 = c.conc(k2), c3 = c.conc(k3);
        thd = thd + (f0 + -1.0 * not_spec) * c0;
        thd = thd + (f1 + -1.0 * not_spec) * c1;
        thd = thd + (f2 + -1.0 * not_spec) * c2;
        thd = thd + (f3 + -1.0 * not_spec) * c3;
    }
    for (; s < e1; ++s)
        thd = thd + (gp(m.thd_eff)[s] + -1.0 * not_spec) * c.conc(gp(m.thd_spec)[s]);
    // EARLY: the entry's indices come with the values kept from phase 1
    const int r = EARLY ? get_pre().r : gp(m.thd_map)[t];
    const int f = EARLY ? get_pre().f : gp(m.fall_mask)[r];
    const double T = st.T, Tinv = st.Traw_inv;
    const double modT = gp(c.bt.thd_mod)[t];
    const double modE = (conp && tt == 3) ? 0.0 : modT;
    if (f < 0) {
        c.ci(t) = thd;
        if (JAC) {
            c.nfac(t) = 1.0;
            c.dTf(t) = -1.0 * st.P * modT * RU_INV * Tinv * Tinv * st.V;
            if (conp) {
                double mo = modE;
                if (tt != 3) mo = mo * st.P * st.rt_inv - thd;
                c.dEf(t) = mo;
            } else {
                c.dEf(t) = modE * st.rt_inv * st.V;
            }
        }
        return;
    }
    // falloff / chemically activated
    const ThdPre pre = get_pre();
    const int si = gp(m.simple_mask)[r];
    const double kf = pre.kf, kff = pre.kff;
    const int ftype = gp(m.fall_type)[f];
    const double kinf = ftype ? kff : kf;
    const double k0 = ftype ? kf : kff;
    const double Pr = thd * k0 / kinf;
    const int blend = gp(m.fall_blend)[f];
    double Fi = 1.0, Fc = 0.0, A = 0.0, B = 0.0, X = 0.0;
    if (blend == 2) {
        const double logPr = log10(fmax(SMALL, Pr));
        double logFc;
        if constexpr (JAC) {
            Fc = pre.Fc;
            logFc = pre.logFc;
        } else {
            const int i = gp(m.fall_param)[f];
            const double a = gp(m.troe_a)[i];
            Fc = a * exp(fmin(EXP_MAX, -1.0 * st.Tc * gp(m.troe_T1inv)[i])) +
                 (1.0 + -1.0 * a) * exp(fmin(EXP_MAX, -1.0 * st.Tc * gp(m.troe_T3inv)[i]));
            const double T2 = gp(m.troe_T2)[i];
            if (T2 != 0.0) Fc = Fc + exp(fmin(EXP_MAX, -1.0 * T2 / st.Tc));
            logFc = log10(fmax(SMALL, Fc));
        }
        B = -1.1762 * logFc + -1.0 * 0.14 * logPr + 0.806;
        A = -0.67 * logFc + logPr + -0.4;
        Fi = pow(fmax(SMALL, Fc), 1.0 / (A * A / (B * B) + 1.0));
    } else if (blend == 3) {
        const int i = gp(m.fall_param)[f];
        const double logPr = log10(fmax(SMALL, Pr));
        X = 1.0 / (logPr * logPr + 1.0);
        const double base = gp(m.sri_a)[i] * exp(fmin(EXP_MAX, -gp(m.sri_b)[i] * st.Tinv)) +
                            exp(fmin(EXP_MAX, -st.Tc / gp(m.sri_c)[i]));
        Fi = pow(base, X);
        const double d = gp(m.sri_d)[i], e = gp(m.sri_e)[i];
        if (d != 1.0) Fi = Fi * d;
        if (e != 0.0) Fi = Fi * pow(st.Tc, e);
    }
    double ci = Fi / (1.0 + Pr);
    if (!ftype) ci = ci * Pr;
    c.ci(t) = ci;
    if (!JAC) return;

    // d(ci)/dn_j factor, create_jacobian.py:3780-3890
    {
        double dFi = 0.0;
        if (blend == 2) {
            dFi = A * A + B * B;
            dFi = -2.0 * A * B * (0.14 * A + B) * pre.lnFc /
                  (fmax(SMALL, Pr) * dFi * dFi * LN10);
        } else if (blend == 3) {
            const int i = gp(m.fall_param)[f];
            const double lv = log(fmax(SMALL, gp(m.sri_a)[i] * exp(fmin(EXP_MAX, -gp(m.sri_b)[i] / T)) +
                                                  exp(fmin(EXP_MAX, -T / gp(m.sri_c)[i]))));
            dFi = -2.0 * X * X * lv * log(fmax(SMALL, Pr)) / (fmax(SMALL, Pr) * LN10SQ);
        }
        double Fi_fac = dFi;
        if (!ftype) Fi_fac = Pr * Fi_fac + 1.0;
        c.nfac(t) = k0 * (Fi * Fi_fac + -1.0 * ci) / (kinf * (Pr + 1.0));
    }
    // d(ci)/dT, create_jacobian.py:362-882
    {
        const double thd_fac = -1.0 * st.P * modT * RU_INV * Tinv * Tinv;
        double beta_0, beta_inf, Ta_0, Ta_inf;
        if (ftype) {
            beta_0 = gp(m.simple_b)[si]; Ta_0 = gp(m.simple_Ta)[si];
            beta_inf = gp(m.fall_b)[f]; Ta_inf = gp(m.fall_Ta)[f];
        } else {
            beta_0 = gp(m.fall_b)[f]; Ta_0 = gp(m.fall_Ta)[f];
            beta_inf = gp(m.simple_b)[si]; Ta_inf = gp(m.simple_Ta)[si];
        }
        const double theta_Pr = Tinv * (beta_0 - beta_inf + (Ta_0 - Ta_inf) * Tinv);
        const double theta_no_Pr = thd_fac * k0 / kinf;
        double dFi = 0.0;
        if (blend == 2) {
            const double dFcent = pre.dFcent;
            const double logFcent = pre.lnFc;
            const double absq = A * A + B * B;
            const double absqsq = absq * absq;
            dFi = -1.0 * B *
                  (2.0 * A * Fc * (0.14 * A + B) * (Pr * theta_Pr + theta_no_Pr) * logFcent +
                   Pr * dFcent * (2.0 * A * (1.1762 * A + -1.0 * 0.67 * B) * logFcent +
                                  -1.0 * B * absq * LN10)) /
                  (Fc * fmax(SMALL, Pr) * absqsq * LN10);
        } else if (blend == 3) {
            const int i = gp(m.fall_param)[f];
            const double a = gp(m.sri_a)[i], b = gp(m.sri_b)[i], e_ = gp(m.sri_e)[i];
            const double cinv = 1.0 / gp(m.sri_c)[i];
            const double ec = exp(fmin(EXP_MAX, -T * cinv));
            const double eb = exp(fmin(EXP_MAX, -b * Tinv));
            const double lv = log(fmax(SMALL, a * eb + ec));
            dFi = -X * (ec * cinv - a * b * Tinv * Tinv * eb) / (a * eb + ec) + e_ * Tinv -
                  2 * X * X * lv * (Pr * theta_Pr + theta_no_Pr) * log(fmax(SMALL, Pr)) /
                      (fmax(SMALL, Pr) * LN10SQ);
        }
        double d = ci * (-1.0 * (Pr * theta_Pr + theta_no_Pr) / (Pr + 1.0) + dFi);
        if (!ftype) d = d + theta_Pr * ci + Fi * theta_no_Pr / (Pr + 1.0);
        c.dTf(t) = d * st.V;
    }
    // d(ci)/dE, create_jacobian.py:2117-2588
    {
        const int not_unity = tt != 3;
        double mod = conp ? modE * st.P * st.rt_inv * k0 / kinf : modE * st.rt_inv * k0 / kinf;
        const double theta = conp ? (mod + -1.0 * Pr * not_unity) : mod;
        double dFi = 0.0;
        if (blend == 2) {
            double absqsq = A * A + B * B;
            absqsq = absqsq * absqsq;
            dFi = -2.0 * A * B * pre.lnFc * (0.14 * A + B) * theta /
                  (fmax(SMALL, Pr) * absqsq * LN10);
        } else if (blend == 3) {
            const int i = gp(m.fall_param)[f];
            const double a = gp(m.sri_a)[i], b = gp(m.sri_b)[i], cinv = 1.0 / gp(m.sri_c)[i];
            double inner;
            if (conp) {
                const double exp_cinv = exp(fmin(EXP_MAX, T * cinv));
                const double exp_b = exp(fmin(EXP_MAX, b * Tinv));
                const double exp_bc = exp(fmin(EXP_MAX, -b * Tinv - T * cinv));
                inner = log(fmax(SMALL, (a * exp_cinv + exp_b) * exp_bc));
            } else {
                const double exp_b = exp(fmin(EXP_MAX, -b * Tinv));
                const double exp_cinv = exp(fmin(EXP_MAX, -T * cinv));
                inner = log(fmax(SMALL, a * exp_b + exp_cinv));
            }
            dFi = -2 * X * X * theta * inner * log(fmax(SMALL, Pr)) / (fmax(SMALL, Pr) * LN10SQ);
        }
        const double num = conp ? (-mod + Pr * not_unity) : -mod;
        double d = ci * (num / (Pr + 1) + dFi);
        if (!ftype) {
            if (conp) d = d + Fi * mod / (Pr + 1) + -1.0 * not_unity * ci;
            else d = d + Fi * mod / (Pr + 1);
        }
        c.dEf(t) = conp ? d : d * st.V;
    }
}

template <bool JAC>
__device__ void thd_task(Ctx& c, const State& st, int t) {
    thd_post<JAC>(c, st, t, [&]() { return thd_pre<JAC>(c, st, t); });
}

__device__ __forceinline__ double plog_par(const DevMech& m, int q, int p, int k) {
    return gp(m.plog_params)[((size_t)q * m.n_plog + p) * m.plog_max_p + k];
}

// PLOG rate constant (rate_subs.py:2084-2248) and, for the Jacobian,
// d ln k / dT (create_jacobian.py:2823-2976) and d ln k / dP
// (create_jacobian.py:1158-1300)
template <bool JAC>
__device__ void plog_task(Ctx& c, const State& st, int p) {
    const DevMech& m = c.m;
    const int mp = m.plog_max_p;
    const int numP = gp(m.plog_np)[p] - 1;
    const double logP = st.logP;
    {
        const bool lower = logP <= plog_par(m, 0, p, 0);
        const bool upper = logP > plog_par(m, 0, p, numP);
        int lo = 0, hi = 0;
        if (upper) { lo = numP; hi = numP; }
        for (int k = 0; k < mp - 1; ++k)
            if (k < numP && logP > plog_par(m, 0, p, k) && logP <= plog_par(m, 0, p, k + 1)) {
                lo = k;
                hi = k + 1;
            }
        const double logk1 = plog_par(m, 1, p, lo) + st.logT * plog_par(m, 2, p, lo) -
                             plog_par(m, 3, p, lo) * st.Tinv;
        const double logk2 = plog_par(m, 1, p, hi) + st.logT * plog_par(m, 2, p, hi) -
                             plog_par(m, 3, p, hi) * st.Tinv;
        double kt = logk1;
        if (!(lower || upper))
            kt = (-logk1 + logk2) * (logP - plog_par(m, 0, p, lo)) /
                     (plog_par(m, 0, p, hi) - plog_par(m, 0, p, lo)) +
                 kt;
        c.ksp(p) = exp(fmin(EXP_MAX, kt));
    }
    if (!JAC) return;
    const double Tinv = st.Traw_inv;
    int lo = 0, hi = numP;
    for (int k = 0; k < mp - 1; ++k)
        if (k <= numP && logP > plog_par(m, 0, p, k) && logP <= plog_par(m, 0, p, k + 1)) {
            lo = k;
            hi = k + 1;
        }
    double dkT;
    if (logP > plog_par(m, 0, p, numP))
        dkT = (plog_par(m, 2, p, hi) + plog_par(m, 3, p, hi) * Tinv) * Tinv;
    else
        dkT = (plog_par(m, 2, p, lo) + plog_par(m, 3, p, lo) * Tinv) * Tinv;
    const bool inr = logP > plog_par(m, 0, p, 0) && logP <= plog_par(m, 0, p, numP);
    if (inr)
        dkT = dkT + Tinv * (logP - plog_par(m, 0, p, lo)) *
                        (plog_par(m, 2, p, hi) - plog_par(m, 2, p, lo) +
                         (plog_par(m, 3, p, hi) - plog_par(m, 3, p, lo)) * Tinv) /
                        (plog_par(m, 0, p, hi) - plog_par(m, 0, p, lo));
    double dkP = 0.0;
    if (!c.conp && inr) {
        const double logT = log(fmax(SMALL, st.T));
        dkP = (plog_par(m, 1, p, hi) - plog_par(m, 1, p, lo) +
               logT * (plog_par(m, 2, p, hi) - plog_par(m, 2, p, lo)) -
               (plog_par(m, 3, p, hi) - plog_par(m, 3, p, lo)) * Tinv) /
              (st.P * (plog_par(m, 0, p, hi) - plog_par(m, 0, p, lo)));
    }
    c.dkT(p) = dkT;
    c.dkP(p) = dkP;
}

// Chebyshev polynomials by the recurrence T_p = 2 x T_{p-1} - T_{p-2}
// (first = x: first kind, first = 2x: second kind)
struct ChebPoly {
    double x, pm2, pm1;
    int p;
    __device__ ChebPoly(double x_, double first) : x(x_), pm2(1.0), pm1(first), p(0) {}
    __device__ double next() {
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

// Chebyshev rate constant (rate_subs.py:1904-2081) and its temperature /
// pressure derivatives (create_jacobian.py:2823-2976, :1158-1300)
template <bool JAC>
__device__ void cheb_task(Ctx& c, const State& st, int q) {
    const DevMech& m = c.m;
    const int maxT = m.cheb_max_t, maxP = m.cheb_max_p;
    const double Pmin = gp(m.cheb_plim)[2 * q], Pmax = gp(m.cheb_plim)[2 * q + 1];
    const double Tmin = gp(m.cheb_tlim)[2 * q], Tmax = gp(m.cheb_tlim)[2 * q + 1];
    const gptr<double> par = gp(m.cheb_params + (size_t)q * maxT * maxP);
    const double Pred = (-Pmax - Pmin + 2.0 * st.logP) / (Pmax - Pmin);
    const int kidx = m.n_plog + q;
    {
        const double Tred = (-Tmax - Tmin + 2.0 * st.Tinv) / (Tmax - Tmin);
        const int numP = gp(m.cheb_np)[q], numT = gp(m.cheb_nt)[q];
        double kt = 0.0;
        ChebPoly tp(Tred, Tred);
        for (int t = 0; t < maxT; ++t) {
            const double tt = tp.next();
            double temp = 0.0;
            ChebPoly pp(Pred, Pred);
            for (int k = 0; k < maxP; ++k) {
                const double pk = pp.next();
                if (k < numP) temp = temp + pk * par[t * maxP + k];
            }
            if (t < numT) kt = kt + tt * temp;
        }
        c.ksp(kidx) = exp(LN10 * fmin(EXP_MAX, kt));
    }
    if (!JAC) return;
    const double Tinv = st.Traw_inv;
    const double Tred = (2 * Tinv - Tmax - Tmin) / (Tmax - Tmin);
    const double Pred2 = (2 * st.logP - Pmax - Pmin) / (Pmax - Pmin);
    double dkT, dkP = 0.0;
    {
        const int numP = gp(m.cheb_np)[q], numT = gp(m.cheb_nt)[q] - 1;
        double dkf = 0;
        ChebPoly tp(Tred, 2 * Tred);
        for (int t = 0; t < maxT - 1; ++t) {
            const double tt = tp.next();
            double temp = 0;
            ChebPoly pp(Pred2, Pred2);
            for (int k = 0; k < maxP; ++k) {
                const double pk = pp.next();
                if (k < numP) temp = temp + pk * par[(t + 1) * maxP + k];
            }
            if (t < numT) dkf = dkf + (t + 1) * tt * temp;
        }
        dkT = -dkf * 2 * LN10 * Tinv * Tinv / (Tmax - Tmin);
    }
    if (!c.conp) {
        const int numP = gp(m.cheb_np)[q] - 1, numT = gp(m.cheb_nt)[q];
        double dkf = 0;
        ChebPoly tp(Tred, Tred);
        for (int t = 0; t < maxT; ++t) {
            const double tt = tp.next();
            double temp = 0;
            ChebPoly pp(Pred2, 2 * Pred2);
            for (int k = 0; k < maxP - 1; ++k) {
                const double pk = pp.next();
                if (k < numP) temp = temp + (k + 1) * pk * par[t * maxP + k + 1];
            }
            if (t < numT) dkf = dkf + tt * temp;
        }
        dkP = dkf * 2 * LN10 / (st.P * (Pmax - Pmin));
    }
    c.dkT(kidx) = dkT;
    c.dkP(kidx) = dkP;
}

// tasks are dealt to the waves in contiguous chunks of the kind-ordered
// list, so each wave sees as few code paths as possible
template <bool JAC, int WPS>
__device__ __forceinline__ void phase_pdep(Ctx& c, const State& st, int wave, int lane) {
    const DevMech& m = c.m;
    const int n1 = m.n_thd, n2 = n1 + m.n_plog, n3 = n2 + m.n_cheb;
    const int per = (n3 + WPS - 1) / WPS;
    const int end = min(n3, (wave + 1) * per);
    for (int i = wave * per + lane; i < end; i += 64) {
        if (i < n1) thd_task<JAC>(c, st, gp(c.bt.thd_order)[i]);
        else if (i < n2) plog_task<JAC>(c, st, i - n1);
        else cheb_task<JAC>(c, st, i - n2);
    }
}

// ------------------------------------------------------------- phase 3
// group flags of 64 consecutive reaction records (built by the runtime)
constexpr int GF_CNT = 0xf;          // largest species count of the group
constexpr int GF_REV = 1 << 4;       // a reversible reaction
constexpr int GF_THD = 1 << 5;       // a third-body / falloff reaction
constexpr int GF_SPECIAL = 1 << 6;   // a PLOG / Chebyshev reaction
constexpr int GF_NS = 1 << 7;        // the bath gas among some reaction's species
constexpr int GF_PAIRS = 1 << 8;     // third-body columns
constexpr int GF_POW = 1 << 9;       // Arrhenius form A T^b

// c^e for e >= 0 as the reference's product chain, branch-free up to e = 3
__device__ __forceinline__ double powu(double v, int e) {
    double r = e >= 1 ? v : 1.0;
    r = e >= 2 ? r * v : r;
    r = e >= 3 ? r * v : r;
    for (int i = 3; i < e; ++i) r *= v;
    return r;
}

// rate constant, reverse rate constant and the forward / reverse rates of
// progress of one reaction (rate_subs.py:2803-2870, :1627-1785,
// :1371-1507), concentrations from LDS
struct Rop {
    double kf, kr, rf, rr, ci;
    double cv[MAXSP];
};

__device__ __forceinline__ Rop rates_of_progress(Ctx& c, const State& st, const Rec& x, int gf) {
    Rop v;
    const int cm = gf & GF_CNT;
#pragma unroll
    for (int q = 0; q < MAXSP; ++q) v.cv[q] = q < cm ? (double)c.conc(x.spec(q)) : 1.0;
    // one exponential serves forms 2-4 (form 3 stores Ta = 0, form 4 b = 0)
    const double ek = exp(fmin(EXP_MAX, x.A + st.logT * x.b + -1.0 * x.Ta * st.Tinv));
    double kf = x.kind >= 2 ? ek : x.A;
    if (gf & GF_POW) {
        if (x.kind == 1) kf = x.A * powi(st.Tc, (int)x.b);
    }
    if (gf & GF_SPECIAL) {
        if (x.kind >= 5) kf = c.ksp(x.kidx);
    }
    v.kf = kf;
    double rf = kf;
#pragma unroll
    for (int q = 0; q < MAXSP; ++q)
        if (q < cm) rf = rf * powu(v.cv[q], x.nur(q));
    v.rf = rf;
    v.kr = 0.0;
    v.rr = 0.0;
    if (gf & GF_REV) {
        // equilibrium constant from the Gibbs terms (padding slots add 0)
        double B_sum = 0.0;
#pragma unroll
        for (int q = 0; q < MAXSP; ++q) {
            const int net_nu = x.nu(q);
            if (q < cm && net_nu != 0) B_sum = B_sum + (double)net_nu * c.b(x.spec(q));
        }
        B_sum = exp(fmin(EXP_MAX, B_sum));
        const double P_val = !(x.nsum > 0) ? RU / PA : PA / RU;
        const double P_sum = powu(P_val, x.nsum < 0 ? -x.nsum : x.nsum);
        const double kr = kf / (P_sum * B_sum);
        double rr = kr;
#pragma unroll
        for (int q = 0; q < MAXSP; ++q)
            if (q < cm) rr = rr * powu(v.cv[q], x.nup(q));
        v.kr = x.rev ? kr : 0.0;
        v.rr = x.rev ? rr : 0.0;
    }
    v.ci = 1.0;
    if (gf & GF_THD) {
        if (x.thd >= 0) v.ci = c.ci(x.thd);
    }
    return v;
}

// nu_k * rop into the production rates (get_spec_rates, rate_subs.py:1079-1167);
// slots without a net coefficient add 0 to the reaction's own row
__device__ __forceinline__ void scatter_wdot(Ctx& c, const Rec& x, double net, int gf) {
    const int cm = gf & GF_CNT;
#pragma unroll
    for (int q = 0; q < MAXSP; ++q) {
        if (q >= cm) break;
        const int nu = x.nu(q);
        lds_add(&c.wdot(nu != 0 ? x.spec(q) : x.home), (double)nu * net);
    }
}

// Where a reaction's Jacobian contributions go: added into a dense NN x NN
// array in LDS (mechanisms whose state fits the CU's LDS), or, for larger
// mechanisms, stored one slot each into the state's contribution buffer in
// device memory, whose slots the host groups by output entry (at most GC per
// entry, summed in slot order by a later gather).  Entries with more
// contributions than that (diagonals of frequent species) get an LDS slot
// that the contributions are added into instead.
struct JacAcc {
    lds_double* lds;
    double* g;
    lds_double *col0, *col1;  // device-memory variant: columns T and V|P in LDS,
    lds_double* lng;          // and the entries with many contributions
};

// one contribution: to the dense LDS entry, or to the buffer slot that the
// record's slot table holds for (column slot, row slot)
template <bool G>
__device__ __forceinline__ void jput(const JacAcc& a, gptr<int> pos, int slot,
                                     int dense, double v) {
    if constexpr (!G) {
        lds_add(a.lds + dense, v);
    } else {
        const int e = pos[slot];
        if (e >= 0) a.g[e] = v;
        else if (e < -1) lds_add(a.lng + (-2 - e), v);
    }
}

// all per-reaction Jacobian contributions of one reaction: dRopidnj(+_ns)
// (create_jacobian.py:4169-4484), dRopi*_dT(+_ns) (:2978-2992 and the _ns
// kernels), dRopi*_dE(+_ns) (:1040-1080, :1158-1300) and the third-body
// parts through the per-entry factors of phase 2, scattered into LDS.
// Slots that contribute nothing add 0 to an entry of the reaction's own
// row instead of branching.
template <bool G>
__device__ __forceinline__ void reaction_jacobian(Ctx& c, const State& st, const Rec& x,
                                                  const Rop& v, double ropi, int gf,
                                                  const JacAcc& acc,
                                                  gptr<int> ent) {
    const int ns = c.ns, NN = c.NN;
    const int cm = gf & GF_CNT;
    const int conp = c.conp;
    const double Tinv = st.Traw_inv;
    const double kr_i = v.kr;
    // rows: species with a net coefficient other than the bath gas;
    // columns: species other than the bath gas
    int roff[MAXSP], coff[MAXSP];
    double nuk[MAXSP];
    const int hrow = __umul24(x.home + 2, NN), hcol = x.home + 2;
#pragma unroll
    for (int q = 0; q < MAXSP; ++q) {
        const int spec = x.spec(q), nu = x.nu(q);
        const bool in = q < x.cnt && spec != ns;
        coff[q] = in ? spec + 2 : hcol;
        const bool rowok = in && nu != 0;
        roff[q] = rowok ? __umul24(spec + 2, NN) : hrow;
        nuk[q] = rowok ? (double)nu : 0.0;
    }
    // d(prod_i c_i^nu_i)/dc_j = nu_j c_j^(nu_j-1) prod_{i != j} c_i^nu_i
    // with the concentrations floored at 1e-300: suffix products first,
    // then one forward sweep that scatters column j as soon as it is known
    double sfx_f[MAXSP], sfx_r[MAXSP];
    {
        double af = 1.0, ar = 1.0;
#pragma unroll
        for (int q = MAXSP - 1; q >= 0; --q) {
            sfx_f[q] = af;
            sfx_r[q] = ar;
            if (q < cm) {
                const double cg = fmax(SMALL, v.cv[q]);
                af *= powu(cg, x.nur(q));
                ar *= powu(cg, x.nup(q));
            }
        }
    }
    double Sf_l = 0.0, Sr_l = 0.0;
    {
        double pf = 1.0, pr = 1.0;
#pragma unroll
        for (int j = 0; j < MAXSP; ++j) {
            if (j >= cm) break;
            const double cg = fmax(SMALL, v.cv[j]);
            const int a = x.nur(j), b = x.nup(j);
            const double df = a ? a * powu(cg, a > 0 ? a - 1 : 0) : 0.0;
            const double dr = b ? b * powu(cg, b > 0 ? b - 1 : 0) : 0.0;
            const double Sf = pf * df * sfx_f[j];
            const double Sr = pr * dr * sfx_r[j];
            if (j == x.cnt - 1) { Sf_l = Sf; Sr_l = Sr; }
            pf *= powu(cg, a);
            pr *= powu(cg, b);
            const bool colok = j < x.cnt && x.spec(j) != ns;
            const double d = colok ? (v.kf * Sf + -1.0 * kr_i * Sr) * v.ci : 0.0;
#pragma unroll
            for (int k = 0; k < MAXSP; ++k) {
                if (k >= cm) break;
                jput<G>(acc, ent, j * MAXSP + k, roff[k] + coff[j], nuk[k] * d);
            }
        }
    }
    // temperature and extra-variable derivatives of the rate of progress
    double dkT = (x.b + x.Ta * Tinv) * Tinv, dkP = 0.0;
    if (gf & GF_SPECIAL) {
        if (x.kind > 4) { dkT = c.dkT(x.kidx); dkP = c.dkP(x.kidx); }
    }
    double dT, dE = 0.0;
    {
        double dd = v.rf * dkT;
        if (gf & GF_REV) {
            double dB = 0.0;
#pragma unroll
            for (int q = 0; q < MAXSP; ++q)
                if (q < cm) dB = dB + (double)x.nu(q) * c.db(x.spec(q));
            if (x.rev) dd = dd + -1.0 * v.rr * (dkT + -1.0 * dB);
        }
        dT = dd * v.ci * st.V;
    }
    if (conp) {
        int nu_f = -1, nu_r = -1;
#pragma unroll
        for (int q = 0; q < MAXSP; ++q)
            if (q < cm) { nu_f += x.nur(q); nu_r += x.nup(q); }
        double dd = -1.0 * nu_f * v.rf;
        if (x.rev) dd = dd + nu_r * v.rr;
        if (x.thd >= 0) dd = dd * v.ci;
        dE = dd;
    } else if (gf & GF_SPECIAL) {
        if (x.kind > 4) dE = (v.rf - v.rr) * dkP * st.V;
    }
    double u = 0.0;
    if (gf & GF_NS) {
        // bath gas among the reaction's species (always its last slot): the
        // _ns kernels, d/dn_ns = -d/dn_j for every species column j
        int last = 0;
#pragma unroll
        for (int q = 0; q < MAXSP; ++q)
            if (q == x.cnt - 1) last = x.spec(q);
        if (last == ns) {
            dT += (Sr_l * kr_i + -1.0 * Sf_l * v.kf) * st.V * v.ci * st.P / (RU * st.T * st.T);
            const double fac = conp ? st.P * st.rt_inv : st.V * st.rt_inv;
            dE += (Sf_l * v.kf - Sr_l * kr_i) * v.ci * fac;
            u = (kr_i * Sr_l + -1.0 * v.kf * Sf_l) * v.ci;
        }
    }
    if (gf & GF_THD) {
        if (x.thd >= 0) {
            // third-body columns: nu_k * d(ci)/dn_j * rop, accumulated per
            // (efficiency set, row species) pair and spread in phase 4
            const int t = x.thd;
            const double trop = ropi * c.nfac(t);
            u += gp(c.bt.effns)[t] * trop;
            dT += c.dTf(t) * ropi;
            dE += c.dEf(t) * ropi;
            if (gf & GF_PAIRS) {
                const int4 p0 = ld_int4(c.bt.thd_pairs, 2 * t);
                const int2 p1 = ld_int2(c.bt.thd_pairs, 4 * t + 2);
                const int pp[MAXSP] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y};
#pragma unroll
                for (int q = 0; q < MAXSP; ++q)
                    if (q < cm && pp[q] >= 0) lds_add(&c.Cp(pp[q]), (double)x.nu(q) * trop);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < MAXSP; ++k) {
        if (k >= cm) break;
        if constexpr (G) {
            // columns T and V|P of the species rows stay in LDS
            const int spec = x.spec(k);
            const bool rowok = k < x.cnt && spec != ns && x.nu(k) != 0;
            const int row = rowok ? spec : x.home;
            lds_add(acc.col0 + row, nuk[k] * dT);
            lds_add(acc.col1 + row, nuk[k] * dE);
        } else {
            jput<G>(acc, ent, 6 * MAXSP + k, roff[k], nuk[k] * dT);
            jput<G>(acc, ent, 7 * MAXSP + k, roff[k] + 1, nuk[k] * dE);
        }
    }
    if (gf & (GF_NS | GF_THD)) {
#pragma unroll
        for (int k = 0; k < MAXSP; ++k) {
            if (k >= cm) break;
            const int spec = x.spec(k);
            const bool rowok = k < x.cnt && spec != ns && x.nu(k) != 0;
            lds_add(&c.U(rowok ? spec : x.home), nuk[k] * u);
        }
    }
}

// ------------------------------------------------------------- kernels

struct Layout {
    size_t sN, sI;  // stride between states, between entries of a state
};

// XCD-aware state of a workgroup: workgroup g runs on XCD g % 8, so give
// each XCD a contiguous range of states; neighbouring states then run on
// one XCD at the same time and their state-fastest (F order) reads and
// writes meet in that XCD's L2.  The grid is 8 * ceil(n / 8) workgroups.
__device__ __forceinline__ long state_of_block(long g, long n) {
    const long per = (n + 7) / 8;
    return (g % 8) * per + g / 8;
}

__device__ __forceinline__ State load_state(int conp, double param, double x0, double x1) {
    State st;
    st.T = x0;
    if (conp) { st.P = param; st.V = x1; }
    else { st.P = x1; st.V = param; }
    st.Tc = clampT(st.T);
    st.logT = log(st.Tc);
    st.Tinv = 1.0 / st.Tc;
    st.logP = log(fmax(SMALL, st.P));
    st.Traw_inv = 1.0 / st.T;
    st.rt_inv = 1.0 / (RU * st.T);
    st.V_inv = 1.0 / fmax(V_MIN, st.V);
    return st;
}

// species rates of state s (one workgroup of 64 * WPS threads)
template <int WPS>
__device__ __forceinline__ void rates_state(const BlockArgs* __restrict__ A, int conp, long n,
                                            long s, const double* __restrict__ param,
                                            const double* __restrict__ phi,
                                            double* __restrict__ dphi, Layout L, double* smem) {
    constexpr int NT = 64 * WPS;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    Ctx c = make_ctx(A, smem, conp);
    const int Ns = c.Ns, Nr = c.m.Nr;
    if (tid == 0) { c.conc(Ns) = 1.0; c.b(Ns) = 0.0; }
    const double* ps = phi + s * L.sN;
    auto ph = [&](int i) { return ps[i * L.sI]; };
    const State st = load_state(conp, param[s], ph(0), ph(1));
    phase_species<false, NT>(c, st, tid, ph);
    __syncthreads();
    bath_gas(c, st, lane);
    phase_pdep<false, WPS>(c, st, wave, lane);
    __syncthreads();
    for (int pass = 0; pass * NT < Nr; ++pass) {
        const int r = pass * NT + tid;
        const int gf = sld(c.bt.grp_flags, pass * WPS + wave);
        if (r < Nr) {
            const Rec x = load_rec(c.bt, r);
            const Rop v = rates_of_progress(c, st, x, gf);
            scatter_wdot(c, x, (v.rf - v.rr) * v.ci, gf);
        }
    }
    __syncthreads();
    // dphi (rate_subs.py:709-1076); the three sums in the reference's
    // sequential order, one lane each
    if (wave == 0) {
        const double V_val = fmax(V_MIN, st.V);
        double acc = 0.0;
        if (lane < 3) {
            for (int k = 0; k < Ns; ++k) {
                const double w = c.wdot(k);
                const double mwf = k < Ns - 1 ? 1.0 + -1.0 * sld(c.m.mw_post, k) : 0.0;
                const double a = lane == 0 ? (double)c.conc(k) : (lane == 1 ? (double)c.energy(k) : mwf);
                const double bb = lane == 0 ? (double)c.heat(k) : w;
                acc = __dadd_rn(acc, __dmul_rn(a, bb));
            }
        }
        const double lower = __shfl(acc, 0), upper = __shfl(acc, 1), dE = __shfl(acc, 2);
        const double Tdot = 0.0 + -1.0 * upper / lower;
        double Edot;
        if (conp) {
            Edot = V_val * Tdot / st.Tc;
            Edot = Edot + V_val * dE * st.Tc * RU / st.P;
        } else {
            Edot = st.P * Tdot / st.Tc;
            Edot = Edot + st.Tc * RU * dE;
        }
        double* out = dphi + s * L.sN;
        for (int k = lane; k < Ns + 1; k += 64) {
            double val;
            if (k == 0) val = Tdot;
            else if (k == 1) val = Edot;
            else val = V_val * c.wdot(k - 2);
            out[k * L.sI] = val;
        }
    }
}

// analytical Jacobian of state s (one workgroup of 64 * WPS threads)
template <int WPS>
__device__ __forceinline__ void jacobian_state(const BlockArgs* __restrict__ A, int conp, long n,
                                               long s, const double* __restrict__ param,
                                               const double* __restrict__ phi,
                                               double* __restrict__ jac, Layout L, int out_C,
                                               double* smem) {
    constexpr int NT = 64 * WPS;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    Ctx c = make_ctx(A, smem, conp);
    const int Ns = c.Ns, NN = c.NN, ns = c.ns, Nr = c.m.Nr, nout = c.m.jac_nnz;
    lds_double* J = c.J();
    // issue priority over the F-order transposes that share the CUs (they
    // wait on memory, these waves on their dependent chains): 0.3 % faster
    __builtin_amdgcn_s_setprio(2);
    const JacAcc acc{J, nullptr, nullptr, nullptr};
    // rows 0 and 1 of the accumulator hold the Gibbs terms b and their T
    // derivatives db until the reaction pass is over (they are first
    // written in phase 5), so only rows 2 .. NN - 1 are cleared here
    if (tid == 0) { c.conc(Ns) = 1.0; c.b(Ns) = 0.0; c.db(Ns) = 0.0; }
    const double* ps = phi + s * L.sN;
    auto ph = [&](int i) { return ps[i * L.sI]; };
    const State st = load_state(conp, param[s], ph(0), ph(1));
    {
        typedef double v2d __attribute__((ext_vector_type(2)));
        typedef __attribute__((address_space(3))) v2d lds_v2d;
        lds_v2d* J2 = (lds_v2d*)J;
        const int half = (NN * NN) / 2;
        for (int e = NN + tid; e < half; e += NT) J2[e] = v2d{0.0, 0.0};
        if (tid == 0 && (NN * NN) % 2) J[NN * NN - 1] = 0.0;
        for (int p = tid; p < c.bt.n_pairs; p += NT) c.Cp(p) = 0.0;
    }
    // pressure-dependent entries: when the species leave enough threads
    // idle in phase 1 (GRI: 53 species, 41 entries, 128 threads), the last
    // threads evaluate the temperature-only parts of the falloff entries and
    // the whole PLOG / Chebyshev entries there, and finish the falloff
    // entries after the barrier
    const int n1 = c.m.n_thd, n2 = n1 + c.m.n_plog, n3 = n2 + c.m.n_cheb;
    const bool early = Ns + n3 <= NT;
    const int task = tid - (NT - n3);
    ThdPre pre{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0, -1};
    phase_species<true, NT, true>(c, st, tid, ph);
    // with the entries on the last threads, wave 0 sums the bath gas in
    // phase 1 (its lane 0 stores it before the barrier)
    if (early && wave == 0) bath_gas_phi(c, st, lane, ph);
    if (early && task >= 0) {
        if (task < n1) pre = thd_pre<true>(c, st, gp(c.bt.thd_order)[task]);
        else if (task < n2) plog_task<true>(c, st, task - n1);
        else cheb_task<true>(c, st, task - n2);
    }
    __syncthreads();
    if (!early) bath_gas_phi(c, st, lane, ph);
    if (early) {
        if (task >= 0 && task < n1)
            thd_post<true, true>(c, st, gp(c.bt.thd_order)[task], [&]() { return pre; });
    } else {
        phase_pdep<true, WPS>(c, st, wave, lane);
    }
    __syncthreads();
    // phase 3: reactions
    for (int pass = 0; pass * NT < Nr; ++pass) {
        const int r = pass * NT + tid;
        const int gf = sld(c.bt.grp_flags, pass * WPS + wave);
        if (r < Nr) {
            const Rec x = load_rec(c.bt, r);
            const Rop v = rates_of_progress(c, st, x, gf);
            const double ropi = v.rf - v.rr;
            scatter_wdot(c, x, ropi * v.ci, gf);
            reaction_jacobian<false>(c, st, x, v, ropi, gf, acc, nullptr);
        }
    }
    __syncthreads();
    // phase 4: third-body columns d(ci)/dn_j (create_jacobian.py
    // dci_thd_dnj / dci_lind / dci_sri / dci_troe): every non-zero
    // (pair, column) of the host's list adds coefficient * efficiency
    {
        for (int i = tid; i < c.bt.n_combo; i += NT) {
            const int4 e = ld_int4(c.bt.combo, i);
            lds_add(J + e.y, c.Cp(e.x) * __hiloint2double(e.w, e.z));
        }
    }
    // b and db are spent: clear rows 0 and 1 for phase 5
    for (int e = tid; e < 2 * NN; e += NT) J[e] = 0.0;
    // temperature rate and total heat capacity (every wave), with the
    // species' thermodynamic properties recomputed here; wave 0 keeps them
    // in LDS (over the third-body arrays of the reaction pass) for phase 5
    double heat_tot, Tdot;
    {
        double lo = 0.0, up = 0.0;
        for (int k = lane; k < Ns; k += 64) {
            const Thermo h = nasa_thermo(nasa_coeffs(c.m, st.Tc, k), st.Tc, conp);
            lo += c.conc(k) * h.heat;
            up += h.energy * c.wdot(k);
            if (wave == 0) {
                c.energy(k) = h.energy;
                c.heat(k) = h.heat;
                c.dheat(k) = h.dheat;
            }
        }
        heat_tot = wave_sum(lo);
        Tdot = 0.0 + -1.0 * wave_sum(up) / heat_tot;
    }
    __syncthreads();
    // phase 5: rows 0 and 1 of the species columns (dTdot_dnj, dEdot_dnj,
    // create_jacobian.py:3266-3484): lanes over columns, waves over blocks
    // of rows; the bath-gas sums U join every species column of their row
    const double T = st.T, Tinv = st.Traw_inv;
    const double fixed = conp ? st.P : st.V;
    const double extra = conp ? st.V : st.P;
    const double e_ns = c.energy(ns), h_ns = c.heat(ns);
    const double inv_vh = 1.0 / (st.V * heat_tot);
    {
        const int per = (Ns - 1 + WPS - 1) / WPS;
        const int k0 = wave * per, k1 = min(Ns - 1, k0 + per);
        for (int j = lane; j < Ns - 1; j += 64) {
            double sum = 0.0, sum2 = 0.0;
#pragma unroll 4
            for (int k = k0; k < k1; ++k) {
                const double jv = J[(k + 2) * NN + j + 2] + c.U(k);
                const double mw = sld(c.m.mw_post, k);
                sum = sum + (c.energy(k) + -1.0 * e_ns * mw) * jv;
                sum2 = sum2 + (1.0 + -1.0 * mw) * jv;
            }
            double j0 = -1.0 * sum * inv_vh;
            if (wave == 0) j0 = j0 + -1.0 * Tdot * (c.heat(j) + -1.0 * h_ns) * inv_vh;
            const double j1 = T * RU * sum2 / fixed + extra * j0 * Tinv;
            if (WPS == 1) {
                J[j + 2] = j0;
                J[NN + j + 2] = j1;
            } else {
                lds_add(J + j + 2, j0);
                lds_add(J + NN + j + 2, j1);
            }
        }
    }
    if (wave == WPS - 1) {
        // columns 0 and 1 of rows 0 and 1 (create_jacobian.py:1595-2114)
        const double Vinv = 1.0 / st.V;
        double rate_sum = 0.0, sumT = 0.0, dTsumE = 0.0, specsum = 0.0, sumE = 0.0,
               sumEE = 0.0, dTsum = 0.0;
        for (int i = lane; i < Ns - 1; i += 64) {
            const double j0 = J[(i + 2) * NN];
            const double j1 = J[(i + 2) * NN + 1];
            const double mw = gp(c.m.mw_post)[i];
            const double w = c.wdot(i);
            rate_sum += Vinv * j0 * (-1.0 * c.energy(i) + e_ns * mw) +
                        w * (-1.0 * c.heat(i) + mw * h_ns);
            dTsum += (h_ns * Tinv + -1.0 * c.dheat(i)) * c.conc(i);
            if (conp) {
                sumT += (1.0 + -1.0 * mw) * (Vinv * j0 + Tinv * w);
                dTsumE += (c.heat(i) + -1.0 * h_ns) * c.conc(i);
                specsum += (c.energy(i) + -1.0 * e_ns * mw) * (j1 + -1.0 * w);
            } else {
                sumT += (1.0 - mw) * (j0 * (T / st.V) + w);
                sumE += (c.energy(i) - e_ns * mw) * j1;
            }
            sumEE += (1 - mw) * j1;
        }
        rate_sum = wave_sum(rate_sum);
        dTsum = wave_sum(dTsum) + (h_ns * Tinv + -1.0 * c.dheat(ns)) * c.conc(ns);
        sumT = wave_sum(sumT);
        sumEE = wave_sum(sumEE);
        if (conp) {
            dTsumE = wave_sum(dTsumE);
            specsum = wave_sum(specsum);
        } else {
            sumE = wave_sum(sumE);
        }
        if (lane == 0) {
            const double j00 = (Tdot * dTsum + rate_sum) / heat_tot;
            J[0] = j00;
            double j10;
            if (conp) {
                j10 = st.V * Tinv * (j00 + -1.0 * Tinv * Tdot);
                j10 = j10 + RU * T * st.V * sumT / st.P;
            } else {
                j10 = st.P * (j00 - Tdot * Tinv) * Tinv;
                j10 = j10 + RU * sumT;
            }
            J[NN] = j10;
            double j01;
            if (conp) {
                j01 = (Tdot * dTsumE + -1.0 * specsum) * (1.0 / (heat_tot * st.V));
            } else {
                const double spec_inv = 1.0 / heat_tot;
                j01 = -(Tdot * h_ns / (RU * T)) * spec_inv;
                j01 = j01 - (sumE * Vinv) * spec_inv;
            }
            J[1] = j01;
            const double var = conp ? st.V : st.P;
            const double par = conp ? st.P : st.V;
            const double j11 = (var * j01 + Tdot) / T;
            J[NN + 1] = j11 + RU * T * sumEE / par;
        }
    }
    __syncthreads();
    // phase 6: stored entries of this state, eight per thread in flight;
    // a C-order row of even length from a 16-byte aligned output goes out
    // as pairs of entries (16-byte index loads and stores)
    if (out_C && (nout & 1) == 0 && ((size_t)jac & 15) == 0) {
        typedef double v2d __attribute__((ext_vector_type(2)));
        constexpr int UNR = 4;
        const int np = nout >> 1;
        v2d* row = (v2d*)(jac + s * (size_t)nout);
        for (int q0 = tid; q0 < np; q0 += NT * UNR) {
            int4 p[UNR];
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                const int q = q0 + u * NT;
                p[u] = q < np ? ld_int4(c.bt.out_pos, q) : make_int4(0, -1, 0, -1);
            }
            v2d val[UNR];
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                double a = J[p[u].x], b = J[p[u].z];
                if (p[u].y >= 0) a += c.U(p[u].y);
                if (p[u].w >= 0) b += c.U(p[u].w);
                val[u] = v2d{a, b};
            }
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                const int q = q0 + u * NT;
                if (q < np) row[q] = val[u];
            }
        }
    } else {
        constexpr int UNR = 8;
        for (int e0 = tid; e0 < nout; e0 += NT * UNR) {
            int2 p[UNR];
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                const int e = e0 + u * NT;
                p[u] = e < nout ? ld_int2(c.bt.out_pos, e) : make_int2(0, -1);
            }
            double val[UNR];
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                val[u] = J[p[u].x];
                if (p[u].y >= 0) val[u] += c.U(p[u].y);
            }
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                const int e = e0 + u * NT;
                if (e < nout) {
                    if (out_C) jac[s * (size_t)nout + e] = val[u];
                    else jac[(size_t)e * n + s] = val[u];
                }
            }
        }
    }
}

// gather of the device-memory variant: entries per thread in flight, and
// the contributions of each loaded together
constexpr int GE = 8, GC = pjk_host::GATHER_MAX;

// Load of data stored by other waves of this workgroup (after their stores
// completed and a barrier): workgroup scope, served by this CU's caches.
// (Agent scope would bypass the XCD's L2 and read from memory.)
__device__ __forceinline__ double wg_load(const double* p) {
    return __hip_atomic_load(gp(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Every memory operation of this wave completed before the barrier that
// follows, so that the next phase of the workgroup sees it.
__device__ __forceinline__ void wait_memory() { __builtin_amdgcn_s_waitcnt(0); }

// Analytical Jacobian of state s for mechanisms whose dense accumulator does
// not fit in LDS.  The species-level arrays stay in LDS; each reaction stores
// its contributions into the state's buffer cb (one slot per contribution,
// slots grouped by output entry); a gather then sums every stored entry's
// run, adds its third-body columns and bath-gas sum, and writes the entry to
// out (the state's nout entries in output order).  Same arithmetic as
// jacobian_state; the rows of T and the extra variable sum the stored
// species entries of each column plus the bath-gas sums of every row.
template <int WPS>
__device__ __forceinline__ void jacobian_global_state(const BlockArgs* __restrict__ A, int conp,
                                                      long s, const double* __restrict__ param,
                                                      const double* __restrict__ phi,
                                                      double* __restrict__ cb,
                                                      double* __restrict__ out, Layout L,
                                                      double* smem, long ostride = 1) {
    constexpr int NT = 64 * WPS;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    Ctx c = make_ctx(A, smem, conp);
    const int Ns = c.Ns, NN = c.NN, ns = c.ns, Nr = c.m.Nr, nout = c.m.jac_nnz;
    lds_double* cs0 = c.sm + c.L.jac;  // column sums of the T row
    lds_double* cs1 = cs0 + NN;        // and of the extra-variable row
    lds_double* col0 = c.sm + c.L.aux; // columns 0 and 1 of the species rows
    lds_double* col1 = col0 + NN;
    lds_double* lng = col1 + NN;       // entries with many contributions
    const JacAcc acc{nullptr, cb, col0, col1, lng};
    const bool approx = c.bt.approx_sums != 0;
    if (tid == 0) { c.conc(Ns) = 1.0; c.b(Ns) = 0.0; c.db(Ns) = 0.0; }
    for (int k = tid; k < Ns - 1; k += NT) c.mw(k) = gp(c.m.mw_post)[k];
    for (int p = tid; p < c.bt.n_pairs; p += NT) c.Cp(p) = 0.0;
    const double* ps = phi + s * L.sN;
    auto ph = [&](int i) { return ps[i * L.sI]; };
    const State st = load_state(conp, param[s], ph(0), ph(1));
    phase_species<true, NT>(c, st, tid, ph);
    __syncthreads();
    bath_gas(c, st, lane);
    phase_pdep<true, WPS>(c, st, wave, lane);
    __syncthreads();
    // nmol is no longer read: its space takes columns T and V|P and the
    // many-contribution entries
    for (int i = tid; i < 2 * NN + c.bt.n_long; i += NT) col0[i] = 0.0;
    __syncthreads();
    // phase 3: reactions
    for (int pass = 0; pass * NT < Nr; ++pass) {
        const int r = pass * NT + tid;
        const int gf = sld(c.bt.grp_flags, pass * WPS + wave);
        if (r < Nr) {
            const Rec x = load_rec(c.bt, r);
            const Rop v = rates_of_progress(c, st, x, gf);
            const double ropi = v.rf - v.rr;
            scatter_wdot(c, x, ropi * v.ci, gf);
            reaction_jacobian<true>(c, st, x, v, ropi, gf, acc,
                                    gp(c.bt.rec_pos + (size_t)r * (8 * MAXSP)));
        }
    }
    __syncthreads();
    // the reaction-only arrays are dead: their space takes the column sums
    for (int k = tid; k < 2 * NN; k += NT) cs0[k] = 0.0;
    // third-body columns (create_jacobian.py dci_*_dnj): each (pair, column)
    // of the host's list is one more contribution of its entry
    {
        for (int i = tid; i < c.bt.n_combo; i += NT) {
            const int4 m4 = ld_int4(c.bt.gcombo, i);
            const double v = c.Cp(m4.x) * __hiloint2double(m4.w, m4.z);
            if (m4.y >= 0) cb[m4.y] = v;
            else lds_add(lng + (-2 - m4.y), v);
        }
    }
    wait_memory();
    __syncthreads();
    double heat_tot, Tdot;
    {
        double lo = 0.0, up = 0.0;
        for (int k = lane; k < Ns; k += 64) {
            lo += c.conc(k) * c.heat(k);
            up += c.energy(k) * c.wdot(k);
        }
        heat_tot = wave_sum(lo);
        Tdot = 0.0 + -1.0 * wave_sum(up) / heat_tot;
    }
    const double T = st.T, Tinv = st.Traw_inv;
    const double fixed = conp ? st.P : st.V;
    const double extra = conp ? st.V : st.P;
    const double e_ns = c.energy(ns), h_ns = c.heat(ns);
    const double inv_vh = 1.0 / (st.V * heat_tot);
    // phase 4: gather the stored species-row entries.  An entry's buffer
    // contributions (at most GC) are contiguous; a thread takes GE entries
    // at a time with all their contributions in flight together and sums
    // each in slot order; entries with more contributions are in LDS
    {
        const gptr<int> es = gp(c.bt.ent_start);
        const gptr<int> rc = gp(c.bt.ent_rc);
        const gptr<int> el = gp(c.bt.ent_long);
        for (int eb = tid; eb < nout; eb += NT * GE) {
            int trc[GE], a[GE], b[GE], lg[GE];
#pragma unroll
            for (int u = 0; u < GE; ++u) {
                const int e = eb + u * NT;
                const bool in = e < nout;
                trc[u] = in ? rc[e] : 0;
                a[u] = in ? es[e] : 0;
                b[u] = in ? es[e + 1] : 0;
                lg[u] = in ? el[e] : -1;
            }
            double x[GE][GC];
#pragma unroll
            for (int u = 0; u < GE; ++u)
#pragma unroll
                for (int i = 0; i < GC; ++i)
                    x[u][i] = a[u] + i < b[u] ? wg_load(cb + a[u] + i) : 0.0;
#pragma unroll
            for (int u = 0; u < GE; ++u) {
                const int e = eb + u * NT;
                const int row = trc[u] >> 16, col = trc[u] & 0xffff;
                if (row < 2) continue;  // rows 0 and 1 come last (and padding)
                const int k = row - 2;
                double v;
                if (col >= 2) {
                    if (lg[u] >= 0) {
                        v = lng[lg[u]];
                    } else {
                        v = x[u][0];
#pragma unroll
                        for (int i = 1; i < GC; ++i) v += x[u][i];
                    }
                    const double mw = c.mw(k);
                    const double vu = v + c.U(k);
                    const double vs = approx ? vu : v;
                    lds_add(cs0 + col - 2, (c.energy(k) + -1.0 * e_ns * mw) * vs);
                    lds_add(cs1 + col - 2, (1.0 + -1.0 * mw) * vs);
                    v = vu;
                } else {
                    v = col == 0 ? col0[k] : col1[k];
                }
                out[e * ostride] = v;
            }
        }
    }
    // the bath-gas sums join every species column of their row (with an
    // approximate pattern only the stored entries, summed above)
    double su0 = 0.0, su1 = 0.0;
    for (int k = lane; k < Ns - 1 && !approx; k += 64) {
        const double u = c.U(k), mw = c.mw(k);
        su0 += (c.energy(k) + -1.0 * e_ns * mw) * u;
        su1 += (1.0 + -1.0 * mw) * u;
    }
    su0 = wave_sum(su0);
    su1 = wave_sum(su1);
    __syncthreads();
    // phase 5: rows 0 and 1 of the species columns (dTdot_dnj, dEdot_dnj,
    // create_jacobian.py:3266-3484)
    const gptr<int> amap = gp(c.bt.acc_map);
    for (int j = tid; j < Ns - 1; j += NT) {
        double j0 = -1.0 * (cs0[j] + su0) * inv_vh;
        j0 = j0 + -1.0 * Tdot * (c.heat(j) + -1.0 * h_ns) * inv_vh;
        const double j1 = T * RU * (cs1[j] + su1) / fixed + extra * j0 * Tinv;
        out[amap[j + 2] * ostride] = j0;
        out[amap[NN + j + 2] * ostride] = j1;
    }
    if (wave == WPS - 1) {
        // columns 0 and 1 of rows 0 and 1 (create_jacobian.py:1595-2114)
        const double Vinv = 1.0 / st.V;
        double rate_sum = 0.0, sumT = 0.0, dTsumE = 0.0, specsum = 0.0, sumE = 0.0,
               sumEE = 0.0, dTsum = 0.0;
        for (int i = lane; i < Ns - 1; i += 64) {
            const double j0 = col0[i];
            const double j1 = col1[i];
            const double mw = c.mw(i);
            const double w = c.wdot(i);
            rate_sum += Vinv * j0 * (-1.0 * c.energy(i) + e_ns * mw) +
                        w * (-1.0 * c.heat(i) + mw * h_ns);
            dTsum += (h_ns * Tinv + -1.0 * c.dheat(i)) * c.conc(i);
            if (conp) {
                sumT += (1.0 + -1.0 * mw) * (Vinv * j0 + Tinv * w);
                dTsumE += (c.heat(i) + -1.0 * h_ns) * c.conc(i);
                specsum += (c.energy(i) + -1.0 * e_ns * mw) * (j1 + -1.0 * w);
            } else {
                sumT += (1.0 - mw) * (j0 * (T / st.V) + w);
                sumE += (c.energy(i) - e_ns * mw) * j1;
            }
            sumEE += (1 - mw) * j1;
        }
        rate_sum = wave_sum(rate_sum);
        dTsum = wave_sum(dTsum) + (h_ns * Tinv + -1.0 * c.dheat(ns)) * c.conc(ns);
        sumT = wave_sum(sumT);
        sumEE = wave_sum(sumEE);
        if (conp) {
            dTsumE = wave_sum(dTsumE);
            specsum = wave_sum(specsum);
        } else {
            sumE = wave_sum(sumE);
        }
        if (lane == 0) {
            const double j00 = (Tdot * dTsum + rate_sum) / heat_tot;
            out[amap[0] * ostride] = j00;
            double j10;
            if (conp) {
                j10 = st.V * Tinv * (j00 + -1.0 * Tinv * Tdot);
                j10 = j10 + RU * T * st.V * sumT / st.P;
            } else {
                j10 = st.P * (j00 - Tdot * Tinv) * Tinv;
                j10 = j10 + RU * sumT;
            }
            out[amap[NN] * ostride] = j10;
            double j01;
            if (conp) {
                j01 = (Tdot * dTsumE + -1.0 * specsum) * (1.0 / (heat_tot * st.V));
            } else {
                const double spec_inv = 1.0 / heat_tot;
                j01 = -(Tdot * h_ns / (RU * T)) * spec_inv;
                j01 = j01 - (sumE * Vinv) * spec_inv;
            }
            out[amap[1] * ostride] = j01;
            const double var = conp ? st.V : st.P;
            const double par = conp ? st.P : st.V;
            const double j11 = (var * j01 + Tdot) / T;
            out[amap[NN + 1] * ostride] = j11 + RU * T * sumEE / par;
        }
    }
}

// Occupancy targets (waves per SIMD) that the register allocator is held to.
// Unconstrained, the Jacobian kernels take 171-181 VGPRs (2 waves per SIMD);
// held to 3 waves (168 VGPRs; the dense kernel spills 12 registers) GRI
// runs 12 % faster, H2/O2 25 % and the device-memory variant on 150 species
// 15 %.  The species-rate kernel at 5 waves (96 VGPRs and a 10-register
// spill, from 111 VGPRs and 4 waves) gains 11 %.  Four waves for the
// Jacobian kernels spill 48-56 registers and lose.
#ifndef PJB_JAC_WAVES
#define PJB_JAC_WAVES 3
#endif
#ifndef PJB_RATES_WAVES
#define PJB_RATES_WAVES 5
#endif
#ifndef PJB_GLOB_WAVES
#define PJB_GLOB_WAVES 3
#endif

// OCC: the occupancy target (waves per SIMD) the register allocator is held
// to; the default is the measured best for GRI-Mech 3.0, the other values
// are instantiated for the autotuner (2 and 4 for the Jacobian kernels with
// one and two waves per state, 4 and 6 for the one-wave species rates)
template <int WPS, int OCC = PJB_RATES_WAVES>
__global__ void __launch_bounds__(64 * WPS) __attribute__((amdgpu_waves_per_eu(OCC)))
rates_block_kernel(const BlockArgs* __restrict__ A, int conp, long n,
                   const double* __restrict__ param, const double* __restrict__ phi,
                   double* __restrict__ dphi, Layout L) {
    extern __shared__ double smem[];
    const long s = state_of_block(blockIdx.x, n);
    if (s >= n) return;
    rates_state<WPS>(A, conp, n, s, param, phi, dphi, L, smem);
}

template <int WPS, int OCC = PJB_JAC_WAVES>
__global__ void __launch_bounds__(64 * WPS) __attribute__((amdgpu_waves_per_eu(OCC)))
jacobian_block_kernel(const BlockArgs* __restrict__ A, int conp, long n,
                      const double* __restrict__ param, const double* __restrict__ phi,
                      double* __restrict__ jac, Layout L, int out_C) {
    extern __shared__ double smem[];
    const long s = state_of_block(blockIdx.x, n);
    if (s >= n) return;
    jacobian_state<WPS>(A, conp, n, s, param, phi, jac, L, out_C, smem);
}

// states s0 .. s0 + count - 1, one workgroup each; state s0 + i stores its
// contributions at cbuf + i * ncontrib and its entries at out + i * nout (the
// output itself in C order, a chunk buffer in F order)
template <int WPS>
__global__ void __launch_bounds__(64 * WPS) __attribute__((amdgpu_waves_per_eu(PJB_GLOB_WAVES)))
jacobian_global_kernel(const BlockArgs* __restrict__ A, int conp, long s0, long count,
                       const double* __restrict__ param, const double* __restrict__ phi,
                       double* __restrict__ cbuf, double* __restrict__ out, Layout L) {
    extern __shared__ double smem[];
    const long i = blockIdx.x;
    if (i >= count) return;
    jacobian_global_state<WPS>(A, conp, s0 + i, param, phi,
                               cbuf + i * (size_t)A->bt.n_contrib,
                               out + i * (size_t)A->m.jac_nnz, L, smem);
}

// chunk buffer [count][nout] (entries of a state contiguous) -> F-order
// output [nout][n], states s0 .. s0 + count - 1; 64 x 64 tiles through LDS so
// that reads and writes are both whole lines.  With nout, n and s0 even every
// access is a 16-byte pair (entries_to_f2_kernel), otherwise single doubles.
__global__ void __launch_bounds__(256)
entries_to_f2_kernel(const double* __restrict__ src, long count, int nout,
                     double* __restrict__ dst, long n, long s0) {
    typedef double v2d __attribute__((ext_vector_type(2)));
    __shared__ double tile[64][65];
    const long t0 = (long)blockIdx.x * 64;
    const int e0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 8 rows per pass
    for (int r = ty; r < 64; r += 8) {
        const long st = t0 + r;
        const int e = e0 + 2 * tx;
        if (st < count && e < nout) {
            const v2d v = *(const v2d*)(src + st * nout + e);  // nout even: e + 1 < nout
            tile[r][2 * tx] = v.x;
            tile[r][2 * tx + 1] = v.y;
        }
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 8) {
        const int e = e0 + r;
        const long st = t0 + 2 * tx;
        if (e < nout && st < count) {
            if (st + 1 < count) {
                *(v2d*)(dst + (size_t)e * n + s0 + st) = v2d{tile[2 * tx][r], tile[2 * tx + 1][r]};
            } else {
                dst[(size_t)e * n + s0 + st] = tile[2 * tx][r];
            }
        }
    }
}

__global__ void __launch_bounds__(256)
entries_to_f_kernel(const double* __restrict__ src, long count, int nout,
                    double* __restrict__ dst, long n, long s0) {
    __shared__ double tile[64][65];
    const long t0 = (long)blockIdx.x * 64;
    const int e0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) {
        const long st = t0 + r;
        const int e = e0 + tx;
        if (st < count && e < nout) tile[r][tx] = src[st * nout + e];
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        const int e = e0 + r;
        const long st = t0 + tx;
        if (st < count && e < nout) dst[(size_t)e * n + s0 + st] = tile[tx][r];
    }
}

// The same transpose in 32 x 32 tiles (8.4 KB of LDS), small enough to run
// beside the LDS-heavy Jacobian kernel on the same CUs: the pipelined F-order
// path transposes one chunk while the next chunk's Jacobians are computed.
// 16-byte pairs throughout (nout, n, s0 even).
__global__ void __launch_bounds__(128)
entries_to_f2_small_kernel(const double* __restrict__ src, long count, int nout,
                           double* __restrict__ dst, long n, long s0) {
    typedef double v2d __attribute__((ext_vector_type(2)));
    __shared__ double tile[32][33];
    const long t0 = (long)blockIdx.x * 32;
    const int e0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;  // 8 rows per pass
    for (int r = ty; r < 32; r += 8) {
        const long st = t0 + r;
        const int e = e0 + 2 * tx;
        if (st < count && e < nout) {
            const v2d v = *(const v2d*)(src + st * nout + e);  // nout even: e + 1 < nout
            tile[r][2 * tx] = v.x;
            tile[r][2 * tx + 1] = v.y;
        }
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int e = e0 + r;
        const long st = t0 + 2 * tx;
        if (e < nout && st < count) {
            if (st + 1 < count) {
                *(v2d*)(dst + (size_t)e * n + s0 + st) = v2d{tile[2 * tx][r], tile[2 * tx + 1][r]};
            } else {
                dst[(size_t)e * n + s0 + st] = tile[2 * tx][r];
            }
        }
    }
}

// The other direction, for the specialised Jacobian module in C order: its
// F-layout store kernel writes a chunk's entries as F-order lines (the entry
// at F-layout position p of chunk state i at src[p * count + i], whole lines
// per wave, non-temporal stores) and this kernel writes them out as C-order
// rows (state s0 + i at dst[(s0 + i) * nout]), C-order entry e read from the
// line of its F-layout position fpos[e].  32 x 32 tiles through LDS: a wave
// reads two 256-byte runs of one entry's states and writes two 256-byte runs
// of one state's entries.  Both arrays are touched once: non-temporal loads
// and stores.
__global__ void __launch_bounds__(256)
entries_to_c_kernel(const double* __restrict__ src, long count, int nout,
                    double* __restrict__ dst, long s0, const int* __restrict__ fpos) {
    __shared__ double tile[32][33];
    const long i0 = (long)blockIdx.x * 32;
    const int e0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 8 rows per pass
    for (int r = ty; r < 32; r += 8) {
        const int e = e0 + r;
        const long i = i0 + tx;
        if (e < nout && i < count)
            tile[r][tx] = __builtin_nontemporal_load(src + (size_t)fpos[e] * count + i);
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const long i = i0 + r;
        const int e = e0 + tx;
        if (i < count && e < nout)
            __builtin_nontemporal_store(tile[tx][r], dst + (size_t)(s0 + i) * nout + e);
    }
}

// Ragged batch: states of several mechanisms / option sets in one launch.
// Member i owns workgroups [block0_i, block0_i + 8 ceil(n_i / 8)); inside
// its range the XCD-aware state order of the single-mechanism launch applies.
struct RaggedMember {
    const BlockArgs* A;
    int conp, out_C;
    long n, block0;
    const double *param, *phi;
    double* out;
    Layout L;
};
struct RaggedArgs {
    int count;
    RaggedMember m[pjk_host::RAGGED_MAX];
};

__device__ __forceinline__ int ragged_member(const RaggedArgs& R, long b) {
    int i = 0;
    for (int k = 1; k < R.count; ++k)
        if (R.m[k].block0 <= b) i = k;
    return i;
}

template <int WPS>
__global__ void __launch_bounds__(64 * WPS) __attribute__((amdgpu_waves_per_eu(PJB_RATES_WAVES)))
rates_ragged_kernel(RaggedArgs R) {
    extern __shared__ double smem[];
    const int i = ragged_member(R, blockIdx.x);
    const RaggedMember& M = R.m[i];
    const long s = state_of_block(blockIdx.x - M.block0, M.n);
    if (s >= M.n) return;
    rates_state<WPS>(M.A, M.conp, M.n, s, M.param, M.phi, M.out, M.L, smem);
}

template <int WPS>
__global__ void __launch_bounds__(64 * WPS) __attribute__((amdgpu_waves_per_eu(PJB_JAC_WAVES)))
jacobian_ragged_kernel(RaggedArgs R) {
    extern __shared__ double smem[];
    const int i = ragged_member(R, blockIdx.x);
    const RaggedMember& M = R.m[i];
    const long s = state_of_block(blockIdx.x - M.block0, M.n);
    if (s >= M.n) return;
    jacobian_state<WPS>(M.A, M.conp, M.n, s, M.param, M.phi, M.out, M.L, M.out_C, smem);
}

// ------------------------------------------ the ragged Jacobian in one launch
// Members of either Jacobian variant -- dense LDS accumulator or the
// contributions gathered in device memory -- share one launch: member i owns
// workgroups [b0_i, b0_i + 8 ceil(n_i / 8)), the gather members first (the
// most expensive states are dispatched first), and inside its range the
// XCD-aware state order of the single-mechanism launch applies.  A gather
// state needs a contribution buffer for as long as its workgroup runs; the
// buffers form a pool with a busy flag per slot, split into one part per XCD
// (read from the XCC_ID hardware register), each part holding twice the
// workgroups that can be resident on that XCD: a workgroup claims a free
// slot of its XCD's part with a compare-and-swap and frees it when it is
// done, so a free slot always exists and the search ends.  A slot is only
// ever reused within one XCD, whose L2 then holds its only copy, so handing
// it on needs no write-back of the L2.  (A persistent loop over work items
// around the per-state code spills 185 VGPRs; one state per workgroup
// spills about as few as the single-mechanism kernels.)
struct QueueMember {
    const BlockArgs* A;
    int conp, out_C, gather;
    long n, b0;                 // states, first workgroup
    const double *param, *phi;
    double* out;
    Layout L;
    long nout;                  // gather: stored values per state
};
struct QueueArgs {
    int count;
    QueueMember m[pjk_host::RAGGED_MAX];
    int* busy;                  // a count of failed claims, then 8 * nslots slot flags
    int nslots;                 // slots per XCD
    double* slots;              // gather members: nslots contribution buffers
    long slot;                  // doubles per slot
};

constexpr long SLOT_PROBES = 1L << 24;  // claim attempts before giving up

template <int WPS>
__global__ void __launch_bounds__(64 * WPS) __attribute__((amdgpu_waves_per_eu(PJB_JAC_WAVES)))
jacobian_queue_kernel(QueueArgs Q) {
    extern __shared__ double smem[];
    __shared__ int slot_sh;
    int mi = 0;
    for (int k = 1; k < Q.count; ++k)
        if (Q.m[k].b0 <= (long)blockIdx.x) mi = k;
    const QueueMember& M = Q.m[mi];
    const long s = state_of_block(blockIdx.x - M.b0, M.n);
    if (s >= M.n) return;
    if (!M.gather) {
        jacobian_state<WPS>(M.A, M.conp, M.n, s, M.param, M.phi, M.out, M.L, M.out_C, smem);
        return;
    }
    if (threadIdx.x == 0) {
        // hwreg(HW_REG_XCC_ID): the XCD this workgroup runs on, 0-7
        const int xcc = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 7;
        int* flags = Q.busy + 1 + (size_t)xcc * Q.nslots;
        int i = (int)((blockIdx.x >> 3) % (unsigned)Q.nslots), got = -1;
        for (long t = 0; t < SLOT_PROBES; ++t) {
            if (atomicCAS(flags + i, 0, 1) == 0) {
                got = xcc * Q.nslots + i;
                break;
            }
            i = i + 1 == Q.nslots ? 0 : i + 1;
        }
        if (got < 0) atomicAdd(Q.busy, 1);
        slot_sh = got;
    }
    __syncthreads();
    const int slot = __builtin_amdgcn_readfirstlane(slot_sh);
    if (slot < 0) return;
    double* cb = Q.slots + (size_t)slot * Q.slot;
    double* out = M.out_C ? M.out + s * M.nout : M.out + s;
    jacobian_global_state<WPS>(M.A, M.conp, s, M.param, M.phi, cb, out, M.L, smem,
                               M.out_C ? 1 : M.n);
    // every store to the buffer done (in this XCD's L2) before the slot is
    // handed on; the next owner reads only what it writes itself
    wait_memory();
    __syncthreads();
    if (threadIdx.x == 0) atomicExch(Q.busy + 1 + slot, 0);
}

}  // namespace pjb

namespace pjk_host {

BlockLds block_lds(const DevMech& m, bool jac, int bt_pairs, bool dense, int bt_long) {
    BlockLds L{};
    int o = 0;
    const int Ns = m.Ns, NN = m.NN;
    auto take = [&](int cnt) {
        int at = o;
        o += cnt > 0 ? cnt : 1;
        return at;
    };
    if (jac && !dense) {
        // device-memory variant: arrays whose lifetimes do not overlap share
        // space.  nmol (until the bath-gas sum) lies under columns T and V|P
        // and the many-contribution entries (from the reaction pass on); the
        // column sums of rows 0 and 1 (gather) lie under the arrays only the
        // reaction pass reads.
        L.conc = take(Ns + 1);
        L.wdot = take(Ns + 1);
        L.energy = take(Ns);
        L.heat = take(Ns);
        L.dheat = take(Ns);
        L.U = take(Ns);
        L.mw = take(Ns - 1);
        L.Cp = take(bt_pairs);
        L.aux = take(2 * NN + bt_long);
        L.nmol = L.aux;
        const int r0 = o;
        L.b = take(Ns + 1);
        L.db = take(Ns + 1);
        L.ci = take(m.n_thd);
        L.ksp = take(m.n_plog + m.n_cheb);
        L.nfac = take(m.n_thd);
        L.dTf = take(m.n_thd);
        L.dEf = take(m.n_thd);
        L.dkT = take(m.n_plog + m.n_cheb);
        L.dkP = take(m.n_plog + m.n_cheb);
        if (o - r0 >= 2 * NN) L.jac = r0;
        else L.jac = take(2 * NN);
        L.total = o;
        return L;
    }
    if (jac) {
        // dense accumulator: the arrays only the pressure-dependent setup
        // and the reaction pass read (third-body factors, PLOG/Chebyshev
        // rate constants) are reused afterwards for the energies, heat
        // capacities and their derivatives; the Gibbs terms b and db lie in
        // rows 0 and 1 of the accumulator until phase 5 writes those rows;
        // no moles or molecular weights in LDS
        L.conc = take(Ns + 1);
        L.wdot = take(Ns + 1);
        L.U = take(Ns);
        L.Cp = take(bt_pairs);
        const int r0 = o;
        L.ci = take(m.n_thd);
        L.ksp = take(m.n_plog + m.n_cheb);
        L.nfac = take(m.n_thd);
        L.dTf = take(m.n_thd);
        L.dEf = take(m.n_thd);
        L.dkT = take(m.n_plog + m.n_cheb);
        L.dkP = take(m.n_plog + m.n_cheb);
        o = r0 + std::max(o - r0, 3 * Ns);
        L.energy = r0;
        L.heat = r0 + Ns;
        L.dheat = r0 + 2 * Ns;
        L.nmol = -1;
        L.mw = -1;
        o = (o + 1) & ~1;  // 16-byte alignment of the accumulator
        L.jac = take(NN * NN);
        L.b = L.jac;
        L.db = L.jac + NN;
        L.total = o;
        return L;
    }
    L.conc = take(Ns + 1);
    L.nmol = take(Ns);
    L.b = take(Ns + 1);
    L.wdot = take(Ns + 1);
    L.energy = take(Ns);
    L.heat = take(Ns);
    L.ci = take(m.n_thd);
    L.ksp = take(m.n_plog + m.n_cheb);
    L.total = o;
    return L;
}

size_t block_lds_bytes(const DevMech& m, bool jac, int n_pairs, bool dense, int n_long) {
    return (size_t)block_lds(m, jac, n_pairs, dense, n_long).total * sizeof(double);
}

template <class K>
static hipError_t set_lds(K kernel, size_t lds) {
    if (lds > 64 * 1024)
        return hipFuncSetAttribute((const void*)kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    return hipSuccess;
}

hipError_t launch_rates_block(const DevMech& m, const BlockArgs* args, const BlockLds& BL, int wps,
                              int conp, char order, long n, const double* param,
                              const double* phi, double* dphi, hipStream_t stream, int occ) {
    pjb::Layout L = order == 'F' ? pjb::Layout{1, (size_t)n} : pjb::Layout{(size_t)m.NN, 1};
    size_t lds = (size_t)BL.total * sizeof(double);
    const long grid = 8 * ((n + 7) / 8);
    if (!occ) occ = PJB_RATES_WAVES;
#define PJB_RATES(W, O)                                                                   \
    if (wps == W && occ == O) {                                                           \
        hipError_t e = set_lds(pjb::rates_block_kernel<W, O>, lds);                      \
        if (e != hipSuccess) return e;                                                    \
        hipLaunchKernelGGL((pjb::rates_block_kernel<W, O>), dim3(grid), dim3(64 * W), lds, \
                           stream, args, conp, n, param, phi, dphi, L);                   \
        return hipGetLastError();                                                         \
    }
    PJB_RATES(1, PJB_RATES_WAVES)
    PJB_RATES(2, PJB_RATES_WAVES)
    PJB_RATES(4, PJB_RATES_WAVES)
    PJB_RATES(8, PJB_RATES_WAVES)
    PJB_RATES(1, 4)
    PJB_RATES(1, 6)
#undef PJB_RATES
    return hipErrorInvalidValue;
}

bool rates_block_has(int wps, int occ) {
    return !occ || occ == PJB_RATES_WAVES ? (wps == 1 || wps == 2 || wps == 4 || wps == 8)
                                          : (wps == 1 && (occ == 4 || occ == 6));
}

bool jacobian_block_has(int wps, int occ) {
    return !occ || occ == PJB_JAC_WAVES ? (wps == 1 || wps == 2 || wps == 4 || wps == 8)
                                        : ((wps == 1 || wps == 2) && (occ == 2 || occ == 4));
}

// chunk of cnt states' entries (C order) -> F-order output columns s0 ..
static hipError_t launch_entries_to_f(const double* chunk, long cnt, int nout, double* jac, long n,
                                      long s0, hipStream_t stream, bool small = false) {
    dim3 grid((unsigned)((cnt + 63) / 64), (unsigned)((nout + 63) / 64));
    const bool pairs = nout % 2 == 0 && n % 2 == 0 && s0 % 2 == 0 &&
                       (uintptr_t)chunk % 16 == 0 && (uintptr_t)jac % 16 == 0;
    if (pairs && small)
        hipLaunchKernelGGL(pjb::entries_to_f2_small_kernel,
                           dim3((unsigned)((cnt + 31) / 32), (unsigned)((nout + 31) / 32)),
                           dim3(128), 0, stream, chunk, cnt, nout, jac, n, s0);
    else if (pairs)
        hipLaunchKernelGGL(pjb::entries_to_f2_kernel, grid, dim3(256), 0, stream, chunk, cnt,
                           nout, jac, n, s0);
    else
        hipLaunchKernelGGL(pjb::entries_to_f_kernel, grid, dim3(256), 0, stream, chunk, cnt, nout,
                           jac, n, s0);
    return hipGetLastError();
}

// states s0 .. s0 + cnt - 1 of F-order inputs with n states, entries in C
// order into out (cnt * nout)
// the Jacobian block kernel of (wps, occ) over n states (a grid of
// 8 ceil(n / 8) workgroups)
static hipError_t launch_jac_kernel(int wps, int occ, size_t lds, long grid, hipStream_t stream,
                                    const BlockArgs* args, int conp, long n, const double* param,
                                    const double* phi, double* out, pjb::Layout L, int out_C) {
    if (!occ) occ = PJB_JAC_WAVES;
#define PJB_JACK(W, O)                                                                    \
    if (wps == W && occ == O) {                                                           \
        hipError_t e = set_lds(pjb::jacobian_block_kernel<W, O>, lds);                   \
        if (e != hipSuccess) return e;                                                    \
        hipLaunchKernelGGL((pjb::jacobian_block_kernel<W, O>), dim3(grid), dim3(64 * W),  \
                           lds, stream, args, conp, n, param, phi, out, L, out_C);        \
        return hipGetLastError();                                                         \
    }
    PJB_JACK(1, PJB_JAC_WAVES)
    PJB_JACK(2, PJB_JAC_WAVES)
    PJB_JACK(4, PJB_JAC_WAVES)
    PJB_JACK(8, PJB_JAC_WAVES)
    PJB_JACK(1, 2)
    PJB_JACK(1, 4)
    PJB_JACK(2, 2)
    PJB_JACK(2, 4)
#undef PJB_JACK
    return hipErrorInvalidValue;
}

// states s0 .. s0 + cnt - 1 of F-order inputs with n states, entries in C
// order into out (cnt * nout)
static hipError_t launch_jacobian_block_chunk(const DevMech& m, const BlockArgs* args,
                                              const BlockLds& BL, int wps, int conp, long n,
                                              long s0, long cnt, const double* param,
                                              const double* phi, double* out,
                                              hipStream_t stream, int occ) {
    pjb::Layout L{1, (size_t)n};
    size_t lds = (size_t)BL.total * sizeof(double);
    const long grid = 8 * ((cnt + 7) / 8);
    return launch_jac_kernel(wps, occ, lds, grid, stream, args, conp, cnt, param + s0, phi + s0,
                             out, L, 1);
}

hipError_t launch_entries_to_c(const double* src, long count, int nout, double* jac, long s0,
                               const int* fpos, hipStream_t stream) {
    if (count <= 0 || nout <= 0) return hipSuccess;
    dim3 grid((unsigned)((count + 31) / 32), (unsigned)((nout + 31) / 32));
    hipLaunchKernelGGL(pjb::entries_to_c_kernel, grid, dim3(256), 0, stream, src, count, nout,
                       jac, s0, fpos);
    return hipGetLastError();
}

hipError_t launch_jacobian_block(const DevMech& m, const BlockArgs* args, const BlockLds& BL,
                                 int wps, int conp, char order, long n, const double* param,
                                 const double* phi, double* jac, double* chunk_buf,
                                 long chunk_states, hipStream_t stream, const FPipeline* pipe,
                                 int occ) {
    if (order == 'F' && chunk_buf && pipe && n > chunk_states) {
        // two chunk buffers: the transpose of chunk k (on the second stream)
        // runs while chunk k + 1's Jacobians are computed; a buffer is
        // reused once its transpose has finished, and the caller's stream
        // waits for the last transpose
        if (chunk_states < 64 || chunk_states % 64) return hipErrorInvalidValue;
        const int nout = m.jac_nnz;
        long k = 0;
        for (long s0 = 0; s0 < n; s0 += chunk_states, ++k) {
            const long cnt = std::min(chunk_states, n - s0);
            const int b = (int)(k & 1);
            double* buf = chunk_buf + (size_t)b * chunk_states * nout;
            hipError_t e = hipSuccess;
            if (k >= 2) e = hipStreamWaitEvent(stream, pipe->transposed[b], 0);
            if (e == hipSuccess)
                e = launch_jacobian_block_chunk(m, args, BL, wps, conp, n, s0, cnt, param, phi,
                                                buf, stream, occ);
            if (e == hipSuccess) e = hipEventRecord(pipe->computed[b], stream);
            if (e == hipSuccess) e = hipStreamWaitEvent(pipe->aux, pipe->computed[b], 0);
            if (e == hipSuccess) e = launch_entries_to_f(buf, cnt, nout, jac, n, s0, pipe->aux, true);
            if (e == hipSuccess) e = hipEventRecord(pipe->transposed[b], pipe->aux);
            if (e != hipSuccess) return e;
        }
        return hipStreamWaitEvent(stream, pipe->transposed[(k - 1) & 1], 0);
    }
    if (order == 'F' && chunk_buf) {
        // F order through C-order chunks: each state's entries are written as
        // one contiguous run, then a tiled transpose writes whole lines
        if (chunk_states < 64 || chunk_states % 64) return hipErrorInvalidValue;
        const int nout = m.jac_nnz;
        for (long s0 = 0; s0 < n; s0 += chunk_states) {
            const long cnt = std::min(chunk_states, n - s0);
            hipError_t e = launch_jacobian_block_chunk(m, args, BL, wps, conp, n, s0, cnt, param,
                                                       phi, chunk_buf, stream, occ);
            if (e != hipSuccess) return e;
            e = launch_entries_to_f(chunk_buf, cnt, nout, jac, n, s0, stream);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    pjb::Layout L = order == 'F' ? pjb::Layout{1, (size_t)n} : pjb::Layout{(size_t)m.NN, 1};
    size_t lds = (size_t)BL.total * sizeof(double);
    const long grid = 8 * ((n + 7) / 8);
    return launch_jac_kernel(wps, occ, lds, grid, stream, args, conp, n, param, phi, jac, L,
                             order == 'C');
}

hipError_t launch_jacobian_global(const DevMech& m, const BlockArgs* args, const BlockLds& BL,
                                  int wps, int conp, char order, long n, const double* param,
                                  const double* phi, double* jac, double* chunk_buf,
                                  long chunk_states, size_t n_contrib, hipStream_t stream) {
    pjb::Layout L = order == 'F' ? pjb::Layout{1, (size_t)n} : pjb::Layout{(size_t)m.NN, 1};
    size_t lds = (size_t)BL.total * sizeof(double);
    const int nout = m.jac_nnz;
    if (!chunk_buf || chunk_states < 64 || chunk_states % 64) return hipErrorInvalidValue;
    // chunk buffer: contributions of chunk_states states, then (F order)
    // their entries
    double* cbuf = chunk_buf;
    double* ebuf = chunk_buf + (size_t)chunk_states * n_contrib;
    for (long s0 = 0; s0 < n; s0 += chunk_states) {
        const long cnt = std::min(chunk_states, n - s0);
        double* out = order == 'F' ? ebuf : jac + (size_t)s0 * nout;
#define PJB_GJ(W)                                                                         \
    case W: {                                                                             \
        hipError_t e = set_lds(pjb::jacobian_global_kernel<W>, lds);                     \
        if (e != hipSuccess) return e;                                                    \
        hipLaunchKernelGGL(pjb::jacobian_global_kernel<W>, dim3(cnt), dim3(64 * W), lds,  \
                           stream, args, conp, s0, cnt, param, phi, cbuf, out, L);        \
        break;                                                                            \
    }
        switch (wps) {
            PJB_GJ(1)
            PJB_GJ(2)
            PJB_GJ(4)
            PJB_GJ(8)
            default: return hipErrorInvalidValue;
        }
#undef PJB_GJ
        hipError_t e = hipGetLastError();
        if (e == hipSuccess && order == 'F') e = launch_entries_to_f(ebuf, cnt, nout, jac, n, s0, stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_queue(int wps, int count, const QueueLaunch* members, size_t lds, int* busy,
                        int nslots, double* slots, long slot, hipStream_t stream) {
    if (count < 1 || count > RAGGED_MAX || nslots < 1) return hipErrorInvalidValue;
    pjb::QueueArgs Q{};
    Q.count = count;
    Q.busy = busy;
    Q.nslots = nslots;
    Q.slots = slots;
    Q.slot = slot;
    long blocks = 0;
    for (int i = 0; i < count; ++i) {
        const QueueLaunch& q = members[i];
        pjb::Layout L = q.order == 'F' ? pjb::Layout{1, (size_t)q.n}
                                       : pjb::Layout{(size_t)q.NN, 1};
        Q.m[i] = pjb::QueueMember{q.args, q.conp, q.order == 'C', q.gather, q.n, blocks, q.param,
                                  q.phi, q.out, L, q.nout};
        blocks += 8 * ((q.n + 7) / 8);
    }
    if (blocks == 0) return hipSuccess;
#define PJB_Q(W)                                                                           \
    case W: {                                                                              \
        hipError_t e = set_lds(pjb::jacobian_queue_kernel<W>, lds);                        \
        if (e != hipSuccess) return e;                                                     \
        hipLaunchKernelGGL(pjb::jacobian_queue_kernel<W>, dim3(blocks), dim3(64 * W), lds, \
                           stream, Q);                                                     \
        break;                                                                             \
    }
    switch (wps) {
        PJB_Q(2)
        PJB_Q(4)
        PJB_Q(8)
        default: return hipErrorInvalidValue;
    }
#undef PJB_Q
    return hipGetLastError();
}

int queue_resident(int wps, size_t lds) {
    int nb = 0;
    hipError_t e = hipSuccess;
#define PJB_QO(W)                                                                         \
    case W:                                                                               \
        e = set_lds(pjb::jacobian_queue_kernel<W>, lds);                                  \
        if (e == hipSuccess)                                                              \
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(                             \
                &nb, pjb::jacobian_queue_kernel<W>, 64 * W, lds);                         \
        break;
    switch (wps) {
        PJB_QO(2)
        PJB_QO(4)
        PJB_QO(8)
        default: return 0;
    }
#undef PJB_QO
    return e == hipSuccess ? nb : 0;
}

hipError_t launch_ragged(int kernel, int wps, int count, const RaggedLaunch* members,
                         size_t lds, hipStream_t stream) {
    if (count < 1 || count > RAGGED_MAX) return hipErrorInvalidValue;
    pjb::RaggedArgs R{};
    R.count = count;
    long blocks = 0;
    for (int i = 0; i < count; ++i) {
        const RaggedLaunch& q = members[i];
        pjb::Layout L = q.order == 'F' ? pjb::Layout{1, (size_t)q.n}
                                       : pjb::Layout{(size_t)q.NN, 1};
        R.m[i] = pjb::RaggedMember{q.args, q.conp, q.order == 'C', q.n, blocks, q.param, q.phi,
                                   q.out, L};
        blocks += 8 * ((q.n + 7) / 8);
    }
    if (blocks == 0) return hipSuccess;
#define PJB_RAG(W)                                                                        \
    case W: {                                                                             \
        if (kernel == 0) {                                                                \
            hipError_t e = set_lds(pjb::rates_ragged_kernel<W>, lds);                    \
            if (e != hipSuccess) return e;                                                \
            hipLaunchKernelGGL(pjb::rates_ragged_kernel<W>, dim3(blocks), dim3(64 * W),   \
                               lds, stream, R);                                           \
        } else {                                                                          \
            hipError_t e = set_lds(pjb::jacobian_ragged_kernel<W>, lds);                 \
            if (e != hipSuccess) return e;                                                \
            hipLaunchKernelGGL(pjb::jacobian_ragged_kernel<W>, dim3(blocks), dim3(64 * W), \
                               lds, stream, R);                                           \
        }                                                                                 \
        break;                                                                            \
    }
    switch (wps) {
        PJB_RAG(1)
        PJB_RAG(2)
        PJB_RAG(4)
        PJB_RAG(8)
        default: return hipErrorInvalidValue;
    }
#undef PJB_RAG
    return hipGetLastError();
}

}  // namespace pjk_host
