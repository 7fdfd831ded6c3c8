// Batched ROS2 integration of the species-rate ODE (include/pyjac_integrator.h).
//
// Every state carries its own time, step and status.  One step attempt of a
// batch is five launches: species rates f(y) and the full C-order Jacobian J
// (this library's kernels), ros2_stage1_kernel (I - gamma h J in LDS, LU with
// partial pivoting, k1, the stage state y + h k1), species rates at the stage
// state, and ros2_stage2_kernel (k2, error estimate, accept or reject, next
// step).  States that are done are skipped by the stage kernels; the host
// stops when no state is active.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/pyjac_integrator.h"

namespace pjros {

constexpr int NT = 256;        // threads per state
constexpr int NMAX = 120;      // largest NN (the iteration matrix in LDS)
constexpr double GAMMA = 1.7071067811865475244;  // 1 + 1/sqrt(2)

enum Status { ACTIVE = 0, DONE = 1, FAILED = 2 };

struct Chunk {
    long n;
    int NN;
    const double* param;
    double* y;       // [n][NN] the states, updated in place
    double* f0;      // [n][NN] rates at y
    double* J;       // [n][NN][NN] Jacobian, then its LU factors
    int* piv;        // [n][NN] pivot rows
    double* k1;      // [n][NN]
    double* y1;      // [n][NN] stage state
    double* f1;      // [n][NN] rates at the stage state
    double* t;       // [n]
    double* h;       // [n]
    int* status;     // [n]
    int* attempts;   // [n]
    int* accepted;   // [n]
    int* rejected;   // [n]
    int* rejlast;    // [n] the last attempt was rejected
    int* active;     // [1] states still active after the last attempt
};

struct Params {
    double t_end, rtol, atol, h0;
    int fixed;
    long max_steps;
};

__device__ __forceinline__ double wave_sum_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ double wave_max_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// x = (LU)^-1 P b for one state, by wave 0: lane l holds rows l and l + 64
// (NN <= 128); A is the LU factor in LDS (unit lower L below the diagonal),
// piv the row exchanges of the factorisation, in order
__device__ void lu_solve_wave(const double* A, const int* piv, int N, int lane, double& b0,
                              double& b1) {
    auto get = [&](int r) { return r < 64 ? __shfl(b0, r) : __shfl(b1, r - 64); };
    for (int k = 0; k < N; ++k) {
        const int p = piv[k];
        if (p != k) {
            const double vk = get(k), vp = get(p);
            if (lane == (k & 63)) { if (k < 64) b0 = vp; else b1 = vp; }
            if (lane == (p & 63)) { if (p < 64) b0 = vk; else b1 = vk; }
        }
    }
    const int r0 = lane, r1 = lane + 64;
    for (int k = 0; k < N; ++k) {  // L y = P b
        const double xk = get(k);
        if (r0 > k && r0 < N) b0 -= A[r0 * N + k] * xk;
        if (r1 > k && r1 < N) b1 -= A[r1 * N + k] * xk;
    }
    for (int k = N - 1; k >= 0; --k) {  // U x = y
        const double xk = get(k) / A[k * N + k];
        if (lane == (k & 63)) { if (k < 64) b0 = xk; else b1 = xk; }
        if (r0 < k) b0 -= A[r0 * N + k] * xk;
        if (r1 < k && r1 < N) b1 -= A[r1 * N + k] * xk;
    }
}

// I - gamma h J -> LU (partial pivoting) in LDS, k1, stage state
__global__ void __launch_bounds__(NT) ros2_stage1_kernel(Chunk c, Params P) {
    extern __shared__ double A[];  // [N][N]
    __shared__ int pv[NMAX];
    __shared__ double hs_sh;
    const long s = blockIdx.x;
    if (s >= c.n) return;
    const int N = c.NN, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double* y = c.y + s * N;
    double* y1 = c.y1 + s * N;
    if (c.status[s] != ACTIVE) {
        for (int i = tid; i < N; i += NT) y1[i] = y[i];
        return;
    }
    const double* f0 = c.f0 + s * N;
    if (wave == 0) {
        double h = c.h[s];
        if (h <= 0.0) {
            // first step: 1e-2 of the time the initial rates take to move any
            // component by its error weight
            double r = 0.0;
            for (int i = lane; i < N; i += 64)
                r = fmax(r, fabs(f0[i]) / (P.atol + P.rtol * fabs(y[i])));
            r = wave_max_d(r);
            h = r > 0.0 ? 1e-2 / r : P.t_end;
            h = fmin(h, P.t_end - c.t[s]);
            if (lane == 0) c.h[s] = h;
        }
        if (lane == 0) hs_sh = h;
    }
    __syncthreads();
    const double gh = -GAMMA * hs_sh;
    const double* J = c.J + s * (size_t)N * N;
    for (int e = tid; e < N * N; e += NT) {
        const int i = e / N, j = e - i * N;
        A[e] = gh * J[e] + (i == j ? 1.0 : 0.0);
    }
    __syncthreads();
    for (int k = 0; k < N; ++k) {
        if (wave == 0) {
            double best = -1.0;
            int bi = k;
            for (int i = k + lane; i < N; i += 64) {
                const double v = fabs(A[i * N + k]);
                if (v > best) { best = v; bi = i; }
            }
            for (int o = 32; o > 0; o >>= 1) {
                const double ob = __shfl_xor(best, o);
                const int oi = __shfl_xor(bi, o);
                if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
            }
            if (lane == 0) pv[k] = bi;
        }
        __syncthreads();
        const int p = pv[k];
        if (p != k)
            for (int j = tid; j < N; j += NT) {
                const double tmp = A[k * N + j];
                A[k * N + j] = A[p * N + j];
                A[p * N + j] = tmp;
            }
        __syncthreads();
        const double pivot = A[k * N + k];
        for (int i = k + 1 + tid; i < N; i += NT) A[i * N + k] /= pivot;
        __syncthreads();
        const int m = N - k - 1;
        for (int e = tid; e < m * m; e += NT) {
            const int i = k + 1 + e / m, j = k + 1 + e % m;
            A[i * N + j] -= A[i * N + k] * A[k * N + j];
        }
        __syncthreads();
    }
    double* LU = c.J + s * (size_t)N * N;
    for (int e = tid; e < N * N; e += NT) LU[e] = A[e];
    for (int i = tid; i < N; i += NT) c.piv[s * N + i] = pv[i];
    if (wave == 0) {
        double b0 = lane < N ? f0[lane] : 0.0;
        double b1 = lane + 64 < N ? f0[lane + 64] : 0.0;
        lu_solve_wave(A, pv, N, lane, b0, b1);
        const double h = hs_sh;
        double* k1 = c.k1 + s * N;
        if (lane < N) { k1[lane] = b0; y1[lane] = y[lane] + h * b0; }
        if (lane + 64 < N) { k1[lane + 64] = b1; y1[lane + 64] = y[lane + 64] + h * b1; }
    }
}

// k2, error estimate, step decision
__global__ void __launch_bounds__(NT) ros2_stage2_kernel(Chunk c, Params P) {
    extern __shared__ double A[];
    __shared__ int pv[NMAX];
    const long s = blockIdx.x;
    if (s >= c.n || c.status[s] != ACTIVE) return;
    const int N = c.NN, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double* LU = c.J + s * (size_t)N * N;
    for (int e = tid; e < N * N; e += NT) A[e] = LU[e];
    for (int i = tid; i < N; i += NT) pv[i] = c.piv[s * N + i];
    __syncthreads();
    if (wave != 0) return;
    const double* k1 = c.k1 + s * N;
    const double* f1 = c.f1 + s * N;
    double* y = c.y + s * N;
    const double k10 = lane < N ? k1[lane] : 0.0;
    const double k11 = lane + 64 < N ? k1[lane + 64] : 0.0;
    double b0 = lane < N ? f1[lane] - 2.0 * k10 : 0.0;
    double b1 = lane + 64 < N ? f1[lane + 64] - 2.0 * k11 : 0.0;
    lu_solve_wave(A, pv, N, lane, b0, b1);
    const double h = c.h[s];
    double yn0 = 0.0, yn1 = 0.0, sq = 0.0;
    if (lane < N) {
        yn0 = y[lane] + 1.5 * h * k10 + 0.5 * h * b0;
        const double e = 0.5 * h * (k10 + b0);
        const double w = P.atol + P.rtol * fmax(fabs(y[lane]), fabs(yn0));
        sq += (e / w) * (e / w);
    }
    if (lane + 64 < N) {
        yn1 = y[lane + 64] + 1.5 * h * k11 + 0.5 * h * b1;
        const double e = 0.5 * h * (k11 + b1);
        const double w = P.atol + P.rtol * fmax(fabs(y[lane + 64]), fabs(yn1));
        sq += (e / w) * (e / w);
    }
    const double err = sqrt(wave_sum_d(sq) / N);
    const bool finite = err == err && err < 1e300;
    const bool accept = P.fixed ? finite : (finite && err <= 1.0);
    if (accept) {
        if (lane < N) y[lane] = yn0;
        if (lane + 64 < N) y[lane + 64] = yn1;
    }
    if (lane == 0) {
        double t = c.t[s], hn = h;
        if (accept) {
            t += h;
            c.accepted[s] += 1;
            if (!P.fixed) {
                double fac = 0.9 / sqrt(fmax(err, 1e-10));
                fac = fmin(c.rejlast[s] ? 1.0 : 5.0, fmax(0.2, fac));
                hn = h * fac;
            }
            c.rejlast[s] = 0;
        } else {
            c.rejected[s] += 1;
            if (!P.fixed) hn = h * (finite ? fmax(0.2, 0.9 / sqrt(err)) : 0.2);
            c.rejlast[s] = 1;
        }
        const int att = c.attempts[s] + 1;
        c.attempts[s] = att;
        int st = ACTIVE;
        if (P.t_end - t <= 1e-12 * P.t_end) {
            st = DONE;
            t = P.t_end;
        } else {
            hn = fmin(hn, P.t_end - t);
            if (att >= P.max_steps) st = FAILED;
        }
        c.t[s] = t;
        c.h[s] = hn;
        c.status[s] = st;
        if (st == ACTIVE) atomicAdd(c.active, 1);
    }
}

__global__ void ros2_init_kernel(Chunk c, Params P) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= c.n) return;
    c.t[s] = 0.0;
    c.h[s] = P.h0 > 0.0 ? fmin(P.h0, P.t_end) : 0.0;
    c.status[s] = P.t_end > 0.0 ? ACTIVE : DONE;
    c.attempts[s] = 0;
    c.accepted[s] = 0;
    c.rejected[s] = 0;
    c.rejlast[s] = 0;
}

__global__ void ros2_zero_kernel(int* p) {
    if (threadIdx.x == 0) p[0] = 0;
}

}  // namespace pjros

struct pyjac_integrator {
    pyjac_handle* h = nullptr;  // full C-order Jacobian and the species rates
    int NN = 0;
    char* buf = nullptr;
    size_t buf_bytes = 0;
    int* d_active = nullptr;
    int* h_active = nullptr;  // pinned
};

namespace {

size_t per_state_bytes(int NN) {
    const size_t N = NN;
    return (N * N + 5 * N + 2) * sizeof(double) + (N + 5) * sizeof(int);
}

}  // namespace

extern "C" {

int pyjac_hip_integrator_create(const pyjac_mech_tables* tables, int conp, pyjac_integrator** out) {
    if (!tables || !out) return PYJAC_ERR_ARG;
    *out = nullptr;
    const int NN = tables->n_species + 1;
    if (NN > pjros::NMAX) return PYJAC_ERR_UNSUPPORTED;
    pyjac_integrator* it = new pyjac_integrator();
    it->NN = NN;
    int rc = pyjac_hip_create(tables, conp, 'C', PYJAC_JAC_FULL, PYJAC_JAC_EXACT, &it->h);
    if (rc) {
        delete it;
        return rc;
    }
    if (hipMalloc(&it->d_active, sizeof(int)) != hipSuccess ||
        hipHostMalloc(&it->h_active, sizeof(int), hipHostMallocDefault) != hipSuccess) {
        pyjac_hip_integrator_destroy(it);
        return PYJAC_ERR_ALLOC;
    }
    *out = it;
    return PYJAC_OK;
}

int pyjac_hip_integrator_destroy(pyjac_integrator* it) {
    if (!it) return PYJAC_OK;
    if (it->h) pyjac_hip_destroy(it->h);
    (void)hipFree(it->buf);
    (void)hipFree(it->d_active);
    if (it->h_active) (void)hipHostFree(it->h_active);
    delete it;
    return PYJAC_OK;
}

int pyjac_hip_integrate_dev(pyjac_integrator* it, size_t n, const double* d_param, double* d_phi,
                            double t_end, double rtol, double atol, double h0, int fixed,
                            long max_steps, void* stream, pyjac_integrate_stats* stats) {
    if (!it || (n && (!d_param || !d_phi)) || !(t_end >= 0.0) || !(rtol >= 0.0) ||
        !(atol > 0.0) || max_steps < 1 || (fixed && !(h0 > 0.0)))
        return PYJAC_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const int N = it->NN;
    const size_t per = per_state_bytes(N);
    const long cs = (long)std::max<size_t>(1, std::min<size_t>(n, ((size_t)2 << 30) / per));
    if (!it->buf || it->buf_bytes < (size_t)cs * per) {
        (void)hipFree(it->buf);
        it->buf = nullptr;
        it->buf_bytes = 0;
        if (hipMalloc(&it->buf, (size_t)cs * per) != hipSuccess) return PYJAC_ERR_ALLOC;
        it->buf_bytes = (size_t)cs * per;
    }
    const size_t lds = (size_t)N * N * sizeof(double);
    if (lds > 64 * 1024) {
        if (hipFuncSetAttribute((const void*)pjros::ros2_stage1_kernel,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess ||
            hipFuncSetAttribute((const void*)pjros::ros2_stage2_kernel,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
            return PYJAC_ERR_HIP;
    }
    pyjac_integrate_stats st{0, 0, 0, 0, 0};
    pjros::Params P{t_end, rtol, atol, h0, fixed, max_steps};
    std::vector<int> h_att, h_acc, h_rej, h_status;
    for (size_t s0 = 0; s0 < n; s0 += cs) {
        const long cnt = (long)std::min<size_t>(cs, n - s0);
        pjros::Chunk c{};
        c.n = cnt;
        c.NN = N;
        c.param = d_param + s0;
        c.y = d_phi + s0 * N;
        double* d = (double*)it->buf;
        c.f0 = d;  d += (size_t)cnt * N;
        c.J = d;   d += (size_t)cnt * N * N;
        c.k1 = d;  d += (size_t)cnt * N;
        c.y1 = d;  d += (size_t)cnt * N;
        c.f1 = d;  d += (size_t)cnt * N;
        c.t = d;   d += cnt;
        c.h = d;   d += cnt;
        int* ip = (int*)d;
        c.piv = ip;      ip += (size_t)cnt * N;
        c.status = ip;   ip += cnt;
        c.attempts = ip; ip += cnt;
        c.accepted = ip; ip += cnt;
        c.rejected = ip; ip += cnt;
        c.rejlast = ip;  ip += cnt;
        c.active = it->d_active;
        hipLaunchKernelGGL(pjros::ros2_init_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256),
                           0, s, c, P);
        for (long att = 0; att < max_steps; ++att) {
            int rc = pyjac_hip_species_rates_dev(it->h, cnt, c.param, c.y, c.f0, s);
            if (!rc) rc = pyjac_hip_jacobian_dev(it->h, cnt, c.param, c.y, c.J, s);
            if (rc) return rc;
            hipLaunchKernelGGL(pjros::ros2_stage1_kernel, dim3((unsigned)cnt), dim3(pjros::NT),
                               lds, s, c, P);
            rc = pyjac_hip_species_rates_dev(it->h, cnt, c.param, c.y1, c.f1, s);
            if (rc) return rc;
            hipLaunchKernelGGL(pjros::ros2_zero_kernel, dim3(1), dim3(64), 0, s, c.active);
            hipLaunchKernelGGL(pjros::ros2_stage2_kernel, dim3((unsigned)cnt), dim3(pjros::NT),
                               lds, s, c, P);
            if (hipGetLastError() != hipSuccess) return PYJAC_ERR_HIP;
            ++st.launches;
            if (hipMemcpyAsync(it->h_active, c.active, sizeof(int), hipMemcpyDeviceToHost, s) !=
                    hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess)
                return PYJAC_ERR_HIP;
            if (*it->h_active == 0) break;
        }
        h_att.resize(cnt);
        h_acc.resize(cnt);
        h_rej.resize(cnt);
        h_status.resize(cnt);
        if (hipMemcpy(h_att.data(), c.attempts, cnt * sizeof(int), hipMemcpyDeviceToHost) ||
            hipMemcpy(h_acc.data(), c.accepted, cnt * sizeof(int), hipMemcpyDeviceToHost) ||
            hipMemcpy(h_rej.data(), c.rejected, cnt * sizeof(int), hipMemcpyDeviceToHost) ||
            hipMemcpy(h_status.data(), c.status, cnt * sizeof(int), hipMemcpyDeviceToHost))
            return PYJAC_ERR_HIP;
        for (long i = 0; i < cnt; ++i) {
            st.accepted += h_acc[i];
            st.rejected += h_rej[i];
            st.max_steps = std::max<long>(st.max_steps, h_att[i]);
            st.failed += h_status[i] != pjros::DONE;
        }
    }
    if (stats) *stats = st;
    return PYJAC_OK;
}

}  // extern "C"
