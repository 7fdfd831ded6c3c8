// Batched ROS2 integration of the species-rate ODE (include/pyjac_integrator.h).
//
// Every state carries its own time, step and status.  One step attempt of a
// batch is five launches: species rates f(y) and the full C-order Jacobian J
// (this library's kernels), ros2_stage1_kernel (I - gamma h J in LDS, LU with
// partial pivoting, k1, the stage state y + h k1), species rates at the stage
// state, and ros2_stage2_kernel (k2, error estimate, accept or reject, next
// step).  States that are done are skipped by the stage kernels; the host
// stops when no state is active.
//
// The batch works on a compact copy of its states.  Whenever a sixteenth of
// the states in it have finished, ros2_compact_kernel writes the finished
// ones back to the caller's array (adding their counts to the statistics)
// and packs the active ones into the other copy, so that the later attempts
// -- often many more for the stiffest few states -- launch only over the
// states still integrating.
//
// The iteration matrix I - gamma h J lives in LDS for NN up to 120; for the
// larger mechanisms (up to 512) it is formed and factorised in device memory,
// in place of the state's Jacobian, by the same workgroup, and the solves
// hold ceil(NN / 64) rows per lane.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/pyjac_integrator.h"

namespace pjros {

constexpr int NT = 256;        // threads per state
constexpr int NMAX = 120;      // largest NN with the iteration matrix in LDS
constexpr int NGMAX = 512;     // largest NN with the iteration matrix in device memory
constexpr double GAMMA = 1.7071067811865475244;  // 1 + 1/sqrt(2)

enum Status { ACTIVE = 0, DONE = 1, FAILED = 2 };

// the per-state data that moves with a state when the batch is compacted
struct StateSet {
    double* y;       // [n][NN]
    double* param;   // [n]
    double* t;       // [n]
    double* h;       // [n]
    int* status;     // [n]
    int* attempts;   // [n]
    int* accepted;   // [n]
    int* rejected;   // [n]
    int* rejlast;    // [n]
    int* orig;       // [n] the state's index in the caller's array
};

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

// The same solve for N <= 64 R, lane l holding rows l + 64 r (r < R), for
// the iteration matrix of the larger mechanisms (in device memory)
template <int R>
__device__ void lu_solve_wave_r(const double* A, const int* piv, int N, int lane, double (&b)[R]) {
    auto get = [&](int row) {
        double v = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const double x = __shfl(b[r], row & 63);
            if ((row >> 6) == r) v = x;
        }
        return v;
    };
    auto put = [&](int row, double v) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (lane == (row & 63) && (row >> 6) == r) b[r] = v;
    };
    for (int k = 0; k < N; ++k) {
        const int p = piv[k];
        if (p != k) {
            const double vk = get(k), vp = get(p);
            put(k, vp);
            put(p, vk);
        }
    }
    for (int k = 0; k < N; ++k) {  // L y = P b
        const double xk = get(k);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = lane + 64 * r;
            if (row > k && row < N) b[r] -= A[(size_t)row * N + k] * xk;
        }
    }
    for (int k = N - 1; k >= 0; --k) {  // U x = y
        const double xk = get(k) / A[(size_t)k * N + k];
        put(k, xk);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = lane + 64 * r;
            if (row < k) b[r] -= A[(size_t)row * N + k] * xk;
        }
    }
}

// Gaussian elimination with partial pivoting of the N x N matrix A (LDS or
// device memory, this workgroup's own): pv[k] the row exchanged with row k,
// L (unit, below the diagonal) and U in place
__device__ void lu_factor_wg(double* A, int* pv, int N, int tid, int lane, int wave) {
    for (int k = 0; k < N; ++k) {
        if (wave == 0) {
            double best = -1.0;
            int bi = k;
            for (int i = k + lane; i < N; i += 64) {
                const double v = fabs(A[(size_t)i * N + k]);
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
                const double tmp = A[(size_t)k * N + j];
                A[(size_t)k * N + j] = A[(size_t)p * N + j];
                A[(size_t)p * N + j] = tmp;
            }
        __syncthreads();
        const double pivot = A[(size_t)k * N + k];
        for (int i = k + 1 + tid; i < N; i += NT) A[(size_t)i * N + k] /= pivot;
        __syncthreads();
        const int m = N - k - 1;
        for (int e = tid; e < m * m; e += NT) {
            const int i = k + 1 + e / m, j = k + 1 + e % m;
            A[(size_t)i * N + j] -= A[(size_t)i * N + k] * A[(size_t)k * N + j];
        }
        __syncthreads();
    }
}

// first-step size of a state (h <= 0: 1e-2 of the time the initial rates
// take to move any component by its error weight), by wave 0
__device__ void first_step(const Chunk& c, const Params& P, long s, const double* y,
                           const double* f0, int lane, double* hs_sh) {
    const int N = c.NN;
    double h = c.h[s];
    if (h <= 0.0) {
        double r = 0.0;
        for (int i = lane; i < N; i += 64)
            r = fmax(r, fabs(f0[i]) / (P.atol + P.rtol * fabs(y[i])));
        r = wave_max_d(r);
        h = r > 0.0 ? 1e-2 / r : P.t_end;
        h = fmin(h, P.t_end - c.t[s]);
        if (lane == 0) c.h[s] = h;
    }
    if (lane == 0) *hs_sh = h;
}

// Mechanisms beyond NMAX: I - gamma h J formed in place of J in device
// memory, factorised there (this workgroup's own rows, visible to it after
// each barrier), k1 and the stage state; R rows per lane in the solve
template <int R>
__global__ void __launch_bounds__(NT) ros2_stage1_global_kernel(Chunk c, Params P) {
    __shared__ int pv[NGMAX];
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
    if (wave == 0) first_step(c, P, s, y, f0, lane, &hs_sh);
    __syncthreads();
    const double gh = -GAMMA * hs_sh;
    double* A = c.J + s * (size_t)N * N;
    for (int e = tid; e < N * N; e += NT) {
        const int i = e / N, j = e - i * N;
        A[e] = gh * A[e] + (i == j ? 1.0 : 0.0);
    }
    __syncthreads();
    lu_factor_wg(A, pv, N, tid, lane, wave);
    for (int i = tid; i < N; i += NT) c.piv[s * N + i] = pv[i];
    if (wave == 0) {
        double b[R];
#pragma unroll
        for (int r = 0; r < R; ++r) b[r] = lane + 64 * r < N ? f0[lane + 64 * r] : 0.0;
        lu_solve_wave_r<R>(A, pv, N, lane, b);
        const double h = hs_sh;
        double* k1 = c.k1 + s * N;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = lane + 64 * r;
            if (row < N) { k1[row] = b[r]; y1[row] = y[row] + h * b[r]; }
        }
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
    if (wave == 0) first_step(c, P, s, y, f0, lane, &hs_sh);
    __syncthreads();
    const double gh = -GAMMA * hs_sh;
    const double* J = c.J + s * (size_t)N * N;
    for (int e = tid; e < N * N; e += NT) {
        const int i = e / N, j = e - i * N;
        A[e] = gh * J[e] + (i == j ? 1.0 : 0.0);
    }
    __syncthreads();
    lu_factor_wg(A, pv, N, tid, lane, wave);
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

__device__ void step_decision(const Chunk& c, const Params& P, long s, double h, double err,
                              bool finite, bool accept);

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
    if (lane == 0) step_decision(c, P, s, h, err, finite, accept);
}

// accept or reject the attempt of state s, the next step size and status
__device__ void step_decision(const Chunk& c, const Params& P, long s, double h, double err,
                              bool finite, bool accept) {
    {
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

// k2 from the factors in device memory, error estimate, step decision
template <int R>
__global__ void __launch_bounds__(64) ros2_stage2_global_kernel(Chunk c, Params P) {
    __shared__ int pv[NGMAX];
    const long s = blockIdx.x;
    if (s >= c.n || c.status[s] != ACTIVE) return;
    const int N = c.NN, lane = threadIdx.x;
    const double* A = c.J + s * (size_t)N * N;
    for (int i = lane; i < N; i += 64) pv[i] = c.piv[s * N + i];
    __syncthreads();
    const double* k1 = c.k1 + s * N;
    const double* f1 = c.f1 + s * N;
    double* y = c.y + s * N;
    double kv[R], b[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int row = lane + 64 * r;
        kv[r] = row < N ? k1[row] : 0.0;
        b[r] = row < N ? f1[row] - 2.0 * kv[r] : 0.0;
    }
    lu_solve_wave_r<R>(A, pv, N, lane, b);
    const double h = c.h[s];
    double yn[R], sq = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int row = lane + 64 * r;
        yn[r] = 0.0;
        if (row < N) {
            yn[r] = y[row] + 1.5 * h * kv[r] + 0.5 * h * b[r];
            const double e = 0.5 * h * (kv[r] + b[r]);
            const double w = P.atol + P.rtol * fmax(fabs(y[row]), fabs(yn[r]));
            sq += (e / w) * (e / w);
        }
    }
    const double err = sqrt(wave_sum_d(sq) / N);
    const bool finite = err == err && err < 1e300;
    const bool accept = P.fixed ? finite : (finite && err <= 1.0);
    if (accept) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (lane + 64 * r < N) y[lane + 64 * r] = yn[r];
    }
    if (lane == 0) step_decision(c, P, s, h, err, finite, accept);
}

// the working copy from the caller's states (one workgroup of 64 per state)
__global__ void __launch_bounds__(64) ros2_init_kernel(StateSet a, long n, int N,
                                                        const double* __restrict__ phi,
                                                        const double* __restrict__ param,
                                                        Params P) {
    const long s = blockIdx.x;
    if (s >= n) return;
    for (int i = threadIdx.x; i < N; i += 64) a.y[s * N + i] = phi[s * N + i];
    if (threadIdx.x == 0) {
        a.param[s] = param[s];
        a.t[s] = 0.0;
        a.h[s] = P.h0 > 0.0 ? fmin(P.h0, P.t_end) : 0.0;
        a.status[s] = P.t_end > 0.0 ? ACTIVE : DONE;
        a.attempts[s] = 0;
        a.accepted[s] = 0;
        a.rejected[s] = 0;
        a.rejlast[s] = 0;
        a.orig[s] = (int)s;
    }
}

// Finished states (every state with flush) go back to the caller's array and
// into the statistics [accepted, rejected, failed, most attempts]; active
// ones are packed into b at positions taken from *count.
__global__ void __launch_bounds__(64) ros2_compact_kernel(StateSet a, StateSet b, long m, int N,
                                                           double* __restrict__ phi, int* count,
                                                           unsigned long long* stats, int flush) {
    __shared__ long pos;
    const long s = blockIdx.x;
    if (s >= m) return;
    const int st = a.status[s];
    const bool keep = !flush && st == ACTIVE;
    if (threadIdx.x == 0) {
        if (keep) {
            pos = atomicAdd(count, 1);
        } else {
            pos = -1;
            atomicAdd(&stats[0], (unsigned long long)a.accepted[s]);
            atomicAdd(&stats[1], (unsigned long long)a.rejected[s]);
            if (st != DONE) atomicAdd(&stats[2], 1ull);
            atomicMax(&stats[3], (unsigned long long)a.attempts[s]);
        }
    }
    __syncthreads();
    const double* src = a.y + s * N;
    if (!keep) {
        double* dst = phi + (long)a.orig[s] * N;
        for (int i = threadIdx.x; i < N; i += 64) dst[i] = src[i];
        return;
    }
    const long d = pos;
    for (int i = threadIdx.x; i < N; i += 64) b.y[d * N + i] = src[i];
    if (threadIdx.x == 0) {
        b.param[d] = a.param[s];
        b.t[d] = a.t[s];
        b.h[d] = a.h[s];
        b.status[d] = st;
        b.attempts[d] = a.attempts[s];
        b.accepted[d] = a.accepted[s];
        b.rejected[d] = a.rejected[s];
        b.rejlast[d] = a.rejlast[s];
        b.orig[d] = a.orig[s];
    }
}

__global__ void ros2_zero_kernel(int* p) {
    if (threadIdx.x == 0) p[0] = 0;
}

__global__ void ros2_zero_stats_kernel(unsigned long long* p) {
    if (threadIdx.x < 4) p[threadIdx.x] = 0;
}

}  // namespace pjros

struct pyjac_integrator {
    pyjac_handle* h = nullptr;  // full C-order Jacobian and the species rates
    int NN = 0;
    char* buf = nullptr;
    size_t buf_bytes = 0;
    int* d_counts = nullptr;                // [0] active after an attempt, [1] packing
    unsigned long long* d_stats = nullptr;  // accepted, rejected, failed, most attempts
    int* h_active = nullptr;                // pinned
};

namespace {

// two copies of the moving per-state data and one of the per-attempt scratch
size_t per_state_bytes(int NN) {
    const size_t N = NN;
    const size_t moving = (N + 3) * sizeof(double) + 6 * sizeof(int);
    const size_t scratch = (N * N + 4 * N) * sizeof(double) + N * sizeof(int);
    return 2 * moving + scratch;
}

char* carve(char*& p, size_t bytes) {
    char* r = p;
    p += (bytes + 255) & ~(size_t)255;
    return r;
}

pjros::StateSet carve_set(char*& p, long cs, int N) {
    pjros::StateSet a;
    a.y = (double*)carve(p, (size_t)cs * N * sizeof(double));
    a.param = (double*)carve(p, (size_t)cs * sizeof(double));
    a.t = (double*)carve(p, (size_t)cs * sizeof(double));
    a.h = (double*)carve(p, (size_t)cs * sizeof(double));
    a.status = (int*)carve(p, (size_t)cs * sizeof(int));
    a.attempts = (int*)carve(p, (size_t)cs * sizeof(int));
    a.accepted = (int*)carve(p, (size_t)cs * sizeof(int));
    a.rejected = (int*)carve(p, (size_t)cs * sizeof(int));
    a.rejlast = (int*)carve(p, (size_t)cs * sizeof(int));
    a.orig = (int*)carve(p, (size_t)cs * sizeof(int));
    return a;
}

}  // namespace

extern "C" {

int pyjac_hip_integrator_create(const pyjac_mech_tables* tables, int conp, pyjac_integrator** out) {
    if (!tables || !out) return PYJAC_ERR_ARG;
    *out = nullptr;
    const int NN = tables->n_species + 1;
    if (NN > pjros::NGMAX) return PYJAC_ERR_UNSUPPORTED;
    pyjac_integrator* it = new pyjac_integrator();
    it->NN = NN;
    int rc = pyjac_hip_create(tables, conp, 'C', PYJAC_JAC_FULL, PYJAC_JAC_EXACT, &it->h);
    if (rc) {
        delete it;
        return rc;
    }
    if (hipMalloc(&it->d_counts, 2 * sizeof(int)) != hipSuccess ||
        hipMalloc(&it->d_stats, 4 * sizeof(unsigned long long)) != hipSuccess ||
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
    (void)hipFree(it->d_counts);
    (void)hipFree(it->d_stats);
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
    const size_t need = (size_t)cs * per + 32 * 256;  // + the alignment of the carved arrays
    if (!it->buf || it->buf_bytes < need) {
        (void)hipFree(it->buf);
        it->buf = nullptr;
        it->buf_bytes = 0;
        if (hipMalloc(&it->buf, need) != hipSuccess) return PYJAC_ERR_ALLOC;
        it->buf_bytes = need;
    }
    // the iteration matrix in LDS up to NMAX, beyond that in device memory
    // (in place of the Jacobian) with R = ceil(N / 64) rows per lane
    // (PYJAC_HIP_INTEGRATOR_DEVICE_LU: the device-memory path for any NN, for
    // tests that compare it with the LDS path)
    const bool in_lds = N <= pjros::NMAX && !getenv("PYJAC_HIP_INTEGRATOR_DEVICE_LU");
    const int R = (N + 63) / 64;
    const size_t lds = in_lds ? (size_t)N * N * sizeof(double) : 0;
    if (lds > 64 * 1024) {
        if (hipFuncSetAttribute((const void*)pjros::ros2_stage1_kernel,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess ||
            hipFuncSetAttribute((const void*)pjros::ros2_stage2_kernel,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
            return PYJAC_ERR_HIP;
    }
    char* p = it->buf;
    pjros::StateSet set[2] = {carve_set(p, cs, N), carve_set(p, cs, N)};
    double* f0 = (double*)carve(p, (size_t)cs * N * sizeof(double));
    double* J = (double*)carve(p, (size_t)cs * N * N * sizeof(double));
    double* k1 = (double*)carve(p, (size_t)cs * N * sizeof(double));
    double* y1 = (double*)carve(p, (size_t)cs * N * sizeof(double));
    double* f1 = (double*)carve(p, (size_t)cs * N * sizeof(double));
    int* piv = (int*)carve(p, (size_t)cs * N * sizeof(int));
    if ((size_t)(p - it->buf) > it->buf_bytes) return PYJAC_ERR_ALLOC;

    pyjac_integrate_stats st{0, 0, 0, 0, 0};
    pjros::Params P{t_end, rtol, atol, h0, fixed, max_steps};
    hipLaunchKernelGGL(pjros::ros2_zero_stats_kernel, dim3(1), dim3(64), 0, s, it->d_stats);
    for (size_t s0 = 0; s0 < n; s0 += cs) {
        long m = (long)std::min<size_t>(cs, n - s0);
        double* phi = d_phi + s0 * N;
        int cur = 0;
        hipLaunchKernelGGL(pjros::ros2_init_kernel, dim3((unsigned)m), dim3(64), 0, s, set[0], m,
                           N, phi, d_param + s0, P);
        for (long att = 0; att < max_steps && m > 0; ++att) {
            const pjros::StateSet& a = set[cur];
            pjros::Chunk c{};
            c.n = m;
            c.NN = N;
            c.param = a.param;
            c.y = a.y;
            c.f0 = f0;
            c.J = J;
            c.piv = piv;
            c.k1 = k1;
            c.y1 = y1;
            c.f1 = f1;
            c.t = a.t;
            c.h = a.h;
            c.status = a.status;
            c.attempts = a.attempts;
            c.accepted = a.accepted;
            c.rejected = a.rejected;
            c.rejlast = a.rejlast;
            c.active = it->d_counts;
            int rc = pyjac_hip_species_rates_dev(it->h, m, c.param, c.y, c.f0, s);
            if (!rc) rc = pyjac_hip_jacobian_dev(it->h, m, c.param, c.y, c.J, s);
            if (rc) return rc;
            if (in_lds) {
                hipLaunchKernelGGL(pjros::ros2_stage1_kernel, dim3((unsigned)m), dim3(pjros::NT),
                                   lds, s, c, P);
            } else {
                switch (R) {
#define PJROS_S1(RR)                                                                      \
    case RR:                                                                              \
        hipLaunchKernelGGL(pjros::ros2_stage1_global_kernel<RR>, dim3((unsigned)m),       \
                           dim3(pjros::NT), 0, s, c, P);                                  \
        break;
                    PJROS_S1(1) PJROS_S1(2) PJROS_S1(3) PJROS_S1(4) PJROS_S1(5) PJROS_S1(6) PJROS_S1(7)
                    PJROS_S1(8)
#undef PJROS_S1
                    default: return PYJAC_ERR_UNSUPPORTED;
                }
            }
            rc = pyjac_hip_species_rates_dev(it->h, m, c.param, c.y1, c.f1, s);
            if (rc) return rc;
            hipLaunchKernelGGL(pjros::ros2_zero_kernel, dim3(1), dim3(64), 0, s, c.active);
            if (in_lds) {
                hipLaunchKernelGGL(pjros::ros2_stage2_kernel, dim3((unsigned)m), dim3(pjros::NT),
                                   lds, s, c, P);
            } else {
                switch (R) {
#define PJROS_S2(RR)                                                                      \
    case RR:                                                                              \
        hipLaunchKernelGGL(pjros::ros2_stage2_global_kernel<RR>, dim3((unsigned)m),       \
                           dim3(64), 0, s, c, P);                                         \
        break;
                    PJROS_S2(1) PJROS_S2(2) PJROS_S2(3) PJROS_S2(4) PJROS_S2(5) PJROS_S2(6) PJROS_S2(7)
                    PJROS_S2(8)
#undef PJROS_S2
                    default: return PYJAC_ERR_UNSUPPORTED;
                }
            }
            if (hipGetLastError() != hipSuccess) return PYJAC_ERR_HIP;
            ++st.launches;
            if (hipMemcpyAsync(it->h_active, c.active, sizeof(int), hipMemcpyDeviceToHost, s) !=
                    hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess)
                return PYJAC_ERR_HIP;
            const long active = *it->h_active;
            if (active == 0) break;
            if (active <= m - std::max<long>(1, m / 16)) {
                hipLaunchKernelGGL(pjros::ros2_zero_kernel, dim3(1), dim3(64), 0, s,
                                   it->d_counts + 1);
                hipLaunchKernelGGL(pjros::ros2_compact_kernel, dim3((unsigned)m), dim3(64), 0, s,
                                   set[cur], set[1 - cur], m, N, phi, it->d_counts + 1,
                                   it->d_stats, 0);
                cur = 1 - cur;
                m = active;
            }
        }
        // every state left goes back, finished or not
        if (m > 0)
            hipLaunchKernelGGL(pjros::ros2_compact_kernel, dim3((unsigned)m), dim3(64), 0, s,
                               set[cur], set[1 - cur], m, N, phi, it->d_counts + 1, it->d_stats, 1);
        if (hipGetLastError() != hipSuccess) return PYJAC_ERR_HIP;
    }
    unsigned long long hs[4];
    if (hipMemcpyAsync(hs, it->d_stats, sizeof(hs), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return PYJAC_ERR_HIP;
    st.accepted = (long)hs[0];
    st.rejected = (long)hs[1];
    st.failed = (long)hs[2];
    st.max_steps = (long)hs[3];
    if (stats) *stats = st;
    return PYJAC_OK;
}

}  // extern "C"
