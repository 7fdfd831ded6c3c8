// Finite-difference Jacobian (jac_type finite_difference), after the
// reference's finite_difference_jacobian (core/create_jacobian.py:4523-4896):
// CVODE-style difference-quotient steps, then forward / backward / central
// differences of the species rates of the given order.
//
// Per state (all FP64):
//   f0     = species rates at phi
//   w_i    = 1 / (ATOL + RTOL |phi_i|)                            (:4710-4716)
//   sum    = sum_i (w_i |f0_i|)^2, in index order
//   fac    = sqrt(sum / NN), r0 = 1000 RTOL DBL_EPSILON NN fac    (:4723-4727)
//   r_i    = max(sqrt(DBL_EPSILON) |phi_i|, r0 / w_i)             (:4737-4739)
//   J[j][i] = (sum_k y_k f_j(phi + x_k r_i e_i)) / r_i, k in order (:4755-4800)
// with the reference's coefficient tables (x_k, y_k) (:4596-4617).  Two
// differences from the reference, whose own test marks its finite-difference
// Jacobian broken (tests/test_jacobian.py:2683): the error weights are the
// inverse ones of CVODE's difference quotient that the formula follows (the
// reference takes ATOL + RTOL |phi_i| itself, which makes the steps of small
// entries orders of magnitude too large), and forward orders 3 and 4 use the
// textbook coefficients (the reference's have +3/2 and +1/4, whose sums are
// not zero).
//
// The rates of the perturbed states come from the species-rate kernels
// themselves: a chunk of states is expanded into NN * K perturbed states in
// a C-order buffer, evaluated in one launch, and combined.

#include <hip/hip_runtime.h>

#include <cmath>

#include "kernels.h"

namespace pjfd {

// one state's error-weight sum and steps: r[s][i] for i < NN
__global__ void __launch_bounds__(256)
fd_steps_kernel(long cnt, int NN, const double* __restrict__ phi_c,
                const double* __restrict__ f0, double rtol, double atol,
                double* __restrict__ r) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= cnt) return;
    const double* p = phi_c + s * NN;
    const double* f = f0 + s * NN;
    const double eps = 2.220446049250313e-16;  // DBL_EPSILON
    double sum = 0.0;
    for (int i = 0; i < NN; ++i) {
        const double w = 1.0 / (atol + rtol * fabs(p[i]));
        const double t = w * fabs(f[i]);
        sum = sum + t * t;
    }
    const double fac = sqrt(sum / (double)NN);
    const double r0 = 1000.0 * rtol * eps * (double)NN * fac;
    const double srur = sqrt(eps);
    for (int i = 0; i < NN; ++i) {
        const double w = 1.0 / (atol + rtol * fabs(p[i]));
        r[s * NN + i] = fmax(srur * fabs(p[i]), r0 / w);
    }
}

// gather states s0 .. s0 + cnt - 1 of the caller's layout into C order
__global__ void __launch_bounds__(256)
fd_gather_kernel(long cnt, int NN, long n, long s0, int order_f, const double* __restrict__ param,
                 const double* __restrict__ phi, double* __restrict__ param_c,
                 double* __restrict__ phi_c) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt * NN) return;
    const long s = t / NN;
    const int j = (int)(t % NN);
    phi_c[t] = order_f ? phi[(size_t)j * n + s0 + s] : phi[(s0 + s) * NN + j];
    if (j == 0) param_c[s] = param[s0 + s];
}

// perturbed state (s, i, k) = phi_s with entry i moved by x_k r_i; C order,
// [cnt][NN][K][NN], and its parameter
__global__ void __launch_bounds__(256)
fd_perturb_kernel(long cnt, int NN, int K, const int* __restrict__ xk,
                  const double* __restrict__ param_c, const double* __restrict__ phi_c,
                  const double* __restrict__ r, double* __restrict__ pparam,
                  double* __restrict__ pphi) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long per = (long)NN * K * NN;
    if (t >= cnt * per) return;
    const long s = t / per;
    const long rem = t % per;
    const int i = (int)(rem / ((long)K * NN));
    const int k = (int)((rem / NN) % K);
    const int j = (int)(rem % NN);
    const double v = phi_c[s * NN + j];
    pphi[t] = j == i ? v + (double)xk[k] * r[s * NN + i] : v;
    if (j == 0) pparam[t / NN] = param_c[s];
}

// J[j][i] of states s0 .. into the caller's layout: entry e of dense (j, i)
// through jac_pos' inverse (acc_map), -1 when not stored
__global__ void __launch_bounds__(256)
fd_combine_kernel(long cnt, int NN, int K, const double* __restrict__ yk,
                  const double* __restrict__ pf, const double* __restrict__ r,
                  const int* __restrict__ acc_map, int nout, long n, long s0, int order_f,
                  double* __restrict__ jac) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt * NN * NN) return;
    const long s = t / ((long)NN * NN);
    const int i = (int)((t / NN) % NN);  // column: the perturbed entry
    const int j = (int)(t % NN);         // row
    const int e = acc_map[j * NN + i];
    if (e < 0) return;
    const double* f = pf + ((s * NN + i) * K) * NN + j;
    double v = 0.0;
    for (int k = 0; k < K; ++k) v = v + yk[k] * f[(long)k * NN];
    v = v / r[s * NN + i];
    if (order_f) jac[(size_t)e * n + s0 + s] = v;
    else jac[(size_t)(s0 + s) * nout + e] = v;
}

}  // namespace pjfd

namespace pjk_host {

namespace {
inline unsigned blocks_for(long count) { return (unsigned)((count + 255) / 256); }
}  // namespace

hipError_t launch_fd_gather(long cnt, int NN, long n, long s0, char order, const double* param,
                            const double* phi, double* param_c, double* phi_c,
                            hipStream_t stream) {
    hipLaunchKernelGGL(pjfd::fd_gather_kernel, dim3(blocks_for(cnt * NN)), dim3(256), 0, stream,
                       cnt, NN, n, s0, order == 'F', param, phi, param_c, phi_c);
    return hipGetLastError();
}

hipError_t launch_fd_steps(long cnt, int NN, const double* phi_c, const double* f0, double rtol,
                           double atol, double* r, hipStream_t stream) {
    hipLaunchKernelGGL(pjfd::fd_steps_kernel, dim3(blocks_for(cnt)), dim3(256), 0, stream, cnt, NN,
                       phi_c, f0, rtol, atol, r);
    return hipGetLastError();
}

hipError_t launch_fd_perturb(long cnt, int NN, int K, const int* xk, const double* param_c,
                             const double* phi_c, const double* r, double* pparam, double* pphi,
                             hipStream_t stream) {
    hipLaunchKernelGGL(pjfd::fd_perturb_kernel, dim3(blocks_for(cnt * NN * K * NN)), dim3(256), 0,
                       stream, cnt, NN, K, xk, param_c, phi_c, r, pparam, pphi);
    return hipGetLastError();
}

hipError_t launch_fd_combine(long cnt, int NN, int K, const double* yk, const double* pf,
                             const double* r, const int* acc_map, int nout, long n, long s0,
                             char order, double* jac, hipStream_t stream) {
    hipLaunchKernelGGL(pjfd::fd_combine_kernel, dim3(blocks_for(cnt * NN * NN)), dim3(256), 0,
                       stream, cnt, NN, K, yk, pf, r, acc_map, nout, n, s0, order == 'F', jac);
    return hipGetLastError();
}

}  // namespace pjk_host
