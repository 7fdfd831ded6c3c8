/*
 * pyjac_surrogate.h -- C ABI of the species-rate surrogate kernels
 * (libpyjac_hip.so, gfx950).
 *
 * The surrogate has no counterpart in the reference.  It is an MLP
 * 55 -> 512 -> 512 -> 512 -> 512 -> 53 (SiLU) from log-scaled T, P and mole
 * fractions to asinh-scaled production rates, trained and evaluated by these
 * kernels with BF16 operands on the MFMA units and FP32 accumulation; an
 * error head (output Ns) gates each state, and rejected states fall back to
 * the FP64 species-rate kernel of pyjac_hip.h.
 *
 * Conventions: bf16 arrays are uint16 bit patterns (IEEE binary32 with the
 * low 16 bits rounded away), row-major; every product is C = A * B^T with
 * A [M][K], B [N][K]; M % 128 == 0, N % 64 == 0, K % 32 == 0 (the caller pads
 * with zeros).  All calls are asynchronous on `stream` (a hipStream_t) and
 * return 0, 1 (bad shape) or 2 (launch failure).
 */
#ifndef PYJAC_SURROGATE_H
#define PYJAC_SURROGATE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* epilogues of pyjac_sur_gemm */
#define PYJAC_SUR_HIDDEN 0 /* z = C + bias; a = bf16(silu(z)) -> a [M][lda], at [N][ldat] (opt.), z (opt.) */
#define PYJAC_SUR_LINEAR 1 /* z = C + bias -> z [M][ldz] (fp32) */
#define PYJAC_SUR_DSILU 2  /* d = bf16(C * silu'(zin)) -> a [M][lda], at [N][ldat] (opt.) */
#define PYJAC_SUR_F32 3    /* C -> z [M][ldz] (fp32) */

int pyjac_sur_gemm(int epi, const void *A, int lda, const void *B, int ldb, int M, int N, int K,
                   const float *bias, float *z, int ldz, void *a, int lda_out, void *at, int ldat,
                   const float *zin, int ldzin, void *stream);

/* features [mpad][ldx] (and [ldx][mpad] when xt != NULL) of n states given in
   the rate kernels' layout (colmajor = F order): (log T, log P, log(X_k +
   1e-12) for the Ns species) normalised as (f - mu) / sd, zero padded */
int pyjac_sur_featurize(const double *param, const double *phi, long n, int NN, int conp,
                        int colmajor, const float *mu, const float *sd, int ldx, long mpad,
                        void *x, void *xt, void *stream);

/* training targets y [mpad][ldy] from the FP64 kernel's dphi, normalised
   per state: y_k = wdot_k / S (k < Ns) with S = max_k |wdot_k|, and
   y_{Ns+1} = (log10 S - stat[0]) / stat[1]; mw_ratio [Ns-1] = mw_k / mw_bath
   (bath gas by mass balance) */
int pyjac_sur_targets(const double *param, const double *phi, const double *dphi, long n, int NN,
                      int conp, int colmajor, const float *stat, const double *mw_ratio,
                      int ldy, long mpad, float *y, void *stream);

/* error-head target y[s][Ns] = log10 of the current relative error */
int pyjac_sur_err_target(const float *out, float *y, long n, int ld, int Ns, const float *stat,
                         void *stream);

/* MSE gradient of the first nout outputs -> dz [mpad][ld], dzt [ld][mpad]
   (bf16), loss += sum of squares * inv_count */
int pyjac_sur_mse_grad(const float *out, const float *y, long n, long mpad, int ld, int nout,
                       float inv_count, void *dz, void *dzt, float *loss, void *stream);

/* db[r] = sum_i dzt[r][i], r < rows */
int pyjac_sur_colsum(const void *dzt, long mpad, int rows, float *db, void *stream);

/* Adam step on fp32 w [N][K] (count = N*K) with bias corrections c1, c2 and
   decoupled weight decay wd; refreshes bf16 copies wb [N][K], wbt [K][N] */
int pyjac_sur_adam(float *w, const float *g, float *m1, float *m2, long count, int N, int K,
                   float lr, float b1, float b2, float eps, float c1, float c2, float wd, void *wb,
                   void *wbt, void *stream);

/* bf16 copies of an fp32 [R][C] matrix: wb [R][C] and wbt [C][R] (opt.) */
int pyjac_sur_to_bf16(const float *w, long count, int R, int C, void *wb, void *wbt,
                      void *stream);

/* wdot [n][Ns] (or [Ns][n] when colmajor) = out[s][k] * S with
   S = 10^(out[s][Ns+1] stat[1] + stat[0]); accept[s] = 1 when the error
   head out[s][Ns] < log10_tol */
int pyjac_sur_decode(const float *out, int ld, long n, int NN, int conp, int colmajor,
                     const double *param, const double *phi, const float *stat, float log10_tol,
                     double *wdot, int *accept, void *stream);

/* fallback of rejected states: compact copies of the states idx[0..m)
   (same layout), and the FP64 kernel's production rates of those copies
   written back to rows idx of wdot ([n][Ns], or [Ns][n] when colmajor;
   bath gas by mass conservation, mw_ratio [Ns-1] = mw_k / mw_bath) */
int pyjac_sur_gather(const double *param, const double *phi, long n, int NN, int colmajor,
                     const long *idx, long m, double *param_c, double *phi_c, void *stream);
int pyjac_sur_scatter(const double *param_c, const double *phi_c, const double *dphi_c, long m,
                      int NN, int conp, int colmajor, const double *mw_ratio, const long *idx,
                      double *wdot, long n, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PYJAC_SURROGATE_H */
