// surrogate_kernels.hip -- BF16 MFMA kernels of the species-rate surrogate
// (gfx950): an MLP  55 -> 512 -> 512 -> 512 -> 512 -> 53  with SiLU,
// trained and evaluated by these kernels.
//
// Layouts (row-major, all padded by the host to the tile sizes):
//   activations A_l  [M][K]   bf16   (and, for training, A_l^T [K][M])
//   weights     W_l  [N][K]   bf16   (out x in; plus W_l^T [K][N])
//   pre-activations Z_l [M][N] fp32 (training), output [M][N] fp32
// Every product is C = A * B^T with both operands contiguous along the
// reduction axis ("NT"), so each MFMA operand is one 16-byte read:
//   forward          Z_l    = A_{l-1} W_l^T          (A = A_{l-1}, B = W_l)
//   input gradient   dA_l-1 = dZ_l W_l               (A = dZ_l,    B = W_l^T)
//   weight gradient  dW_l   = dZ_l^T A_{l-1}         (A = dZ_l^T,  B = A_{l-1}^T)
// The epilogues fuse bias + SiLU (forward), the SiLU derivative (backward)
// and write transposed copies where the next product needs them.
//
// MFMA: v_mfma_f32_32x32x16_bf16, 64 lanes, A and B fragments of 8 bf16
// per lane (lane l: row/col l % 32, k = 8 (l / 32) .. + 7), 16 fp32
// accumulators per lane (lane l, slot i: col l % 32,
// row 8 (i / 4) + 4 (l / 32) + i % 4).
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgk {

typedef uint16_t bf16_t;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) uint32_t u32x4;

__device__ __forceinline__ bf16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ float sigmoidf(float z) { return 1.0f / (1.0f + __expf(-z)); }

enum Epi { EPI_HIDDEN = 0, EPI_LINEAR = 1, EPI_DSILU = 2, EPI_F32 = 3 };

struct EpiArgs {
    const float* bias;   // [N]            HIDDEN, LINEAR
    float* z;            // [M][ldz] fp32  HIDDEN (optional), LINEAR, F32
    int ldz;
    bf16_t* a;           // [M][lda] bf16  HIDDEN, DSILU
    int lda;
    bf16_t* at;          // [N][ldat] bf16 HIDDEN (optional), DSILU (optional)
    int ldat;
    const float* zin;    // [M][ldzin] fp32 pre-activation of the layer below, DSILU
    int ldzin;
};

constexpr int BM = 128, BK = 64, PAD = 8;

// LDS of one (BM x BN) block: two stages of the A and B tiles (BK wide, each
// row padded by PAD), reused for the bf16 output tile of the epilogue
template <int BN>
constexpr size_t gemm_lds_bytes() {
    constexpr size_t ops = 2 * (size_t)(BM + BN) * (BK + PAD), outs = (size_t)BM * (BN + 8);
    return (ops > outs ? ops : outs) * sizeof(bf16_t);
}

// C[M][N] = A[M][K] * B[N][K]^T; M % 128 == 0, N % BN == 0, K % 32 == 0.
// Two LDS stages: the next K tile is fetched into registers while the
// current stage feeds the MFMAs, then stored into the other stage, with one
// barrier per K tile.  Four waves, each a (BM / WM) x (BN / WN) block of
// 32 x 32 MFMA tiles.
template <int BN, int EPI>
__global__ void __launch_bounds__(256)
gemm_nt(const bf16_t* __restrict__ A, int lda, const bf16_t* __restrict__ B, int ldb, int M,
        int N, int K, EpiArgs e) {
    constexpr int WN = BN / 64;            // waves along N (2 for BN=128, 1 for BN=64)
    constexpr int WM = 4 / WN;             // waves along M
    constexpr int TM = BM / WM / 32;       // 32x32 tiles per wave along M
    constexpr int TN = BN / WN / 32;       // along N
    constexpr int CA = BM * (BK / 8) / 256;  // 16-byte chunks per thread per tile
    constexpr int CB = BN * (BK / 8) / 256;
    constexpr int LDT = BN + 8;
    constexpr int STAGE = (BM + BN) * (BK + PAD);
    extern __shared__ __attribute__((aligned(16))) bf16_t smem[];
    auto sA = [&](int st) { return (bf16_t (*)[BK + PAD])(smem + st * STAGE); };
    auto sB = [&](int st) { return (bf16_t (*)[BK + PAD])(smem + st * STAGE + BM * (BK + PAD)); };
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const long m0 = (long)blockIdx.y * BM, n0 = (long)blockIdx.x * BN;
    f32x16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[i][j][q] = 0.0f;
    u32x4 ra[CA], rb[CB];
    // tile k0 -> registers; K may be a multiple of 32 only: the second half
    // of the last tile is zero
    auto fetch = [&](int k0) {
#pragma unroll
        for (int c = 0; c < CA; ++c) {
            const int ch = tid + c * 256, r = ch / (BK / 8), q = (ch % (BK / 8)) * 8;
            ra[c] = (k0 + q < K) ? *(const u32x4*)&A[(m0 + r) * lda + k0 + q] : u32x4{0, 0, 0, 0};
        }
#pragma unroll
        for (int c = 0; c < CB; ++c) {
            const int ch = tid + c * 256, r = ch / (BK / 8), q = (ch % (BK / 8)) * 8;
            rb[c] = (k0 + q < K) ? *(const u32x4*)&B[(n0 + r) * ldb + k0 + q] : u32x4{0, 0, 0, 0};
        }
    };
    auto stash = [&](int st) {
        bf16_t (*a)[BK + PAD] = sA(st);
        bf16_t (*b)[BK + PAD] = sB(st);
#pragma unroll
        for (int c = 0; c < CA; ++c) {
            const int ch = tid + c * 256, r = ch / (BK / 8), q = (ch % (BK / 8)) * 8;
            *(u32x4*)&a[r][q] = ra[c];
        }
#pragma unroll
        for (int c = 0; c < CB; ++c) {
            const int ch = tid + c * 256, r = ch / (BK / 8), q = (ch % (BK / 8)) * 8;
            *(u32x4*)&b[r][q] = rb[c];
        }
    };
    fetch(0);
    stash(0);
    __syncthreads();
    int cur = 0;
    for (int k0 = 0; k0 < K; k0 += BK) {
        const bool more = k0 + BK < K;
        if (more) fetch(k0 + BK);
        const bf16_t (*a)[BK + PAD] = sA(cur);
        const bf16_t (*b)[BK + PAD] = sB(cur);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            bf16x8 fa[TM], fb[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) {
                const int r = wm * (TM * 32) + i * 32 + (lane & 31);
                fa[i] = *(const bf16x8*)&a[r][kk + (lane >> 5) * 8];
            }
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                const int r = wn * (TN * 32) + j * 32 + (lane & 31);
                fb[j] = *(const bf16x8*)&b[r][kk + (lane >> 5) * 8];
            }
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
        // the other stage was last read before the previous barrier
        if (more) stash(cur ^ 1);
        __syncthreads();
        cur ^= 1;
    }
    // epilogue.  fp32 results are stored directly (32 lanes write 128
    // contiguous bytes of a row); bf16 results go through LDS so that the
    // row-major tile and its transpose leave in 16-byte stores.
    bf16_t* tile = smem;  // the operand stages are dead after the last barrier
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int cl = wn * (TN * 32) + j * 32 + (lane & 31);
            const long col = n0 + cl;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int rl = wm * (TM * 32) + i * 32 + (q >> 2) * 8 + (lane >> 5) * 4 + (q & 3);
                const long row = m0 + rl;
                const float v = acc[i][j][q];
                if (EPI == EPI_HIDDEN) {
                    const float z = v + e.bias[col];
                    if (e.z) e.z[row * e.ldz + col] = z;
                    tile[rl * LDT + cl] = f2bf(z * sigmoidf(z));
                } else if (EPI == EPI_LINEAR) {
                    e.z[row * e.ldz + col] = v + e.bias[col];
                } else if (EPI == EPI_DSILU) {
                    const float z = e.zin[row * e.ldzin + col];
                    const float sg = sigmoidf(z);
                    tile[rl * LDT + cl] = f2bf(v * sg * (1.0f + z * (1.0f - sg)));
                } else {
                    e.z[row * e.ldz + col] = v;
                }
            }
        }
    if (EPI == EPI_HIDDEN || EPI == EPI_DSILU) {
        __syncthreads();
        for (int ch = tid; ch < BM * (BN / 8); ch += 256) {
            const int r = ch / (BN / 8), q = (ch % (BN / 8)) * 8;
            *(u32x4*)&e.a[(m0 + r) * e.lda + n0 + q] = *(const u32x4*)&tile[r * LDT + q];
        }
        if (e.at) {
            for (int ch = tid; ch < BN * (BM / 8); ch += 256) {
                const int c = ch / (BM / 8), r0 = (ch % (BM / 8)) * 8;
                uint16_t v[8];
#pragma unroll
                for (int jj = 0; jj < 8; ++jj) v[jj] = tile[(r0 + jj) * LDT + c];
                u32x4 w;
                w.x = v[0] | ((uint32_t)v[1] << 16); w.y = v[2] | ((uint32_t)v[3] << 16);
                w.z = v[4] | ((uint32_t)v[5] << 16); w.w = v[6] | ((uint32_t)v[7] << 16);
                *(u32x4*)&e.at[(n0 + c) * e.ldat + m0 + r0] = w;
            }
        }
    }
}

// Features of a state (one thread per state): log T, log P and the logs of
// the mole fractions of all Ns species (bath gas by difference), each
// normalised with the training set's mean and scale, zero padded to 64.
// phi / param in the rate kernels' layout (C: [n][NN], F: [NN][n]).  A
// block's 256 x 64 tile is staged in LDS so that x [mpad][64] (and, for
// training, x^T [64][mpad]) are written with coalesced 16-byte stores.
constexpr int FX = 64;
__global__ void __launch_bounds__(256)
featurize(const double* __restrict__ param, const double* __restrict__ phi, long n, int NN,
          int conp, int colmajor, const float* __restrict__ mu, const float* __restrict__ sd,
          long mpad, bf16_t* __restrict__ x, bf16_t* __restrict__ xt) {
    __shared__ __attribute__((aligned(16))) bf16_t tile[256][FX + 8];
    const long s0 = (long)blockIdx.x * 256;
    const long s = s0 + threadIdx.x;
    const int Ns = NN - 1;
    auto ph = [&](int i) { return colmajor ? phi[(long)i * n + s] : phi[s * NN + i]; };
    if (s < n) {
        const double T = ph(0);
        const double P = conp ? param[s] : ph(1);
        const double V = conp ? ph(1) : param[s];
        double known = 0.0;
        for (int k = 0; k < Ns - 1; ++k) known += fmax(1e-300, ph(2 + k));
        const double ntot = P * V / (8314.4621 * T);
        const double inv = 1.0 / ntot;
        tile[threadIdx.x][0] = f2bf(((float)log(fmax(T, 1.0)) - mu[0]) / sd[0]);
        tile[threadIdx.x][1] = f2bf(((float)log(fmax(P, 1e-300)) - mu[1]) / sd[1]);
        for (int k = 0; k < Ns; ++k) {
            const double nk = k < Ns - 1 ? fmax(1e-300, ph(2 + k)) : fmax(0.0, ntot - known);
            tile[threadIdx.x][2 + k] = f2bf((logf((float)(nk * inv) + 1e-12f) - mu[2 + k]) / sd[2 + k]);
        }
        for (int c = Ns + 2; c < FX; ++c) tile[threadIdx.x][c] = 0;
    } else {
        for (int c = 0; c < FX; ++c) tile[threadIdx.x][c] = 0;
    }
    __syncthreads();
    const long rows = min(256L, mpad - s0);
    // x: rows of 64 bf16 = 8 chunks of 16 bytes
    for (int ch = threadIdx.x; ch < rows * (FX / 8); ch += 256) {
        const int r = ch / (FX / 8), q = (ch % (FX / 8)) * 8;
        *(u32x4*)&x[(s0 + r) * FX + q] = *(const u32x4*)&tile[r][q];
    }
    if (xt) {
        // x^T: for each feature, 256 consecutive states
        for (int ch = threadIdx.x; ch < FX * 32; ch += 256) {
            const int c = ch / 32, r0 = (ch % 32) * 8;
            if (r0 >= rows) continue;
            uint16_t v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = tile[r0 + j][c];
            u32x4 w;
            w.x = v[0] | ((uint32_t)v[1] << 16); w.y = v[2] | ((uint32_t)v[3] << 16);
            w.z = v[4] | ((uint32_t)v[5] << 16); w.w = v[6] | ((uint32_t)v[7] << 16);
            *(u32x4*)&xt[(long)c * mpad + s0 + r0] = w;
        }
    }
}

// Training targets, normalised per state: y_k = wdot_k / S (k < Ns, bath
// gas by mass conservation) with S = max_k |wdot_k|, and
// y_{Ns+1} = (log10 S - stat[0]) / stat[1]; y_Ns (the error head) is set
// during training.  wdot from the FP64 kernel's dphi.
__global__ void targets(const double* __restrict__ param, const double* __restrict__ phi,
                        const double* __restrict__ dphi, long n, int NN, int conp, int colmajor,
                        const float* __restrict__ stat, const double* __restrict__ mw_ratio,
                        int ldy, long mpad, float* __restrict__ y) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= mpad) return;
    const int Ns = NN - 1;
    for (int c = 0; c < ldy; ++c) y[s * ldy + c] = 0.0f;
    if (s >= n) return;
    auto at = [&](const double* a, int i) { return colmajor ? a[(long)i * n + s] : a[s * NN + i]; };
    const double V = conp ? at(phi, 1) : param[s];
    double wns = 0.0, S = 1e-300;
    for (int k = 0; k < Ns - 1; ++k) {
        const double w = at(dphi, 2 + k) / V;
        wns -= mw_ratio[k] * w;
        S = fmax(S, fabs(w));
    }
    S = fmax(S, fabs(wns));
    for (int k = 0; k < Ns - 1; ++k) y[s * ldy + k] = (float)(at(dphi, 2 + k) / V / S);
    y[s * ldy + Ns - 1] = (float)(wns / S);
    y[s * ldy + Ns + 1] = (float)((log10(S) - stat[0]) / stat[1]);
}

// Mean-squared-error gradient of the first nout outputs (rows below n),
// written as dZ (bf16, [M][ld] and [ld][M]); accumulates the loss.
__global__ void mse_grad(const float* __restrict__ out, const float* __restrict__ y, long n,
                         long mpad, int ld, int nout, float inv_count, bf16_t* __restrict__ dz,
                         bf16_t* __restrict__ dzt, float* __restrict__ loss) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= mpad) return;
    float part = 0.0f;
    for (int c = 0; c < ld; ++c) {
        float g = 0.0f;
        if (s < n && c < nout) {
            const float d = out[s * ld + c] - y[s * ld + c];
            part += d * d;
            g = 2.0f * d * inv_count;
        }
        const bf16_t b = f2bf(g);
        dz[s * ld + c] = b;
        dzt[(long)c * mpad + s] = b;
    }
    // wave reduction, one atomic per wave
    for (int o = 32; o >= 1; o >>= 1) part += __shfl_xor(part, o);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, part * inv_count);
}

// db[n] = sum over the batch of dZ (rows of dZ^T), one workgroup per n
__global__ void __launch_bounds__(256)
colsum(const bf16_t* __restrict__ dzt, long mpad, float* __restrict__ db) {
    const long r = blockIdx.x;
    float s = 0.0f;
    for (long i = threadIdx.x; i < mpad; i += 256) s += bf2f(dzt[r * mpad + i]);
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o >= 1; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[r] = red[0];
}

// Adam on fp32 master weights; refreshes the bf16 copies W [N][K] and
// W^T [K][N] (both may be null for the bias vector)
__global__ void adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
                     float* __restrict__ m2, long count, int N, int K, float lr, float b1, float b2,
                     float eps, float c1, float c2, float wd, bf16_t* __restrict__ wb,
                     bf16_t* __restrict__ wbt) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    float wi = w[i];
    wi -= lr * (wd * wi + (a * c1) / (sqrtf(b * c2) + eps));
    w[i] = wi;
    if (wb) {
        const bf16_t q = f2bf(wi);
        wb[i] = q;
        const long r = i / K, c = i % K;
        wbt[c * N + r] = q;
    }
}

// float -> bf16 copy of a [R][C] matrix and its transpose [C][R]
__global__ void to_bf16(const float* __restrict__ w, long count, int R, int C,
                        bf16_t* __restrict__ wb, bf16_t* __restrict__ wbt) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const bf16_t q = f2bf(w[i]);
    wb[i] = q;
    if (wbt) wbt[(i % C) * R + i / C] = q;
}

// Target of the error head (output Ns): log10 of the current relative error
// of the state, max_k |wdot_k - wdot_ref_k| / max_k |wdot_ref_k|, from the
// network's outputs and the targets (both normalised per state)
__global__ void err_target(const float* __restrict__ out, float* __restrict__ y, long n, int ld,
                           int Ns, const float* __restrict__ stat) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    // ratio of the predicted to the reference magnitude
    const float r = exp10f((out[s * ld + Ns + 1] - y[s * ld + Ns + 1]) * stat[1]);
    float emax = 0.0f;
    for (int k = 0; k < Ns; ++k) emax = fmaxf(emax, fabsf(out[s * ld + k] * r - y[s * ld + k]));
    y[s * ld + Ns] = log10f(fmaxf(emax, 1e-7f));
}

// Decode the network's output for states [0, n): S = 10^(out_{Ns+1} stat[1]
// + stat[0]), wdot_k = out_k S for the Ns species ([n][Ns] or [Ns][n]);
// accept[s] = 1 when the error head out_Ns is below log10_tol.  64 states
// per block: their output rows are staged in LDS, then written species by
// species so that F-order rows are coalesced.
__global__ void __launch_bounds__(256)
decode(const float* __restrict__ out, int ld, long n, int NN, int conp, int colmajor,
       const double* __restrict__ param, const double* __restrict__ phi,
       const float* __restrict__ stat, float log10_tol, double* __restrict__ wdot,
       int* __restrict__ accept) {
    __shared__ float tile[64][129];
    __shared__ double S[64];
    const int Ns = NN - 1, cols = Ns + 2;
    const long s0 = (long)blockIdx.x * 64;
    for (int i = threadIdx.x; i < 64 * cols; i += 256) {
        const int r = i / cols, c = i % cols;
        tile[r][c] = s0 + r < n ? out[(s0 + r) * ld + c] : 0.0f;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const long s = s0 + threadIdx.x;
        S[threadIdx.x] = exp10((double)tile[threadIdx.x][Ns + 1] * stat[1] + stat[0]);
        if (s < n) accept[s] = tile[threadIdx.x][Ns] < log10_tol ? 1 : 0;
    }
    __syncthreads();
    const int j = threadIdx.x & 63;
    const long s = s0 + j;
    if (s >= n) return;
    for (int k = threadIdx.x >> 6; k < Ns; k += 4) {
        const double w = (double)tile[j][k] * S[j];
        if (colmajor) wdot[(long)k * n + s] = w;
        else wdot[s * Ns + k] = w;
    }
}

// compact copies of the states idx[0..m) (same layout as the input)
__global__ void gather(const double* __restrict__ param, const double* __restrict__ phi, long n,
                       int NN, int colmajor, const long* __restrict__ idx, long m,
                       double* __restrict__ param_c, double* __restrict__ phi_c) {
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const long s = idx[j];
    param_c[j] = param[s];
    for (int i = 0; i < NN; ++i) {
        if (colmajor) phi_c[(long)i * m + j] = phi[(long)i * n + s];
        else phi_c[j * NN + i] = phi[s * NN + i];
    }
}

// production rates of the compact states from the FP64 kernel's dphi,
// written to rows idx[j] of wdot ([n][Ns] or [Ns][n])
__global__ void scatter(const double* __restrict__ param_c, const double* __restrict__ phi_c,
                        const double* __restrict__ dphi_c, long m, int NN, int conp, int colmajor,
                        const double* __restrict__ mw_ratio, const long* __restrict__ idx,
                        double* __restrict__ wdot, long n) {
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const int Ns = NN - 1;
    const long s = idx[j];
    auto at = [&](const double* a, int i) { return colmajor ? a[(long)i * m + j] : a[j * NN + i]; };
    const double V = conp ? at(phi_c, 1) : param_c[j];
    double wns = 0.0;
    for (int k = 0; k < Ns - 1; ++k) {
        const double w = at(dphi_c, 2 + k) / V;
        wns -= mw_ratio[k] * w;
        if (colmajor) wdot[(long)k * n + s] = w;
        else wdot[s * Ns + k] = w;
    }
    if (colmajor) wdot[(long)(Ns - 1) * n + s] = wns;
    else wdot[s * Ns + Ns - 1] = wns;
}

}  // namespace sgk

// ------------------------------------------------------------------ C ABI
// (declared in include/pyjac_surrogate.h)
extern "C" {

using sgk::bf16_t;

static int launch_gemm(int bn, int epi, const bf16_t* A, int lda, const bf16_t* B, int ldb, int M,
                       int N, int K, const sgk::EpiArgs& e, hipStream_t st) {
    if (M % sgk::BM || K % 32 || (bn != 64 && bn != 128) || N % bn) return 1;
    dim3 grid(N / bn, M / sgk::BM), block(256);
#define SGK_G(BN_, E_)                                                                        \
    if (bn == BN_ && epi == E_) {                                                             \
        constexpr size_t lds = sgk::gemm_lds_bytes<BN_>();                                    \
        static bool attr_##BN_##_##E_ = false;                                                \
        if (!attr_##BN_##_##E_) {                                                             \
            if (hipFuncSetAttribute((const void*)sgk::gemm_nt<BN_, E_>,                       \
                                    hipFuncAttributeMaxDynamicSharedMemorySize,               \
                                    (int)lds) != hipSuccess)                                  \
                return 2;                                                                     \
            attr_##BN_##_##E_ = true;                                                         \
        }                                                                                     \
        hipLaunchKernelGGL((sgk::gemm_nt<BN_, E_>), grid, block, lds, st, A, lda, B, ldb, M,  \
                           N, K, e);                                                          \
        return hipGetLastError() == hipSuccess ? 0 : 2;                                       \
    }
    SGK_G(128, 0) SGK_G(128, 1) SGK_G(128, 2) SGK_G(128, 3)
    SGK_G(64, 0) SGK_G(64, 1) SGK_G(64, 2) SGK_G(64, 3)
#undef SGK_G
    return 1;
}

int pyjac_sur_gemm(int epi, const void* A, int lda, const void* B, int ldb, int M, int N, int K,
                   const float* bias, float* z, int ldz, void* a, int ldaa, void* at, int ldat,
                   const float* zin, int ldzin, void* stream) {
    sgk::EpiArgs e{bias, z, ldz, (bf16_t*)a, ldaa, (bf16_t*)at, ldat, zin, ldzin};
    const int bn = (N % 128 == 0) ? 128 : 64;
    return launch_gemm(bn, epi, (const bf16_t*)A, lda, (const bf16_t*)B, ldb, M, N, K, e,
                       (hipStream_t)stream);
}

int pyjac_sur_featurize(const double* param, const double* phi, long n, int NN, int conp,
                        int colmajor, const float* mu, const float* sd, int ldx, long mpad,
                        void* x, void* xt, void* stream) {
    if (ldx != sgk::FX || NN + 1 > sgk::FX || mpad % 128) return 1;
    hipLaunchKernelGGL(sgk::featurize, dim3((mpad + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, param, phi, n, NN, conp, colmajor, mu, sd, mpad,
                       (bf16_t*)x, (bf16_t*)xt);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int pyjac_sur_targets(const double* param, const double* phi, const double* dphi, long n, int NN,
                      int conp, int colmajor, const float* stat, const double* mw_ratio,
                      int ldy, long mpad, float* y, void* stream) {
    hipLaunchKernelGGL(sgk::targets, dim3((mpad + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       param, phi, dphi, n, NN, conp, colmajor, stat, mw_ratio, ldy, mpad, y);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int pyjac_sur_mse_grad(const float* out, const float* y, long n, long mpad, int ld, int nout,
                       float inv_count, void* dz, void* dzt, float* loss, void* stream) {
    hipLaunchKernelGGL(sgk::mse_grad, dim3((mpad + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, out, y, n, mpad, ld, nout, inv_count, (bf16_t*)dz,
                       (bf16_t*)dzt, loss);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int pyjac_sur_colsum(const void* dzt, long mpad, int rows, float* db, void* stream) {
    hipLaunchKernelGGL(sgk::colsum, dim3(rows), dim3(256), 0, (hipStream_t)stream,
                       (const bf16_t*)dzt, mpad, db);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int pyjac_sur_adam(float* w, const float* g, float* m1, float* m2, long count, int N, int K,
                   float lr, float b1, float b2, float eps, float c1, float c2, float wd, void* wb,
                   void* wbt, void* stream) {
    hipLaunchKernelGGL(sgk::adam, dim3((count + 255) / 256), dim3(256), 0, (hipStream_t)stream, w,
                       g, m1, m2, count, N, K, lr, b1, b2, eps, c1, c2, wd, (bf16_t*)wb,
                       (bf16_t*)wbt);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int pyjac_sur_to_bf16(const float* w, long count, int R, int C, void* wb, void* wbt,
                      void* stream) {
    hipLaunchKernelGGL(sgk::to_bf16, dim3((count + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       w, count, R, C, (bf16_t*)wb, (bf16_t*)wbt);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int pyjac_sur_err_target(const float* out, float* y, long n, int ld, int Ns,
                         const float* wscale, void* stream) {
    hipLaunchKernelGGL(sgk::err_target, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       out, y, n, ld, Ns, wscale);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int pyjac_sur_decode(const float* out, int ld, long n, int NN, int conp, int colmajor,
                     const double* param, const double* phi, const float* wscale, float log10_tol,
                     double* wdot, int* accept, void* stream) {
    if (NN + 1 > 128 || ld < NN + 1) return 1;
    hipLaunchKernelGGL(sgk::decode, dim3((n + 63) / 64), dim3(256), 0, (hipStream_t)stream, out,
                       ld, n, NN, conp, colmajor, param, phi, wscale, log10_tol, wdot, accept);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int pyjac_sur_gather(const double* param, const double* phi, long n, int NN, int colmajor,
                     const long* idx, long m, double* param_c, double* phi_c, void* stream) {
    if (m <= 0) return 0;
    hipLaunchKernelGGL(sgk::gather, dim3((m + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       param, phi, n, NN, colmajor, idx, m, param_c, phi_c);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int pyjac_sur_scatter(const double* param_c, const double* phi_c, const double* dphi_c, long m,
                      int NN, int conp, int colmajor, const double* mw_ratio, const long* idx,
                      double* wdot, long n, void* stream) {
    if (m <= 0) return 0;
    hipLaunchKernelGGL(sgk::scatter, dim3((m + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       param_c, phi_c, dphi_c, m, NN, conp, colmajor, mw_ratio, idx, wdot, n);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

}  // extern "C"
