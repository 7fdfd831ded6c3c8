// runtime.cpp -- C ABI of libpyjac_hip.so (declared in include/pyjac_hip.h).
//
// Replaces the reference's generated C++ Kernel class
// (kernel_utils/common/kernel.hpp.in:86-336, kernel.cpp.in:40-640): a handle
// owns the device copy of the mechanism tables, the per-wave workspace of
// the persistent kernels and the staging buffers of the host-pointer entry
// points.  Every entry point returns a status code; the message of the last
// failure of the calling thread is available from pyjac_hip_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>
#include <type_traits>
#include <cstddef>
#include <cmath>

#include "../../include/pyjac_hip.h"
#include "device_mech.h"
#include "kernels.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(PYJAC_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// Packs host arrays into one blob, remembering where each landed.
struct Packer {
    std::vector<char> bytes;
    template <class T>
    size_t add(const T* p, size_t n) {
        size_t off = (bytes.size() + 255) & ~size_t(255);
        bytes.resize(off + std::max<size_t>(n, 1) * sizeof(T), 0);
        if (n) std::memcpy(bytes.data() + off, p, n * sizeof(T));
        return off;
    }
};

}  // namespace

struct pyjac_handle {
    int conp = 1;
    char order = 'C';
    int jac_format = PYJAC_JAC_SPARSE;
    int jac_type = PYJAC_JAC_EXACT;
    int device = 0;
    int n_cu = 256;
    DevMech m{};
    pjk_host::WaveTables wt{};
    pjk_host::BlockTables bt{};
    bool wave_ok_rates = false, wave_ok_jac = false;
    bool block_ok_rates = false, block_ok_jac = false;
    // Jacobian accumulated in device memory (dense accumulator beyond LDS)
    bool glob_ok_jac = false;
    pjk_host::BlockArgs* d_args_g = nullptr;
    pjk_host::BlockLds lds_g{};
    int wps_g = 0;
    double* d_chunk = nullptr;  // F order: one chunk of states' entries
    long chunk_states = 0;
    // LDS-kernel Jacobian in F order: C-order chunks before the transpose,
    // two of them when the transposes are pipelined on a second stream
    double* d_fchunk = nullptr;
    long fchunk_states = 0;
    bool fpipe_ok = false;
    pjk_host::FPipeline fpipe{};
    // finite-difference Jacobian: options, coefficients (device), buffers
    int fd_mode = PYJAC_FD_FORWARD, fd_order = 1, fd_K = 0;
    double fd_rtol = 1e-8, fd_atol = 1e-15;
    int* d_fd_x = nullptr;
    double* d_fd_y = nullptr;
    double* d_fd = nullptr;
    long fd_states = 0;
    // block kernels: device copies of their arguments, LDS layouts and the
    // tuned number of waves per state (0: default)
    pjk_host::BlockArgs* d_args[2] = {nullptr, nullptr};
    pjk_host::BlockLds lds[2]{};
    int wps[2] = {0, 0};
    char* d_blob = nullptr;
    // workspace of the persistent kernels
    double* d_ws = nullptr;
    size_t ws_bytes = 0;
    // staging for host-pointer calls
    size_t problem_size = 0, chunk = 0;
    double *d_param = nullptr, *d_phi = nullptr, *d_out = nullptr, *d_out2 = nullptr,
           *d_out3 = nullptr;
    size_t staged_states = 0;
    int grid_rates = 0, grid_jac = 0, block = 256;
    int variant[4] = {0, 0, 0, 0};
    // tuned occupancy targets of the block kernels (0: the default build)
    // and the F-order chunk of the Jacobian (0: the default size)
    int occ[2] = {0, 0};
    long fchunk_pref = 0;
    hipStream_t stream = nullptr;
    // names of the tables (version 2), for the introspection entry points
    std::vector<std::string> species_names, reaction_strings;
    // device-side finiteness scans: the input check of the device-pointer
    // entry points and the per-state report; two counters on the device
    int input_check = 0;
    unsigned long long* d_counters = nullptr;
    // mechanism-specialised species-rate module (variant 30): the code
    // object generated for these tables, and their fingerprint
    unsigned long long tables_hash = 0;
    hipModule_t lane_mod = nullptr;
    hipFunction_t lane_fn = nullptr;
    // mechanism-specialised analytical Jacobian module (variant 31): kernel
    // A (rates and per-reaction values into scratch), kernel B per output
    // layout, the scratch values per state and the scratch stride (states
    // per chunk, a literal in the module)
    std::vector<hipModule_t> ljac_mods;
    hipFunction_t ljac_a = nullptr, ljac_b[4] = {nullptr, nullptr, nullptr, nullptr};
    long ljac_nscr = 0, ljac_stride = 0;
    double* d_ljac_scr = nullptr;
    size_t ljac_scr_doubles = 0;
    // C order through F-order lines (variant 32): one chunk's entries, which
    // a tiled transpose writes out as C-order rows, and the F-layout position
    // of each C-order entry (host copy from the tables, device copy made on
    // first use)
    double* d_ljac_lines = nullptr;
    size_t ljac_lines_doubles = 0;
    std::vector<int> ljac_fpos;
    int* d_ljac_fpos = nullptr;
    // the stored sparsity pattern, row * NN + col, whatever the format
    std::vector<int> pattern;
};

namespace {

int nvalues(const pyjac_handle* h) {
    int NN = h->m.NN;
    return h->jac_format == PYJAC_JAC_FULL ? NN * NN : h->m.jac_nnz;
}

// grid for a persistent kernel: enough waves to fill the device, capped
// so that the workspace stays within the budget
int pick_grid(const pyjac_handle* h, size_t per_state_doubles, int variant, long n) {
    int waves_per_cu = 8;
    if (variant > 0) waves_per_cu = variant;  // tuned: waves per CU
    const char* env = getenv("PYJAC_HIP_WAVES_PER_CU");
    if (env && atoi(env) > 0) waves_per_cu = atoi(env);
    long blocks = (long)h->n_cu * waves_per_cu / (h->block / 64);
    size_t per_block = per_state_doubles * (size_t)h->block * sizeof(double);
    size_t budget = (size_t)24 << 30;  // 24 GiB of workspace at most
    long cap = (long)std::max<size_t>(1, budget / std::max<size_t>(per_block, 1));
    blocks = std::min(blocks, cap);
    long need = (n + h->block - 1) / h->block;
    blocks = std::max(1L, std::min(blocks, need));
    return (int)blocks;
}

int ensure_ws(pyjac_handle* h, size_t bytes) {
    if (bytes <= h->ws_bytes) return PYJAC_OK;
    if (h->d_ws) (void)hipFree(h->d_ws);
    h->d_ws = nullptr;
    h->ws_bytes = 0;
    HIP_TRY(hipMalloc(&h->d_ws, bytes));
    h->ws_bytes = bytes;
    return PYJAC_OK;
}

// persistent grid of one-wave workgroups for the wave-per-state kernels
int wave_grid(const pyjac_handle* h, size_t lds_bytes, long n) {
    long per_cu = (long)((160 * 1024) / std::max<size_t>(lds_bytes, 1));
    per_cu = std::max(1L, std::min(per_cu, 32L));
    const char* env = getenv("PYJAC_HIP_WAVES_PER_CU");
    if (env && atoi(env) > 0) per_cu = atoi(env);
    long G = (long)h->n_cu * per_cu;
    G = std::min(G, std::max(8L, (n + 7) / 8 * 8));
    return (int)std::max(8L, G / 8 * 8);
}

// Launch shape of the block-per-state kernels.  Variants: 0 automatic,
// 1 lane-per-state, 2 wave-per-state, 10 + W block-per-state with W waves
// per state (W = 1, 2, 4, 8), 20 + W block-per-state Jacobian accumulated in
// device memory.
constexpr size_t LDS_PER_CU = 160 * 1024;

int block_wps(const pyjac_handle* h, int kernel) {
    int v = h->variant[kernel];
    if (v > 10) return v - 10;
    if (h->wps[kernel]) return h->wps[kernel];
    // measured on MI355X: rates fastest with one wave per state; the
    // Jacobian with two for GRI (325 reactions: 2.5e7/s against 1.9e7/s with
    // one) but one for H2/O2 (27 reactions: 1.4e8/s against 8.2e7/s), whose
    // reactions fill less than one wave's pass
    if (kernel == 1) return h->m.Nr >= 128 ? 2 : 1;
    return 1;
}

// the species rates by the loaded specialised module: asked for, or
// automatic once a module is loaded
bool use_module(const pyjac_handle* h) {
    return h->lane_fn && (h->variant[0] == 0 || h->variant[0] == 30);
}

// the analytical Jacobian by the loaded specialised module: asked for, or
// automatic once a module is loaded (exact Jacobians; the approximate pattern
// and the finite-difference Jacobian keep the table-driven kernels)
int ljac_layout(const pyjac_handle* h) {
    return (h->jac_format == PYJAC_JAC_FULL ? 2 : 0) + (h->order == 'F' ? 1 : 0);
}

// the module's kernels for this handle are loaded (kernel A and the store
// kernel of the handle's layout), for the exact Jacobian
bool jac_module_ready(const pyjac_handle* h) {
    return h->ljac_a && h->ljac_b[ljac_layout(h)] && h->jac_type == PYJAC_JAC_EXACT;
}

// C order through lines (variant 32): the store kernel of the handle's
// format in F layout writes a chunk's entries as lines (entry stride = the
// chunk's states, non-temporal stores), and a tiled transpose writes them out
// as C-order rows, each C-order entry read from the line of its F-layout
// position
bool ljac_lines_ready(const pyjac_handle* h) {
    return h->order == 'C' && h->ljac_a && h->ljac_b[ljac_layout(h) | 1] &&
           h->jac_type == PYJAC_JAC_EXACT && !h->ljac_fpos.empty();
}

// the module runs through lines and the transpose: asked for (variant 32),
// or the automatic choice in C order, where a lane storing its own state's
// entries leaves each store scattered over 64 rows (measured on MI355X,
// 262,144 states: GRI-Mech 3.0 sparse 4.7 ms through lines, 9.5 ms stored
// directly, 6.8 ms for the block kernel; H2/O2 full, 65,536 states, 0.083
// against 0.143 ms)
bool ljac_through_lines(const pyjac_handle* h) {
    if (!ljac_lines_ready(h)) return false;
    return h->variant[1] == 32 || h->variant[1] == 0;
}

// automatic choice once loaded: F order (each wave stores whole lines), C
// order through lines; variant 31 asks for the store kernel of the handle's
// layout writing the output directly, in either order
bool use_jac_module(const pyjac_handle* h) {
    if (ljac_through_lines(h)) return true;
    if (!jac_module_ready(h)) return false;
    return h->variant[1] == 31 || (h->variant[1] == 0 && h->order == 'F');
}

bool use_block(const pyjac_handle* h, int kernel) {
    if (kernel == 0 && use_module(h)) return false;
    if (kernel == 1 && use_jac_module(h)) return false;
    int v = h->variant[kernel];
    bool ok = kernel == 0 ? h->block_ok_rates : h->block_ok_jac;
    return ok && (v == 0 || (v > 10 && v < 20));
}

// the Jacobian accumulated in device memory: asked for, or automatic when
// the dense accumulator does not fit in LDS
bool use_glob(const pyjac_handle* h) {
    int v = h->variant[1];
    if (!h->glob_ok_jac) return false;
    return (v > 20 && v < 30) || (v == 0 && !h->block_ok_jac);
}

int glob_wps(const pyjac_handle* h) {
    int v = h->variant[1];
    if (v > 20) return v - 20;
    if (h->wps_g) return h->wps_g;
    // measured on MI355X: 8 waves for 3000 reactions, 4 for 900, 2 for 325
    const int Nr = h->m.Nr;
    return Nr >= 2000 ? 8 : (Nr >= 600 ? 4 : 2);
}

// F-order output of the LDS block kernel through a C-order chunk and a
// transpose (PYJAC_HIP_FCHUNK states per chunk, 0 for direct F writes)
long fchunk_want(const pyjac_handle* h, long n) {
    if (h->order != 'F') return 0;
    // up to 65536 states and 1 GiB (measured on MI355X, GRI: 47.2 ms per
    // 1M states against 50.5 ms with direct F-order writes)
    const size_t per = (size_t)h->m.jac_nnz * sizeof(double);
    long want = std::min(65536L, (long)(((size_t)1 << 30) / per) / 64 * 64);
    if (h->fchunk_pref > 0) want = h->fchunk_pref;
    const char* env = getenv("PYJAC_HIP_FCHUNK");
    if (env) want = atol(env) / 64 * 64;
    if (want <= 0) return 0;
    return std::max(64L, std::min(want, (n + 63) / 64 * 64));
}

// two chunks and the second stream (default priority, non-blocking) with
// its events, made once per handle
int ensure_fchunk(pyjac_handle* h, long n) {
    const long want = fchunk_want(h, n);
    if (!want || (h->d_fchunk && h->fchunk_states >= want)) return PYJAC_OK;
    (void)hipFree(h->d_fchunk);
    h->d_fchunk = nullptr;
    h->fchunk_states = 0;
    HIP_TRY(hipMalloc(&h->d_fchunk, 2 * (size_t)want * h->m.jac_nnz * sizeof(double)));
    h->fchunk_states = want;
    if (!h->fpipe_ok) {
        pjk_host::FPipeline& p = h->fpipe;
        HIP_TRY(hipStreamCreateWithFlags(&p.aux, hipStreamNonBlocking));
        for (int i = 0; i < 2; ++i) {
            HIP_TRY(hipEventCreateWithFlags(&p.computed[i], hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&p.transposed[i], hipEventDisableTiming));
        }
        h->fpipe_ok = true;
    }
    return PYJAC_OK;
}

// Coefficients of the finite-difference Jacobian, the reference's tables
// (core/create_jacobian.py:4596-4617) with the textbook signs of forward
// orders 3 and 4 (the reference's +3/2 and +1/4 make those sums non-zero)
bool fd_coefficients(int mode, int order, std::vector<int>& x, std::vector<double>& y) {
    x.clear();
    y.clear();
    if (mode == PYJAC_FD_CENTRAL) {
        switch (order) {
            case 2: y = {-0.5, 0.5}; break;
            case 4: y = {1.0 / 12, -2.0 / 3, 2.0 / 3, -1.0 / 12}; break;
            case 6: y = {-1.0 / 60, 3.0 / 20, -3.0 / 4, 3.0 / 4, -3.0 / 20, 1.0 / 60}; break;
            case 8:
                y = {1.0 / 280, -4.0 / 105, 1.0 / 5, -4.0 / 5, 4.0 / 5, -1.0 / 5, 4.0 / 105,
                     -1.0 / 280};
                break;
            default: return false;
        }
        for (int k = -order / 2; k <= order / 2; ++k)
            if (k) x.push_back(k);
        return true;
    }
    if (mode != PYJAC_FD_FORWARD && mode != PYJAC_FD_BACKWARD) return false;
    switch (order) {
        case 1: y = {-1.0, 1.0}; break;
        case 2: y = {-3.0 / 2, 2.0, -1.0 / 2}; break;
        case 3: y = {-11.0 / 6, 3.0, -3.0 / 2, 1.0 / 3}; break;
        case 4: y = {-25.0 / 12, 4.0, -3.0, 4.0 / 3, -1.0 / 4}; break;
        case 5: y = {-137.0 / 60, 5.0, -5.0, 10.0 / 3, -5.0 / 4, 1.0 / 5}; break;
        case 6: y = {-49.0 / 20, 6.0, -15.0 / 2, 20.0 / 3, -15.0 / 4, 6.0 / 5, -1.0 / 6}; break;
        default: return false;
    }
    for (int k = 0; k <= order; ++k) x.push_back(k);
    if (mode == PYJAC_FD_BACKWARD) {
        for (auto& v : x) v = -v;
        for (auto& v : y) v = -v;
    }
    return true;
}

int fd_configure(pyjac_handle* h, int mode, int order, double rtol, double atol) {
    std::vector<int> x;
    std::vector<double> y;
    if (!fd_coefficients(mode, order, x, y))
        return fail(PYJAC_ERR_ARG, "no finite-difference coefficients for mode %d, order %d", mode,
                    order);
    if (!(rtol >= 0.0) || !(atol >= 0.0)) return fail(PYJAC_ERR_ARG, "negative tolerance");
    if (!h->d_fd_x) {
        HIP_TRY(hipMalloc(&h->d_fd_x, 8 * sizeof(int)));
        HIP_TRY(hipMalloc(&h->d_fd_y, 8 * sizeof(double)));
    }
    HIP_TRY(hipMemcpy(h->d_fd_x, x.data(), x.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h->d_fd_y, y.data(), y.size() * sizeof(double), hipMemcpyHostToDevice));
    if ((int)x.size() != h->fd_K) {
        // the buffers depend on the number of coefficients
        (void)hipFree(h->d_fd);
        h->d_fd = nullptr;
        h->fd_states = 0;
    }
    h->fd_mode = mode;
    h->fd_order = order;
    h->fd_K = (int)x.size();
    h->fd_rtol = rtol;
    h->fd_atol = atol;
    return PYJAC_OK;
}

// doubles per state of the finite-difference buffers: C-order state, base
// rates, steps, and the NN * K perturbed states with their rates
size_t fd_doubles_per_state(const pyjac_handle* h) {
    const size_t NN = h->m.NN, K = h->fd_K;
    return 1 + 3 * NN + NN * K + 2 * NN * K * NN;
}

long fd_chunk(const pyjac_handle* h, long n) {
    long c = (long)(((size_t)768 << 20) / (fd_doubles_per_state(h) * sizeof(double))) / 64 * 64;
    return std::max(64L, std::min(c, (n + 63) / 64 * 64));
}

int ensure_fd(pyjac_handle* h, long n) {
    const long want = fd_chunk(h, n);
    if (h->d_fd && h->fd_states >= want) return PYJAC_OK;
    (void)hipFree(h->d_fd);
    h->d_fd = nullptr;
    h->fd_states = 0;
    HIP_TRY(hipMalloc(&h->d_fd, (size_t)want * fd_doubles_per_state(h) * sizeof(double)));
    h->fd_states = want;
    return PYJAC_OK;
}

// chunk buffer of the device-memory variant (about 2 GiB): per state its
// contributions and, in F order, its entries
int ensure_chunk(pyjac_handle* h, long n) {
    const size_t per = ((size_t)h->bt.n_contrib + (h->order == 'F' ? h->m.jac_nnz : 0)) *
                       sizeof(double);
    long want = (long)(((size_t)2 << 30) / std::max<size_t>(per, 1)) / 64 * 64;
    const char* env = getenv("PYJAC_HIP_CHUNK_STATES");  // override, for tests of the chunking
    if (env && atol(env) > 0) want = atol(env) / 64 * 64;
    want = std::max(64L, std::min(want, (n + 63) / 64 * 64));
    if (h->d_chunk && h->chunk_states >= want) return PYJAC_OK;
    (void)hipFree(h->d_chunk);
    h->d_chunk = nullptr;
    h->chunk_states = 0;
    HIP_TRY(hipMalloc(&h->d_chunk, (size_t)want * per));
    h->chunk_states = want;
    return PYJAC_OK;
}

// device buffers a call on n states needs (lane-per-state workspace, chunk
// buffers), allocated ahead of a graph capture (no allocation may happen
// while a stream is captured)
int prepare_lane(pyjac_handle* h, int kernel, long n) {
    if (kernel == 1 && h->jac_type == PYJAC_JAC_FINITE_DIFFERENCE) {
        int rc = ensure_fd(h, n);
        if (rc || use_block(h, 0) || use_module(h)) return rc;
        // the species-rate launches over the perturbed states
        const long m = fd_chunk(h, n) * h->m.NN * h->fd_K;
        size_t per = pjk_host::ws_doubles_per_state(h->m, false);
        int grid = pick_grid(h, per, h->variant[0], m);
        return ensure_ws(h, per * (size_t)grid * h->block * sizeof(double));
    }
    if (kernel == 1 && use_glob(h)) return ensure_chunk(h, n);
    if (kernel == 1 && use_block(h, 1)) return ensure_fchunk(h, n);
    if (use_block(h, kernel) || (kernel == 0 && use_module(h))) return PYJAC_OK;
    if (h->variant[kernel] == 2 && (kernel == 0 ? h->wave_ok_rates : h->wave_ok_jac))
        return PYJAC_OK;
    size_t per = pjk_host::ws_doubles_per_state(h->m, kernel == 1);
    int grid = pick_grid(h, per, h->variant[kernel], n);
    return ensure_ws(h, per * (size_t)grid * h->block * sizeof(double));
}

// the specialised module's kernel: one state per lane, 64 per workgroup;
// entry i of state s at [s * sN + i * sI]
constexpr int LANE_VERSION = 1;
constexpr int LJAC_VERSION = 2;
constexpr int LANE_BLOCK = 64;

int launch_module(pyjac_handle* h, long n, const double* param, const double* phi, double* dphi,
                  char order, hipStream_t s) {
    long sN = order == 'F' ? 1 : h->m.NN, sI = order == 'F' ? n : 1;
    int conp = h->conp;
    void* args[] = {&n, &param, &phi, &dphi, &sN, &sI, &conp};
    const long grid = (n + LANE_BLOCK - 1) / LANE_BLOCK;
    if (grid > 0x7fffffffL) return fail(PYJAC_ERR_ARG, "too many states for one launch");
    HIP_TRY(hipModuleLaunchKernel(h->lane_fn, (unsigned)grid, 1, 1, LANE_BLOCK, 1, 1, 0, s, args,
                                  nullptr));
    return PYJAC_OK;
}

int run_rates(pyjac_handle* h, long n, const double* param, const double* phi, double* dphi,
              hipStream_t s) {
    if (n <= 0) return PYJAC_OK;
    if (use_module(h)) return launch_module(h, n, param, phi, dphi, h->order, s);
    if (use_block(h, 0)) {
        HIP_TRY(pjk_host::launch_rates_block(h->m, h->d_args[0], h->lds[0], block_wps(h, 0),
                                             h->conp, h->order, n, param, phi, dphi, s,
                                             h->occ[0]));
        return PYJAC_OK;
    }
    if (h->wave_ok_rates && h->variant[0] == 2) {
        size_t lds = pjk_host::wave_lds_bytes(h->m, false, 0);
        HIP_TRY(pjk_host::launch_rates_wave(h->m, h->conp, h->order, n, param, phi, dphi, h->wt,
                                            wave_grid(h, lds, n), s));
        return PYJAC_OK;
    }
    size_t per = pjk_host::ws_doubles_per_state(h->m, false);
    int grid = pick_grid(h, per, h->variant[0], n);
    int rc = ensure_ws(h, per * (size_t)grid * h->block * sizeof(double));
    if (rc) return rc;
    HIP_TRY(pjk_host::launch_species_rates(h->m, h->conp, h->order, n, param, phi, dphi,
                                           h->d_ws, grid, h->block, s));
    return PYJAC_OK;
}

// species rates of n states in C order (the buffers of the finite-difference
// Jacobian), whatever the handle's data order
int run_rates_c(pyjac_handle* h, long n, const double* param, const double* phi, double* dphi,
                hipStream_t s) {
    if (use_module(h)) return launch_module(h, n, param, phi, dphi, 'C', s);
    if (use_block(h, 0)) {
        HIP_TRY(pjk_host::launch_rates_block(h->m, h->d_args[0], h->lds[0], block_wps(h, 0),
                                             h->conp, 'C', n, param, phi, dphi, s, h->occ[0]));
        return PYJAC_OK;
    }
    size_t per = pjk_host::ws_doubles_per_state(h->m, false);
    int grid = pick_grid(h, per, h->variant[0], n);
    int rc = ensure_ws(h, per * (size_t)grid * h->block * sizeof(double));
    if (rc) return rc;
    HIP_TRY(pjk_host::launch_species_rates(h->m, h->conp, 'C', n, param, phi, dphi, h->d_ws, grid,
                                           h->block, s));
    return PYJAC_OK;
}

// finite-difference Jacobian (create_jacobian.py:4523-4896), in chunks
int run_fd_jac(pyjac_handle* h, long n, const double* param, const double* phi, double* jac,
               hipStream_t s) {
    int rc = ensure_fd(h, n);
    if (rc) return rc;
    const long NN = h->m.NN, K = h->fd_K, cs = h->fd_states;
    double* param_c = h->d_fd;
    double* phi_c = param_c + cs;
    double* f0 = phi_c + cs * NN;
    double* r = f0 + cs * NN;
    double* pparam = r + cs * NN;
    double* pphi = pparam + cs * NN * K;
    double* pf = pphi + cs * NN * K * NN;
    for (long s0 = 0; s0 < n; s0 += cs) {
        const long cnt = std::min(cs, n - s0);
        HIP_TRY(pjk_host::launch_fd_gather(cnt, (int)NN, n, s0, h->order, param, phi, param_c,
                                           phi_c, s));
        if ((rc = run_rates_c(h, cnt, param_c, phi_c, f0, s))) return rc;
        HIP_TRY(pjk_host::launch_fd_steps(cnt, (int)NN, phi_c, f0, h->fd_rtol, h->fd_atol, r, s));
        HIP_TRY(pjk_host::launch_fd_perturb(cnt, (int)NN, (int)K, h->d_fd_x, param_c, phi_c, r,
                                            pparam, pphi, s));
        if ((rc = run_rates_c(h, cnt * NN * K, pparam, pphi, pf, s))) return rc;
        HIP_TRY(pjk_host::launch_fd_combine(cnt, (int)NN, (int)K, h->d_fd_y, pf, r, h->bt.acc_map,
                                            nvalues(h), n, s0, h->order, jac, s));
    }
    return PYJAC_OK;
}

// the specialised Jacobian module: per chunk of at most `ljac_stride`
// states, kernel A writes the rates' per-reaction and per-species values to
// the scratch array ([value][state in chunk]) and kernel B stores every entry
// at its position p of the handle's layout (jac + (s - jS) * jN + p * jI),
// or, through lines, at its F-layout position in lines of the chunk (jN = 1,
// jS = s0, jI = the chunk's states) that a tiled transpose writes out as
// C-order rows; one state per lane, 64 per workgroup
int grow_buffer(double** buf, size_t* have, size_t need) {
    if (need <= *have) return PYJAC_OK;
    if (*buf) (void)hipFree(*buf);
    *buf = nullptr;
    *have = 0;
    HIP_TRY(hipMalloc(buf, need * sizeof(double)));
    *have = need;
    return PYJAC_OK;
}

int launch_jac_module(pyjac_handle* h, long n, const double* param, const double* phi,
                      double* jac, hipStream_t s) {
    const long stride = h->ljac_stride;
    const long ne = nvalues(h);
    const bool lines = ljac_through_lines(h);
    int rc = grow_buffer(&h->d_ljac_scr, &h->ljac_scr_doubles,
                  (size_t)(h->ljac_nscr - 1) * stride + (size_t)std::min(n, stride));
    if (!rc && lines)
        rc = grow_buffer(&h->d_ljac_lines, &h->ljac_lines_doubles, (size_t)ne * std::min(n, stride));
    if (rc) return rc;
    if (lines && !h->d_ljac_fpos) {
        HIP_TRY(hipMalloc(&h->d_ljac_fpos, h->ljac_fpos.size() * sizeof(int)));
        HIP_TRY(hipMemcpy(h->d_ljac_fpos, h->ljac_fpos.data(), h->ljac_fpos.size() * sizeof(int),
                          hipMemcpyHostToDevice));
    }
    const bool F = h->order == 'F';
    const int layout = lines ? (ljac_layout(h) | 1) : ljac_layout(h);
    long sN = F ? 1 : h->m.NN, sI = F ? n : 1;
    int conp = h->conp;
    double* scr = h->d_ljac_scr;
    for (long s0 = 0; s0 < n; s0 += stride) {
        long cnt = std::min(stride, n - s0);
        const unsigned grid = (unsigned)((cnt + LANE_BLOCK - 1) / LANE_BLOCK);
        void* aa[] = {&cnt, &s0, (void*)&param, (void*)&phi, &scr, &sN, &sI, (void*)&stride, &conp};
        HIP_TRY(hipModuleLaunchKernel(h->ljac_a, grid, 1, 1, LANE_BLOCK, 1, 1, 0, s, aa, nullptr));
        double* out = lines ? h->d_ljac_lines : jac;
        long jN = lines || F ? 1 : ne;
        long jS = lines ? s0 : 0;
        long jI = lines ? cnt : sI;
        void* ab[] = {&cnt, &s0, (void*)&param, (void*)&phi, &scr, &out, &sN, &sI, (void*)&stride,
                      &jN, &jS, &jI, &conp};
        HIP_TRY(hipModuleLaunchKernel(h->ljac_b[layout], grid, 1, 1, LANE_BLOCK, 1, 1, 0, s, ab,
                                      nullptr));
        if (lines)
            HIP_TRY(pjk_host::launch_entries_to_c(out, cnt, (int)ne, jac, s0, h->d_ljac_fpos, s));
    }
    return PYJAC_OK;
}

int run_jac(pyjac_handle* h, long n, const double* param, const double* phi, double* jac,
            hipStream_t s) {
    if (n <= 0) return PYJAC_OK;
    if (h->jac_type == PYJAC_JAC_FINITE_DIFFERENCE) return run_fd_jac(h, n, param, phi, jac, s);
    if (use_jac_module(h)) return launch_jac_module(h, n, param, phi, jac, s);
    if (use_block(h, 1)) {
        int rc = ensure_fchunk(h, n);
        if (rc) return rc;
        const bool chunked = fchunk_want(h, n) > 0;
        HIP_TRY(pjk_host::launch_jacobian_block(h->m, h->d_args[1], h->lds[1], block_wps(h, 1),
                                                h->conp, h->order, n, param, phi, jac,
                                                chunked ? h->d_fchunk : nullptr,
                                                chunked ? h->fchunk_states : 0, s,
                                                chunked && h->fpipe_ok ? &h->fpipe : nullptr,
                                                h->occ[1]));
        return PYJAC_OK;
    }
    if (use_glob(h)) {
        int rc = ensure_chunk(h, n);
        if (rc) return rc;
        HIP_TRY(pjk_host::launch_jacobian_global(h->m, h->d_args_g, h->lds_g, glob_wps(h),
                                                 h->conp, h->order, n, param, phi, jac,
                                                 h->d_chunk, h->chunk_states,
                                                 (size_t)h->bt.n_contrib, s));
        return PYJAC_OK;
    }
    if (h->wave_ok_jac && h->variant[1] == 2) {
        size_t lds = pjk_host::wave_lds_bytes(h->m, true, h->m.NN * h->m.NN);
        HIP_TRY(pjk_host::launch_jacobian_wave(h->m, h->conp, h->order, n, param, phi, jac, h->wt,
                                               wave_grid(h, lds, n), s));
        return PYJAC_OK;
    }
    size_t per = pjk_host::ws_doubles_per_state(h->m, true);
    int grid = pick_grid(h, per, h->variant[1], n);
    int rc = ensure_ws(h, per * (size_t)grid * h->block * sizeof(double));
    if (rc) return rc;
    HIP_TRY(pjk_host::launch_jacobian(h->m, h->conp, h->order, n, param, phi, jac, h->d_ws, grid,
                                      h->block, s));
    return PYJAC_OK;
}

// host <-> device copy of `cnt` states starting at `off` of an array with
// `width` values per state, in the handle's data order
int copy_states(pyjac_handle* h, double* dst, const double* src, size_t total, size_t off,
                size_t cnt, size_t width, bool to_device, hipStream_t s) {
    hipMemcpyKind kind = to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    if (h->order == 'C' || width == 1) {
        if (to_device)
            HIP_TRY(hipMemcpyAsync(dst, src + off * width, cnt * width * sizeof(double), kind, s));
        else
            HIP_TRY(hipMemcpyAsync(dst + off * width, src, cnt * width * sizeof(double), kind, s));
    } else {
        // F order: `width` rows of `total` states on the host, rows of `cnt`
        // states on the device (kernel_utils/c/memcpy_2d.hpp)
        if (to_device)
            HIP_TRY(hipMemcpy2DAsync(dst, cnt * sizeof(double), src + off, total * sizeof(double),
                                     cnt * sizeof(double), width, kind, s));
        else
            HIP_TRY(hipMemcpy2DAsync(dst + off, total * sizeof(double), src, cnt * sizeof(double),
                                     cnt * sizeof(double), width, kind, s));
    }
    return PYJAC_OK;
}

int ensure_staging(pyjac_handle* h) {
    if (!h->problem_size) return fail(PYJAC_ERR_STATE, "handle not resized (problem_size == 0)");
    size_t cnt = h->chunk ? std::min(h->chunk, h->problem_size) : h->problem_size;
    if (cnt <= h->staged_states && h->d_param) return PYJAC_OK;
    (void)hipFree(h->d_param); (void)hipFree(h->d_phi); (void)hipFree(h->d_out); (void)hipFree(h->d_out2);
    (void)hipFree(h->d_out3);
    h->d_param = h->d_phi = h->d_out = h->d_out2 = h->d_out3 = nullptr;
    size_t NN = h->m.NN;
    size_t outw = std::max<size_t>(NN, (size_t)nvalues(h));
    HIP_TRY(hipMalloc(&h->d_param, cnt * sizeof(double)));
    HIP_TRY(hipMalloc(&h->d_phi, cnt * NN * sizeof(double)));
    HIP_TRY(hipMalloc(&h->d_out, cnt * outw * sizeof(double)));
    HIP_TRY(hipMalloc(&h->d_out2, cnt * NN * sizeof(double)));
    HIP_TRY(hipMalloc(&h->d_out3, cnt * NN * sizeof(double)));
    h->staged_states = cnt;
    return PYJAC_OK;
}

// host arrays param[N] and phi[N * NN]: the first state holding a NaN or an
// infinity is an argument error
int check_host_inputs(const pyjac_handle* h, size_t N, const double* param, const double* phi) {
    const size_t NN = h->m.NN;
    for (size_t s = 0; s < N; ++s)
        if (!std::isfinite(param[s]))
            return fail(PYJAC_ERR_ARG, "param of state %zu is not finite", s);
    for (size_t i = 0; i < N * NN; ++i)
        if (!std::isfinite(phi[i])) {
            const size_t s = h->order == 'C' ? i / NN : i % N;
            return fail(PYJAC_ERR_ARG, "phi of state %zu is not finite (NaN or infinity)", s);
        }
    return PYJAC_OK;
}

// scan of a device array (n states, width values each) on stream s: number
// of non-finite states and the first one (n when none); synchronises s
int scan_nonfinite(pyjac_handle* h, size_t n, const double* d, int width, int32_t* flags,
                   size_t* n_bad, size_t* first, hipStream_t s) {
    if (!h->d_counters) HIP_TRY(hipMalloc(&h->d_counters, 2 * sizeof(unsigned long long)));
    const unsigned long long init[2] = {0ull, ~0ull};
    HIP_TRY(hipMemcpyAsync(h->d_counters, init, sizeof(init), hipMemcpyHostToDevice, s));
    HIP_TRY(pjk_host::launch_nonfinite((long)n, width, d, width == 1 ? 'C' : h->order, flags,
                                       h->d_counters, s));
    unsigned long long got[2];
    HIP_TRY(hipMemcpyAsync(got, h->d_counters, sizeof(got), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (n_bad) *n_bad = (size_t)got[0];
    if (first) *first = got[0] ? (size_t)got[1] : n;
    return PYJAC_OK;
}

// the input check of the device-pointer entry points
int check_dev_inputs(pyjac_handle* h, size_t n, const double* param, const double* phi,
                     hipStream_t s) {
    if (!h->input_check || !n) return PYJAC_OK;
    size_t bad = 0, first = 0;
    int rc = scan_nonfinite(h, n, param, 1, nullptr, &bad, &first, s);
    if (rc) return rc;
    if (bad) return fail(PYJAC_ERR_ARG, "param of state %zu is not finite (%zu states)", first, bad);
    rc = scan_nonfinite(h, n, phi, h->m.NN, nullptr, &bad, &first, s);
    if (rc) return rc;
    if (bad)
        return fail(PYJAC_ERR_ARG, "phi of state %zu is not finite (NaN or infinity; %zu states)",
                    first, bad);
    return PYJAC_OK;
}

// 64-bit FNV-1a over the species and reaction counts and the tables a
// specialised species-rate module is generated from, in a fixed order
// (pyjac/libgen/lanegen.py computes the same number)
unsigned long long tables_fingerprint(const pyjac_mech_tables* mt) {
    unsigned long long h = 0xcbf29ce484222325ULL;
    auto add = [&](const void* p, size_t bytes) {
        const unsigned char* c = (const unsigned char*)p;
        for (size_t i = 0; i < bytes; ++i) {
            h ^= c[i];
            h *= 0x100000001b3ULL;
        }
    };
    const int32_t head[2] = {mt->n_species, mt->n_reactions};
    add(head, sizeof(head));
    const size_t Ns = mt->n_species, Nr = mt->n_reactions, nrs = mt->n_rxn_spec;
    const size_t nsi = mt->n_simple, nf = mt->n_fall, ntr = mt->n_troe, nsr = mt->n_sri;
    const size_t nt = mt->n_thd, nts = mt->n_thd_spec, npl = mt->n_plog, nc = mt->n_cheb;
#define PJ_H(field, count) \
    if ((count) > 0) add(mt->field, (size_t)(count) * sizeof(*mt->field))
    PJ_H(rxn_off, Nr + 1); PJ_H(rxn_spec, nrs); PJ_H(nu_prod, nrs); PJ_H(nu_reac, nrs);
    PJ_H(nu_sum, Nr); PJ_H(rev_mask, Nr); PJ_H(thd_mask, Nr); PJ_H(simple_mask, Nr);
    PJ_H(plog_mask, Nr); PJ_H(cheb_mask, Nr); PJ_H(fall_mask, Nr);
    PJ_H(simple_form, nsi); PJ_H(simple_A, nsi); PJ_H(simple_b, nsi); PJ_H(simple_Ta, nsi);
    PJ_H(fall_type, nf); PJ_H(fall_form, nf); PJ_H(fall_blend, nf); PJ_H(fall_param, nf);
    PJ_H(fall_A, nf); PJ_H(fall_b, nf); PJ_H(fall_Ta, nf);
    PJ_H(troe_a, ntr); PJ_H(troe_T1inv, ntr); PJ_H(troe_T2, ntr); PJ_H(troe_T3inv, ntr);
    PJ_H(sri_a, nsr); PJ_H(sri_b, nsr); PJ_H(sri_c, nsr); PJ_H(sri_d, nsr); PJ_H(sri_e, nsr);
    PJ_H(thd_map, nt); PJ_H(thd_type, nt); PJ_H(thd_off, nt + 1); PJ_H(thd_spec, nts);
    PJ_H(thd_eff, nts);
    PJ_H(plog_np, npl); PJ_H(plog_params, 4 * npl * (size_t)mt->plog_max_p);
    PJ_H(cheb_nt, nc); PJ_H(cheb_np, nc);
    PJ_H(cheb_params, nc * (size_t)mt->cheb_max_t * (size_t)mt->cheb_max_p);
    PJ_H(cheb_plim, 2 * nc); PJ_H(cheb_tlim, 2 * nc);
    PJ_H(a_lo, 7 * Ns); PJ_H(a_hi, 7 * Ns); PJ_H(T_mid, Ns); PJ_H(mw_post, Ns - 1);
#undef PJ_H
    return h;
}

}  // namespace

extern "C" {

const char* pyjac_hip_last_error(void) { return g_last_error.c_str(); }

// ---------------------------------------------------------------- tables
// the array fields of pyjac_mech_tables in struct order, with their types
#define PJ_TABLE_ARRAYS(X)                                                               \
    X(int32_t, rxn_off) X(int32_t, rxn_spec) X(int32_t, nu_prod) X(int32_t, nu_reac)    \
    X(int32_t, nu_sum) X(int32_t, rev_mask) X(int32_t, thd_mask) X(int32_t, simple_mask) \
    X(int32_t, plog_mask) X(int32_t, cheb_mask) X(int32_t, fall_mask)                    \
    X(int32_t, simple_form) X(double, simple_A) X(double, simple_b) X(double, simple_Ta)  \
    X(int32_t, fall_map) X(int32_t, fall_type) X(int32_t, fall_form)                     \
    X(int32_t, fall_blend) X(int32_t, fall_param) X(int32_t, fall_to_thd)                \
    X(double, fall_A) X(double, fall_b) X(double, fall_Ta) X(double, troe_a)             \
    X(double, troe_T1inv) X(double, troe_T2) X(double, troe_T3inv) X(double, sri_a)      \
    X(double, sri_b) X(double, sri_c) X(double, sri_d) X(double, sri_e)                  \
    X(int32_t, thd_map) X(int32_t, thd_type) X(int32_t, thd_off) X(int32_t, thd_spec)    \
    X(double, thd_eff) X(int32_t, plog_map) X(int32_t, plog_np) X(double, plog_params)   \
    X(int32_t, cheb_map) X(int32_t, cheb_nt) X(int32_t, cheb_np) X(double, cheb_params)  \
    X(double, cheb_plim) X(double, cheb_tlim) X(double, a_lo) X(double, a_hi)            \
    X(double, T_mid) X(double, mw_post) X(int32_t, rxn_has_ns) X(int32_t, ccs_col_ptr)   \
    X(int32_t, ccs_row_ind) X(int32_t, crs_row_ptr) X(int32_t, crs_col_ind)

// the leading int32 fields, abi_version .. jac_nnz
constexpr int PJ_TABLE_SCALARS = 18;

struct LoadedTables {
    pyjac_mech_tables t;
    std::vector<std::vector<char>> blocks;
    std::vector<std::string> names[2];      // species, reactions
    std::vector<const char*> name_ptrs[2];
};

int pyjac_hip_tables_load(const char* path, pyjac_mech_tables** out) {
    if (!path || !out) return fail(PYJAC_ERR_ARG, "null argument");
    *out = nullptr;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return fail(PYJAC_ERR_IO, "could not open tables file: %s", path);
    auto* L = new LoadedTables();
    std::memset(&L->t, 0, sizeof(L->t));
    int rc = PYJAC_OK;
    auto rd = [&](void* p, size_t n) { return std::fread(p, 1, n, fp) == n; };
    char magic[8];
    int32_t ns = 0, na = 0;
    int32_t sc[PJ_TABLE_SCALARS];
    if (!rd(magic, 8) || !rd(&ns, 4)) rc = fail(PYJAC_ERR_IO, "tables file too short: %s", path);
    else if (std::memcmp(magic, "PJHIPTB1", 8) != 0)
        rc = fail(PYJAC_ERR_ARG, "not a pyJac HIP tables file: %s", path);
    else if (ns != PJ_TABLE_SCALARS || !rd(sc, sizeof(sc)) || !rd(&na, 4))
        rc = fail(PYJAC_ERR_ARG, "tables file %s: %d scalar fields, this library has %d", path, ns,
                  PJ_TABLE_SCALARS);
    if (!rc) {
        int32_t* dst = &L->t.abi_version;  // the 18 leading int32 fields are contiguous
        for (int i = 0; i < PJ_TABLE_SCALARS; ++i) dst[i] = sc[i];
        if (L->t.abi_version != PYJAC_MECH_ABI_VERSION && L->t.abi_version != 1)
            rc = fail(PYJAC_ERR_ARG, "tables file %s: ABI %d, library reads 1 and %d", path,
                      L->t.abi_version, PYJAC_MECH_ABI_VERSION);
        L->t.abi_version = PYJAC_MECH_ABI_VERSION;  // the loaded struct has the name fields
    }
    int idx = 0;
#define PJ_LOAD(T, name)                                                                  \
    if (!rc) {                                                                            \
        uint8_t ty;                                                                       \
        int64_t cnt;                                                                      \
        if (idx >= na || !rd(&ty, 1) || !rd(&cnt, 8) || cnt < 0)                          \
            rc = fail(PYJAC_ERR_IO, "tables file %s: array %d (" #name ") missing", path, idx); \
        else if (ty != (std::is_same<T, double>::value ? 1 : 0))                          \
            rc = fail(PYJAC_ERR_ARG, "tables file %s: array " #name " has another type", path); \
        else {                                                                            \
            L->blocks.emplace_back((size_t)std::max<int64_t>(cnt, 1) * sizeof(T));        \
            if (!rd(L->blocks.back().data(), (size_t)cnt * sizeof(T)))                   \
                rc = fail(PYJAC_ERR_IO, "tables file %s: array " #name " truncated", path); \
            L->t.name = (const T*)L->blocks.back().data();                                \
        }                                                                                 \
        ++idx;                                                                            \
    }
    PJ_TABLE_ARRAYS(PJ_LOAD)
#undef PJ_LOAD
    if (!rc && idx != na)
        rc = fail(PYJAC_ERR_ARG, "tables file %s: %d arrays, this library reads %d", path, na, idx);
    // optional trailer: "PJNAMES1", then the species names and the reaction
    // strings, each list as an int32 count and (int32 length, bytes) per name
    char tmagic[8];
    if (!rc && rd(tmagic, 8)) {
        if (std::memcmp(tmagic, "PJNAMES1", 8) != 0)
            rc = fail(PYJAC_ERR_ARG, "tables file %s: unknown trailer", path);
        const int32_t want[2] = {L->t.n_species, L->t.n_reactions};
        for (int q = 0; q < 2 && !rc; ++q) {
            int32_t cnt = 0;
            if (!rd(&cnt, 4) || cnt != want[q]) {
                rc = fail(PYJAC_ERR_IO, "tables file %s: name list %d has %d entries, expected %d",
                          path, q, cnt, want[q]);
                break;
            }
            for (int32_t i = 0; i < cnt && !rc; ++i) {
                int32_t len = 0;
                if (!rd(&len, 4) || len < 0 || len > (1 << 20)) {
                    rc = fail(PYJAC_ERR_IO, "tables file %s: name %d truncated", path, i);
                    break;
                }
                std::string nm((size_t)len, '\0');
                if (len && !rd(&nm[0], (size_t)len))
                    rc = fail(PYJAC_ERR_IO, "tables file %s: name %d truncated", path, i);
                L->names[q].push_back(nm);
            }
        }
        if (!rc) {
            for (int q = 0; q < 2; ++q)
                for (const auto& nm : L->names[q]) L->name_ptrs[q].push_back(nm.c_str());
            L->t.species_names = L->name_ptrs[0].data();
            L->t.reaction_strings = L->name_ptrs[1].data();
        }
    }
    std::fclose(fp);
    if (rc) {
        delete L;
        return rc;
    }
    static_assert(offsetof(LoadedTables, t) == 0, "tables first");
    *out = &L->t;
    return PYJAC_OK;
}

int pyjac_hip_tables_free(pyjac_mech_tables* tables) {
    delete reinterpret_cast<LoadedTables*>(tables);
    return PYJAC_OK;
}

// read_initial_conditions.cpp.in:41-116: one record per state, C order on disk
int pyjac_hip_read_initial_conditions(const char* filename, size_t num_states, int nn,
                                      double* param, double* phi, char order) {
    if (!filename || nn < 2 || (num_states && (!param || !phi)) || (order != 'C' && order != 'F'))
        return fail(PYJAC_ERR_ARG, "bad argument");
    FILE* fp = std::fopen(filename, "rb");
    if (!fp) return fail(PYJAC_ERR_IO, "could not open file: %s", filename);
    std::vector<double> rec((size_t)nn + 1);
    for (size_t i = 0; i < num_states; ++i) {
        size_t got = std::fread(rec.data(), sizeof(double), rec.size(), fp);
        if (got != rec.size()) {
            std::fclose(fp);
            return fail(PYJAC_ERR_IO,
                        "file (%s) is incorrectly formatted: state %zu needs %d doubles, %zu read",
                        filename, i, nn + 1, got);
        }
        param[i] = rec[1];
        if (order == 'C') {
            phi[i * nn] = rec[0];
            for (int j = 1; j < nn; ++j) phi[i * nn + j] = rec[j + 1];
        } else {
            phi[i] = rec[0];
            for (int j = 1; j < nn; ++j) phi[(size_t)j * num_states + i] = rec[j + 1];
        }
    }
    std::fclose(fp);
    return PYJAC_OK;
}

// write_data.hpp:15-32
int pyjac_hip_write_data(const char* filename, const double* arr, size_t count) {
    if (!filename || (count && !arr)) return fail(PYJAC_ERR_ARG, "bad argument");
    FILE* fp = std::fopen(filename, "wb");
    if (!fp) return fail(PYJAC_ERR_IO, "error opening file for data output: %s", filename);
    size_t put = std::fwrite(arr, sizeof(double), count, fp);
    int closed = std::fclose(fp);
    if (put != count || closed != 0)
        return fail(PYJAC_ERR_IO, "error writing %zu doubles to %s", count, filename);
    return PYJAC_OK;
}

int pyjac_hip_tables_hash(const pyjac_mech_tables* tables, unsigned long long* out) {
    if (!tables || !out) return fail(PYJAC_ERR_ARG, "null argument");
    *out = tables_fingerprint(tables);
    return PYJAC_OK;
}

int pyjac_hip_load_module(pyjac_handle* h, const char* path) {
    if (!h || !path) return fail(PYJAC_ERR_ARG, "null argument");
    hipModule_t mod = nullptr;
    if (hipModuleLoad(&mod, path) != hipSuccess) {
        (void)hipGetLastError();
        return fail(PYJAC_ERR_IO, "could not load code object %s", path);
    }
    auto reject = [&](int code, const char* why) {
        (void)hipGetLastError();
        (void)hipModuleUnload(mod);
        return fail(code, "code object %s: %s", path, why);
    };
    hipDeviceptr_t sig_ptr = nullptr;
    size_t sig_bytes = 0;
    hipFunction_t fn = nullptr;
    unsigned long long sig[4] = {0, 0, 0, 0};
    const bool is_jac = hipModuleGetGlobal(&sig_ptr, &sig_bytes, mod, "pyjac_ljac_sig") == hipSuccess;
    // a failed lookup sets the thread's last HIP error, which the next
    // launch's error check would report: clear it after every probe
    (void)hipGetLastError();
    if (is_jac) {
        // a specialised analytical Jacobian module
        unsigned long long js[7] = {0, 0, 0, 0, 0, 0, 0};
        if (sig_bytes != sizeof(js) || hipMemcpyDtoH(js, sig_ptr, sizeof(js)) != hipSuccess)
            return reject(PYJAC_ERR_ARG, "unreadable Jacobian module signature");
        if (js[1] != (unsigned long long)LJAC_VERSION)
            return reject(PYJAC_ERR_ARG, "generated for another module version");
        if (js[0] != h->tables_hash || js[2] != (unsigned long long)h->m.Ns ||
            js[3] != (unsigned long long)h->m.Nr)
            return reject(PYJAC_ERR_ARG, "generated from other mechanism tables");
        if (h->jac_type == PYJAC_JAC_APPROXIMATE)
            return reject(PYJAC_ERR_UNSUPPORTED, "the module stores the exact Jacobian's pattern");
        // its entries must be exactly the handle's stored pattern
        const long ne = (long)js[5];
        hipDeviceptr_t ent_ptr = nullptr;
        size_t ent_bytes = 0;
        std::vector<int> ent(ne > 0 ? ne : 0);
        if (ne != (long)h->pattern.size() || js[4] < 1 || js[6] < 64 || js[6] % 64 ||
            hipModuleGetGlobal(&ent_ptr, &ent_bytes, mod, "pyjac_ljac_entries") != hipSuccess ||
            ent_bytes != ent.size() * sizeof(int) ||
            hipMemcpyDtoH(ent.data(), ent_ptr, ent_bytes) != hipSuccess)
            return reject(PYJAC_ERR_ARG, "entry list does not match the tables' sparsity pattern");
        std::vector<char> seen((size_t)h->m.NN * h->m.NN, 0);
        for (int e : h->pattern) seen[e] = 1;
        for (int e : ent) {
            const int row = e >> 16, col = e & 0xffff;
            if (row >= h->m.NN || col >= h->m.NN || seen[row * h->m.NN + col] != 1)
                return reject(PYJAC_ERR_ARG, "entry list does not match the tables' sparsity pattern");
            seen[row * h->m.NN + col] = 2;
        }
        // the module is split into parts (kernel A; the store kernel of each
        // layout): attach every kernel this code object holds
        if (h->ljac_nscr && (h->ljac_nscr != (long)js[4] || h->ljac_stride != (long)js[6]))
            return reject(PYJAC_ERR_ARG, "part of a different Jacobian module");
        hipFunction_t a = nullptr, b[4] = {nullptr, nullptr, nullptr, nullptr};
        bool any = hipModuleGetFunction(&a, mod, "pyjac_ljac_a") == hipSuccess;
        for (int L = 0; L < 4; ++L) {
            char name[32];
            std::snprintf(name, sizeof(name), "pyjac_ljac_b%d", L);
            if (hipModuleGetFunction(&b[L], mod, name) == hipSuccess) any = true;
            else b[L] = nullptr;
        }
        (void)hipGetLastError();
        if (!any) return reject(PYJAC_ERR_ARG, "no Jacobian kernel in the code object");
        h->ljac_mods.push_back(mod);
        if (a) h->ljac_a = a;
        for (int L = 0; L < 4; ++L)
            if (b[L]) h->ljac_b[L] = b[L];
        h->ljac_nscr = (long)js[4];
        h->ljac_stride = (long)js[6];
        return PYJAC_OK;
    }
    if (hipModuleGetGlobal(&sig_ptr, &sig_bytes, mod, "pyjac_lane_sig") != hipSuccess ||
        sig_bytes != sizeof(sig)) {
        (void)hipGetLastError();
        return reject(PYJAC_ERR_ARG, "not a specialised species-rate or Jacobian module");
    }
    if (hipMemcpyDtoH(sig, sig_ptr, sizeof(sig)) != hipSuccess)
        return reject(PYJAC_ERR_HIP, "signature unreadable");
    if (sig[1] != (unsigned long long)LANE_VERSION)
        return reject(PYJAC_ERR_ARG, "generated for another module version");
    if (sig[0] != h->tables_hash || sig[2] != (unsigned long long)h->m.Ns ||
        sig[3] != (unsigned long long)h->m.Nr)
        return reject(PYJAC_ERR_ARG, "generated from other mechanism tables");
    if (hipModuleGetFunction(&fn, mod, "pyjac_lane_rates") != hipSuccess)
        return reject(PYJAC_ERR_ARG, "no pyjac_lane_rates kernel");
    if (h->lane_mod) (void)hipModuleUnload(h->lane_mod);
    h->lane_mod = mod;
    h->lane_fn = fn;
    return PYJAC_OK;
}

int pyjac_hip_create(const pyjac_mech_tables* mt, int conp, char order, int jac_format,
                     int jac_type, pyjac_handle** out) {
    if (!mt || !out) return fail(PYJAC_ERR_ARG, "null argument");
    *out = nullptr;
    if (mt->abi_version != PYJAC_MECH_ABI_VERSION && mt->abi_version != 1)
        return fail(PYJAC_ERR_ARG, "mechanism table ABI %d, library reads 1 and %d",
                    mt->abi_version, PYJAC_MECH_ABI_VERSION);
    if (order != 'C' && order != 'F') return fail(PYJAC_ERR_ARG, "order must be 'C' or 'F'");
    if (jac_format != PYJAC_JAC_FULL && jac_format != PYJAC_JAC_SPARSE)
        return fail(PYJAC_ERR_ARG, "bad jac_format %d", jac_format);
    if (jac_type != PYJAC_JAC_EXACT && jac_type != PYJAC_JAC_APPROXIMATE &&
        jac_type != PYJAC_JAC_FINITE_DIFFERENCE)
        return fail(PYJAC_ERR_ARG, "bad jac_type %d", jac_type);
    const int Ns = mt->n_species, Nr = mt->n_reactions, NN = Ns + 1;
    if (Ns < 2 || Nr < 1) return fail(PYJAC_ERR_ARG, "empty mechanism");
    if (mt->n_cheb && (mt->cheb_max_p > 16 || mt->cheb_max_t > 16))
        return fail(PYJAC_ERR_UNSUPPORTED, "Chebyshev degree above 16");

    pyjac_handle* h = new pyjac_handle();
    h->tables_hash = tables_fingerprint(mt);
    if (mt->abi_version >= 2) {
        // names are optional; a table without them gives NULL names
        if (mt->species_names)
            for (int k = 0; k < Ns; ++k)
                h->species_names.emplace_back(mt->species_names[k] ? mt->species_names[k] : "");
        if (mt->reaction_strings)
            for (int r = 0; r < Nr; ++r)
                h->reaction_strings.emplace_back(mt->reaction_strings[r] ? mt->reaction_strings[r]
                                                                         : "");
    }
    h->conp = conp ? 1 : 0;
    h->order = order;
    h->jac_format = jac_format;
    h->jac_type = jac_type;
    (void)hipGetDevice(&h->device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) == hipSuccess && prop.multiProcessorCount > 0)
        h->n_cu = prop.multiProcessorCount;

    for (int row = 0; row < NN; ++row)
        for (int q = mt->crs_row_ptr[row]; q < mt->crs_row_ptr[row + 1]; ++q)
            h->pattern.push_back(row * NN + mt->crs_col_ind[q]);
    // F-layout position of each C-order entry: column-major within the state
    // (full), the CCS index of the CRS entry (sparse)
    if (order == 'C') {
        if (jac_format == PYJAC_JAC_FULL) {
            for (int row = 0; row < NN; ++row)
                for (int col = 0; col < NN; ++col) h->ljac_fpos.push_back(col * NN + row);
        } else {
            std::vector<int> ccs((size_t)NN * NN, -1);
            for (int col = 0; col < NN; ++col)
                for (int q = mt->ccs_col_ptr[col]; q < mt->ccs_col_ptr[col + 1]; ++q)
                    ccs[(size_t)mt->ccs_row_ind[q] * NN + col] = q;
            for (int e : h->pattern) h->ljac_fpos.push_back(ccs[e]);
            for (int q : h->ljac_fpos)
                if (q < 0 || q >= (int)h->pattern.size()) {
                    h->ljac_fpos.clear();  // inconsistent CRS / CCS tables: no lines path
                    break;
                }
        }
    }

    // derived tables
    std::vector<int> thd_only, fall_order, jac_pos;
    for (int t = 0; t < mt->n_thd; ++t)
        if (mt->fall_mask[mt->thd_map[t]] < 0) thd_only.push_back(t);
    for (int blend : {1, 3, 2})
        for (int f = 0; f < mt->n_fall; ++f)
            if (mt->fall_blend[f] == blend) fall_order.push_back(f);
    if (jac_format == PYJAC_JAC_FULL) {
        for (int a = 0; a < NN; ++a)
            for (int b = 0; b < NN; ++b)
                jac_pos.push_back(order == 'C' ? a * NN + b : b * NN + a);
    } else if (order == 'F') {
        for (int col = 0; col < NN; ++col)
            for (int q = mt->ccs_col_ptr[col]; q < mt->ccs_col_ptr[col + 1]; ++q)
                jac_pos.push_back(mt->ccs_row_ind[q] * NN + col);
    } else {
        for (int row = 0; row < NN; ++row)
            for (int q = mt->crs_row_ptr[row]; q < mt->crs_row_ptr[row + 1]; ++q)
                jac_pos.push_back(row * NN + mt->crs_col_ind[q]);
    }

    // packed reaction records of the wave-per-state kernels
    struct I4 { int x, y, z, w; };
    struct D4 { double x, y, z, w; };
    std::vector<I4> rx_i(Nr);
    std::vector<D4> rx_d(Nr);
    std::vector<int> rx_s((size_t)8 * Nr, 0);
    bool packable = Ns < 32767;
    for (int r = 0; r < Nr; ++r) {
        int kind, si = mt->simple_mask[r];
        D4 d{0, 0, 0, 0};
        if (si >= 0) {
            kind = mt->simple_form[si];
            d = D4{mt->simple_A[si], mt->simple_b[si], mt->simple_Ta[si], 0.0};
        } else if (mt->plog_mask[r] >= 0) {
            kind = 5;
            d.w = mt->plog_mask[r];
        } else {
            kind = 6;
            d.w = mt->n_plog + mt->cheb_mask[r];
        }
        int nsum = mt->nu_sum[r];
        int cnt = mt->rxn_off[r + 1] - mt->rxn_off[r];
        if (nsum < -100 || nsum > 100 || cnt > pjk_host::WAVE_MAXSP || cnt < 1) packable = false;
        rx_i[r] = I4{kind | ((nsum + 128) << 8), mt->rev_mask[r], mt->thd_mask[r], cnt};
        rx_d[r] = d;
        for (int q = 0; q < 8; ++q) {
            int idx = mt->rxn_off[r] + q;
            if (q < cnt && q < pjk_host::WAVE_MAXSP) {
                int nr = mt->nu_reac[idx], np = mt->nu_prod[idx];
                if (nr < 0 || nr > 255 || np < 0 || np > 127) packable = false;
                rx_s[(size_t)8 * r + q] = mt->rxn_spec[idx] | (nr << 16) | (np << 24);
            } else {
                rx_s[(size_t)8 * r + q] = Ns;  // padding slot, nu = 0
            }
        }
    }

    // block-kernel tables: reaction records sorted by structure (third body,
    // pressure-dependent kind, reversibility, bath gas, species count) so
    // the lanes of a wave follow the same branches; the simplest come last,
    // where the last group of 64 is usually partly empty but costs its wave
    // a whole pass of the instructions its flags select
    std::vector<int> perm(Nr), pos_of(Nr);
    for (int r = 0; r < Nr; ++r) perm[r] = r;
    auto rkey = [&](int r) {
        int cnt = mt->rxn_off[r + 1] - mt->rxn_off[r];
        int has_ns = cnt > 0 && mt->rxn_spec[mt->rxn_off[r + 1] - 1] == Ns - 1;
        int kind = rx_i[r].x & 0xff;
        return ((mt->thd_mask[r] >= 0) << 12) | ((kind >= 5) << 11) |
               ((mt->fall_mask[r] >= 0) << 10) | ((mt->rev_mask[r] >= 0) << 9) |
               (has_ns << 8) | (kind << 4) | cnt;
    };
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return rkey(a) > rkey(b); });
    std::vector<I4> rec_i(Nr);
    std::vector<D4> rec_d(Nr);
    std::vector<int> rec_s((size_t)8 * Nr);
    for (int q = 0; q < Nr; ++q) {
        int r = perm[q];
        pos_of[r] = q;
        rec_i[q] = rx_i[r];
        rec_i[q].y = mt->rev_mask[r] >= 0 ? 1 : 0;
        rec_d[q] = rx_d[r];
        for (int k = 0; k < 8; ++k) rec_s[(size_t)8 * q + k] = rx_s[(size_t)8 * r + k];
        // slot 6: a species with a Jacobian row (net coefficient, not the
        // bath gas); zero-valued contributions of the reaction go there
        int home = 0;
        for (int idx = mt->rxn_off[r + 1] - 1; idx >= mt->rxn_off[r]; --idx)
            if (mt->rxn_spec[idx] != Ns - 1 && mt->nu_prod[idx] != mt->nu_reac[idx])
                home = mt->rxn_spec[idx];
        rec_s[(size_t)8 * q + 6] = home;
        rec_s[(size_t)8 * q + 7] = 0;
    }
    // third-body tables: kind-ordered entry list (plain, Lindemann, SRI,
    // Troe), efficiency of the bath gas in d/dT and d/dE, d(ci)/dn_j through
    // the bath gas, and the third-body columns as coefficient pairs
    // (distinct efficiency set, row species) with their non-zero columns
    const int nt_ = mt->n_thd;
    std::vector<int> thd_order, thd_pairs((size_t)8 * std::max(nt_, 1), -1);
    std::vector<double> thd_mod(std::max(nt_, 1), 1.0), effns(std::max(nt_, 1), 0.0);
    for (int cls = 0; cls < 4; ++cls)
        for (int t = 0; t < nt_; ++t) {
            int f = mt->fall_mask[mt->thd_map[t]];
            int c = f < 0 ? 0 : (mt->fall_blend[f] == 1 ? 1 : (mt->fall_blend[f] == 3 ? 2 : 3));
            if (c == cls) thd_order.push_back(t);
        }
    std::vector<std::vector<double>> eff_sets;
    std::vector<std::pair<int, int>> pairs;  // (efficiency set, species)
    for (int t = 0; t < nt_; ++t) {
        const int tt = mt->thd_type[t], e0 = mt->thd_off[t], e1 = mt->thd_off[t + 1];
        const bool last_ns = e1 > e0 && mt->thd_spec[e1 - 1] == Ns - 1;
        if (tt == 2) thd_mod[t] = last_ns ? 1.0 : 0.0;
        if (tt == 1 && last_ns) thd_mod[t] = mt->thd_eff[e1 - 1];
        if (last_ns) effns[t] = tt == 2 ? -1.0 : (tt == 1 ? 1.0 + -1.0 * mt->thd_eff[e1 - 1] : 0.0);
        std::vector<double> eff(Ns - 1, 0.0);
        bool any = false;
        for (int e = e0; e < e1; ++e) {
            int j = mt->thd_spec[e];
            if (j == Ns - 1) continue;
            double d = tt == 2 ? 1.0 : (tt == 1 ? mt->thd_eff[e] + -1.0 : 0.0);
            eff[j] = d;
            any = any || d != 0.0;
        }
        if (!any) continue;
        int g = (int)(std::find(eff_sets.begin(), eff_sets.end(), eff) - eff_sets.begin());
        if (g == (int)eff_sets.size()) eff_sets.push_back(eff);
        const int r = mt->thd_map[t];
        for (int q = 0; q < pjk_host::WAVE_MAXSP; ++q) {
            int idx = mt->rxn_off[r] + q;
            if (idx >= mt->rxn_off[r + 1]) break;
            int spec = mt->rxn_spec[idx];
            if (spec == Ns - 1 || mt->nu_prod[idx] == mt->nu_reac[idx]) continue;
            std::pair<int, int> key(g, spec);
            int p = (int)(std::find(pairs.begin(), pairs.end(), key) - pairs.begin());
            if (p == (int)pairs.size()) pairs.push_back(key);
            thd_pairs[(size_t)8 * t + q] = p;
        }
    }
    // flags of each group of 64 consecutive records (see pyjac_block.hip);
    // seven zero entries past the last group cover the waves of a block
    // kernel's last pass that have no group (up to eight waves per state)
    std::vector<int> grp_flags((Nr + 63) / 64 + 7, 0);
    for (int q = 0; q < Nr; ++q) {
        const int r = perm[q];
        const int cnt = mt->rxn_off[r + 1] - mt->rxn_off[r];
        const int kind = rx_i[r].x & 0xff;
        int& f = grp_flags[q / 64];
        f = (f & ~0xf) | std::max(f & 0xf, std::min(cnt, pjk_host::WAVE_MAXSP));
        if (mt->rev_mask[r] >= 0) f |= 1 << 4;
        if (mt->thd_mask[r] >= 0) f |= 1 << 5;
        if (kind >= 5) f |= 1 << 6;
        if (cnt > 0 && mt->rxn_spec[mt->rxn_off[r + 1] - 1] == Ns - 1) f |= 1 << 7;
        if (kind == 1) f |= 1 << 9;
    }
    for (int t = 0; t < nt_; ++t)
        for (int q = 0; q < 8; ++q)
            if (thd_pairs[(size_t)8 * t + q] >= 0) grp_flags[pos_of[mt->thd_map[t]] / 64] |= 1 << 8;
    struct Combo { int p, dst; double e; };
    std::vector<Combo> combos;
    for (size_t p = 0; p < pairs.size(); ++p)
        for (int j = 0; j < Ns - 1; ++j) {
            double e = eff_sets[pairs[p].first][j];
            if (e != 0.0) combos.push_back(Combo{(int)p, (pairs[p].second + 2) * NN + j + 2, e});
        }
    std::vector<int> out_pos((size_t)2 * jac_pos.size());
    for (size_t e = 0; e < jac_pos.size(); ++e) {
        int p = jac_pos[e], row = p / NN, col = p % NN;
        out_pos[2 * e] = p;
        out_pos[2 * e + 1] = (row >= 2 && col >= 2) ? row - 2 : -1;
    }
    // tables of the variant gathering in device memory: dense index ->
    // output entry; each entry's contribution slots (ent_start, in record
    // order, then its third-body columns) or LDS slot (ent_long); each
    // record's and each third-body column's target; each entry's row and
    // column
    std::vector<int> acc_map((size_t)NN * NN, -1);
    for (size_t e = 0; e < jac_pos.size(); ++e) acc_map[jac_pos[e]] = (int)e;
    const size_t nout_ = jac_pos.size();
    bool glob_tables = NN <= 0x7fff;
    // rows 0 and 1 are stored in full (the kernel writes them by entry)
    for (int j = 0; j < NN; ++j) glob_tables = glob_tables && acc_map[j] >= 0 && acc_map[NN + j] >= 0;
    // contributions per entry: the reactions' species-column slots (columns
    // T and V|P stay in LDS) and the third-body columns; entries with more
    // than GATHER_MAX get an LDS slot instead of buffer slots
    const int MS = pjk_host::WAVE_MAXSP;
    std::vector<int> rec_pos((size_t)Nr * 8 * MS, -1);
    auto each_slot = [&](auto&& fn) {
        for (int q = 0; q < Nr; ++q) {
            const int r = perm[q], e0 = mt->rxn_off[r], cnt = mt->rxn_off[r + 1] - e0;
            for (int k = 0; k < cnt && k < MS; ++k) {
                const int sk = mt->rxn_spec[e0 + k];
                if (sk == Ns - 1 || mt->nu_prod[e0 + k] == mt->nu_reac[e0 + k]) continue;
                const size_t rowp = (size_t)(sk + 2) * NN;
                for (int j = 0; j < cnt && j < MS; ++j) {
                    if (mt->rxn_spec[e0 + j] == Ns - 1) continue;
                    // an unstored pair is a structural zero (e.g. the product
                    // columns of an irreversible reaction): nothing to store
                    const int e = acc_map[rowp + mt->rxn_spec[e0 + j] + 2];
                    if (e >= 0) fn(e, ((size_t)q * 8 + j) * MS + k);
                }
            }
        }
    };
    std::vector<Combo> gcombos(combos);
    for (auto& cb : gcombos) {
        cb.dst = acc_map[cb.dst];
        glob_tables = glob_tables && cb.dst >= 0;
    }
    std::vector<long> ent_n(nout_, 0);
    if (packable && glob_tables) {
        each_slot([&](int e, size_t) { ++ent_n[e]; });
        for (const auto& cb : gcombos) ++ent_n[cb.dst];
    }
    std::vector<int> ent_long(nout_, -1), ent_start(nout_ + 1, 0);
    int n_long = 0;
    for (size_t e = 0; e < nout_; ++e) {
        if (ent_n[e] > pjk_host::GATHER_MAX) ent_long[e] = n_long++;
        ent_start[e + 1] = ent_start[e] + (ent_long[e] >= 0 ? 0 : (int)ent_n[e]);
    }
    const long n_contrib = ent_start[nout_];
    glob_tables = glob_tables && n_contrib < (1L << 31);
    {
        std::vector<int> next(ent_start.begin(), ent_start.end() - 1);
        auto target = [&](int e) { return ent_long[e] >= 0 ? -2 - ent_long[e] : next[e]++; };
        if (packable && glob_tables) each_slot([&](int e, size_t at) { rec_pos[at] = target(e); });
        for (auto& cb : gcombos)
            if (cb.dst >= 0) cb.dst = target(cb.dst);
    }
    std::vector<int> ent_rc(nout_);
    for (size_t e = 0; e < nout_; ++e) ent_rc[e] = ((jac_pos[e] / NN) << 16) | (jac_pos[e] % NN);

    Packer P;
    DevMech& m = h->m;
    m.Ns = Ns; m.Nr = Nr; m.NN = NN; m.ns = Ns - 1;
    m.n_rev = mt->n_rev; m.n_thd = mt->n_thd; m.n_fall = mt->n_fall;
    m.n_plog = mt->n_plog; m.plog_max_p = mt->plog_max_p;
    m.n_cheb = mt->n_cheb; m.cheb_max_t = mt->cheb_max_t; m.cheb_max_p = mt->cheb_max_p;
    m.n_rxn_has_ns = mt->n_rxn_has_ns;
    m.jac_nnz = (int)jac_pos.size();
    m.n_thd_only = (int)thd_only.size();
    const int nrs = mt->n_rxn_spec, nts = mt->n_thd_spec;
    const int nf = mt->n_fall, nt = mt->n_thd;
    struct Slot { const void** dst; size_t off; };
    std::vector<Slot> slots;
#define PACK(field, src, count) \
    slots.push_back({(const void**)&m.field, P.add(src, (size_t)(count))})
    PACK(rxn_off, mt->rxn_off, Nr + 1);
    PACK(rxn_spec, mt->rxn_spec, nrs);
    PACK(nu_prod, mt->nu_prod, nrs);
    PACK(nu_reac, mt->nu_reac, nrs);
    PACK(nu_sum, mt->nu_sum, Nr);
    PACK(rev_mask, mt->rev_mask, Nr);
    PACK(thd_mask, mt->thd_mask, Nr);
    PACK(simple_mask, mt->simple_mask, Nr);
    PACK(plog_mask, mt->plog_mask, Nr);
    PACK(cheb_mask, mt->cheb_mask, Nr);
    PACK(fall_mask, mt->fall_mask, Nr);
    PACK(simple_form, mt->simple_form, mt->n_simple);
    PACK(simple_A, mt->simple_A, mt->n_simple);
    PACK(simple_b, mt->simple_b, mt->n_simple);
    PACK(simple_Ta, mt->simple_Ta, mt->n_simple);
    PACK(fall_map, mt->fall_map, nf);
    PACK(fall_type, mt->fall_type, nf);
    PACK(fall_form, mt->fall_form, nf);
    PACK(fall_blend, mt->fall_blend, nf);
    PACK(fall_param, mt->fall_param, nf);
    PACK(fall_to_thd, mt->fall_to_thd, nf);
    PACK(fall_A, mt->fall_A, nf);
    PACK(fall_b, mt->fall_b, nf);
    PACK(fall_Ta, mt->fall_Ta, nf);
    PACK(troe_a, mt->troe_a, mt->n_troe);
    PACK(troe_T1inv, mt->troe_T1inv, mt->n_troe);
    PACK(troe_T2, mt->troe_T2, mt->n_troe);
    PACK(troe_T3inv, mt->troe_T3inv, mt->n_troe);
    PACK(sri_a, mt->sri_a, mt->n_sri);
    PACK(sri_b, mt->sri_b, mt->n_sri);
    PACK(sri_c, mt->sri_c, mt->n_sri);
    PACK(sri_d, mt->sri_d, mt->n_sri);
    PACK(sri_e, mt->sri_e, mt->n_sri);
    PACK(thd_map, mt->thd_map, nt);
    PACK(thd_type, mt->thd_type, nt);
    PACK(thd_off, mt->thd_off, nt + 1);
    PACK(thd_spec, mt->thd_spec, nts);
    PACK(thd_eff, mt->thd_eff, nts);
    PACK(plog_map, mt->plog_map, mt->n_plog);
    PACK(plog_np, mt->plog_np, mt->n_plog);
    PACK(plog_params, mt->plog_params, (size_t)4 * mt->n_plog * mt->plog_max_p);
    PACK(cheb_map, mt->cheb_map, mt->n_cheb);
    PACK(cheb_nt, mt->cheb_nt, mt->n_cheb);
    PACK(cheb_np, mt->cheb_np, mt->n_cheb);
    PACK(cheb_params, mt->cheb_params, (size_t)mt->n_cheb * mt->cheb_max_t * mt->cheb_max_p);
    PACK(cheb_plim, mt->cheb_plim, 2 * mt->n_cheb);
    PACK(cheb_tlim, mt->cheb_tlim, 2 * mt->n_cheb);
    PACK(a_lo, mt->a_lo, 7 * Ns);
    PACK(a_hi, mt->a_hi, 7 * Ns);
    PACK(T_mid, mt->T_mid, Ns);
    PACK(mw_post, mt->mw_post, Ns - 1);
    PACK(rxn_has_ns, mt->rxn_has_ns, mt->n_rxn_has_ns);
    PACK(thd_only, thd_only.data(), thd_only.size());
    PACK(fall_order, fall_order.data(), fall_order.size());
    PACK(jac_pos, jac_pos.data(), jac_pos.size());
#undef PACK
    size_t off_rx_i = P.add(rx_i.data(), rx_i.size());
    size_t off_rx_d = P.add(rx_d.data(), rx_d.size());
    size_t off_rx_s = P.add(rx_s.data(), rx_s.size());
    size_t off_rec_i = P.add(rec_i.data(), rec_i.size());
    size_t off_rec_d = P.add(rec_d.data(), rec_d.size());
    size_t off_rec_s = P.add(rec_s.data(), rec_s.size());
    size_t off_grp_flags = P.add(grp_flags.data(), grp_flags.size());
    size_t off_thd_order = P.add(thd_order.data(), thd_order.size());
    size_t off_thd_mod = P.add(thd_mod.data(), thd_mod.size());
    size_t off_effns = P.add(effns.data(), effns.size());
    size_t off_thd_pairs = P.add(thd_pairs.data(), thd_pairs.size());
    size_t off_combo = P.add(combos.data(), combos.size());
    size_t off_out_pos = P.add(out_pos.data(), out_pos.size());
    size_t off_acc_map = P.add(acc_map.data(), acc_map.size());
    size_t off_rec_pos = P.add(rec_pos.data(), rec_pos.size());
    size_t off_ent_start = P.add(ent_start.data(), ent_start.size());
    size_t off_ent_long = P.add(ent_long.data(), ent_long.size());
    size_t off_gcombo = P.add(gcombos.data(), gcombos.size());
    size_t off_ent_rc = P.add(ent_rc.data(), ent_rc.size());
    hipError_t e = hipMalloc(&h->d_blob, P.bytes.size());
    if (e == hipSuccess)
        e = hipMemcpy(h->d_blob, P.bytes.data(), P.bytes.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(h->d_blob);
        delete h;
        return fail(PYJAC_ERR_HIP, "table upload failed: %s", hipGetErrorString(e));
    }
    for (auto& s : slots) *s.dst = h->d_blob + s.off;
    h->wt.rx_i = h->d_blob + off_rx_i;
    h->wt.rx_d = h->d_blob + off_rx_d;
    h->wt.rx_s = h->d_blob + off_rx_s;
    h->bt.rec_i = h->d_blob + off_rec_i;
    h->bt.rec_d = h->d_blob + off_rec_d;
    h->bt.rec_s = h->d_blob + off_rec_s;
    h->bt.grp_flags = (const int*)(h->d_blob + off_grp_flags);
    h->bt.thd_order = (const int*)(h->d_blob + off_thd_order);
    h->bt.thd_mod = (const double*)(h->d_blob + off_thd_mod);
    h->bt.effns = (const double*)(h->d_blob + off_effns);
    h->bt.thd_pairs = (const int*)(h->d_blob + off_thd_pairs);
    h->bt.combo = h->d_blob + off_combo;
    h->bt.n_pairs = (int)pairs.size();
    h->bt.n_combo = (int)combos.size();
    h->bt.out_pos = h->d_blob + off_out_pos;
    h->bt.acc_map = (const int*)(h->d_blob + off_acc_map);
    h->bt.rec_pos = (const int*)(h->d_blob + off_rec_pos);
    h->bt.ent_start = (const int*)(h->d_blob + off_ent_start);
    h->bt.ent_long = (const int*)(h->d_blob + off_ent_long);
    h->bt.gcombo = h->d_blob + off_gcombo;
    h->bt.ent_rc = (const int*)(h->d_blob + off_ent_rc);
    h->bt.n_contrib = (int)n_contrib;
    h->bt.n_long = n_long;
    // the wave-per-state kernels keep a state's arrays in 64 KiB of LDS at most
    const size_t lds_cap = 64 * 1024;
    h->wave_ok_rates = packable && pjk_host::wave_lds_bytes(m, false, 0) <= lds_cap;
    h->wave_ok_jac = packable && pjk_host::wave_lds_bytes(m, true, NN * NN) <= lds_cap;
    // the block kernels fit one workgroup's arrays into the CU's LDS
    h->block_ok_rates = packable && pjk_host::block_lds_bytes(m, false, h->bt.n_pairs) <= LDS_PER_CU;
    h->block_ok_jac = packable && pjk_host::block_lds_bytes(m, true, h->bt.n_pairs) <= LDS_PER_CU;
    h->glob_ok_jac = packable && glob_tables &&
                     pjk_host::block_lds_bytes(m, true, h->bt.n_pairs, false, n_long) <= LDS_PER_CU;
    // an approximate Jacobian on a sparse pattern: only the kernels that sum
    // rows T and V|P over the stored entries compute it
    h->bt.approx_sums = jac_type == PYJAC_JAC_APPROXIMATE && jac_format == PYJAC_JAC_SPARSE;
    if (h->bt.approx_sums) {
        h->block_ok_jac = false;
        h->wave_ok_jac = false;
        if (!h->glob_ok_jac) {
            pyjac_hip_destroy(h);
            return fail(PYJAC_ERR_UNSUPPORTED,
                        "approximate sparse Jacobian: the mechanism does not fit the kernels "
                        "that compute it");
        }
    }
    if (getenv("PYJAC_HIP_VERBOSE"))
        fprintf(stderr,
                "pyjac_hip: Ns %d Nr %d nnz %d | LDS bytes: rates %zu, Jacobian %zu (dense), "
                "%zu (device-memory variant, %ld contributions) | block %d/%d, device-memory %d\n",
                Ns, Nr, (int)jac_pos.size(), pjk_host::block_lds_bytes(m, false, h->bt.n_pairs),
                pjk_host::block_lds_bytes(m, true, h->bt.n_pairs),
                pjk_host::block_lds_bytes(m, true, h->bt.n_pairs, false, n_long), n_contrib,
                (int)h->block_ok_rates, (int)h->block_ok_jac, (int)h->glob_ok_jac);
    for (int k = 0; k < 3; ++k) {
        pjk_host::BlockLds lay = pjk_host::block_lds(m, k != 0, h->bt.n_pairs, k != 2, n_long);
        pjk_host::BlockArgs a{m, h->bt, lay};
        pjk_host::BlockArgs** dst = k < 2 ? &h->d_args[k] : &h->d_args_g;
        if (k < 2) h->lds[k] = lay;
        else h->lds_g = lay;
        hipError_t e2 = hipMalloc(dst, sizeof(a));
        if (e2 == hipSuccess) e2 = hipMemcpy(*dst, &a, sizeof(a), hipMemcpyHostToDevice);
        if (e2 != hipSuccess) {
            pyjac_hip_destroy(h);
            return fail(PYJAC_ERR_HIP, "argument upload failed: %s", hipGetErrorString(e2));
        }
    }
    if (jac_type == PYJAC_JAC_FINITE_DIFFERENCE) {
        // the reference's defaults: forward, first order, rtol 1e-8, atol 1e-15
        int rc = fd_configure(h, PYJAC_FD_FORWARD, 1, 1e-8, 1e-15);
        if (rc) {
            pyjac_hip_destroy(h);
            return rc;
        }
    }
    *out = h;
    return PYJAC_OK;
}

int pyjac_hip_set_fd_options(pyjac_handle* h, int mode, int order, double rtol, double atol) {
    if (!h) return fail(PYJAC_ERR_ARG, "null handle");
    if (h->jac_type != PYJAC_JAC_FINITE_DIFFERENCE)
        return fail(PYJAC_ERR_STATE, "handle was not created with PYJAC_JAC_FINITE_DIFFERENCE");
    return fd_configure(h, mode, order, rtol, atol);
}

int pyjac_hip_destroy(pyjac_handle* h) {
    if (!h) return PYJAC_OK;
    (void)hipFree(h->d_args[0]);
    (void)hipFree(h->d_args[1]);
    (void)hipFree(h->d_args_g);
    (void)hipFree(h->d_chunk);
    (void)hipFree(h->d_fchunk);
    if (h->fpipe_ok) {
        (void)hipStreamDestroy(h->fpipe.aux);
        for (int i = 0; i < 2; ++i) {
            (void)hipEventDestroy(h->fpipe.computed[i]);
            (void)hipEventDestroy(h->fpipe.transposed[i]);
        }
    }
    (void)hipFree(h->d_fd);
    (void)hipFree(h->d_fd_x);
    (void)hipFree(h->d_fd_y);
    (void)hipFree(h->d_blob);
    (void)hipFree(h->d_ws);
    (void)hipFree(h->d_param);
    (void)hipFree(h->d_phi);
    (void)hipFree(h->d_out);
    (void)hipFree(h->d_out2);
    (void)hipFree(h->d_out3);
    (void)hipFree(h->d_counters);
    if (h->lane_mod) (void)hipModuleUnload(h->lane_mod);
    for (hipModule_t m : h->ljac_mods) (void)hipModuleUnload(m);
    (void)hipFree(h->d_ljac_scr);
    (void)hipFree(h->d_ljac_lines);
    (void)hipFree(h->d_ljac_fpos);
    delete h;
    return PYJAC_OK;
}

int pyjac_hip_resize(pyjac_handle* h, size_t problem_size, size_t work_size) {
    if (!h) return fail(PYJAC_ERR_ARG, "null handle");
    h->problem_size = problem_size;
    h->chunk = work_size;
    return PYJAC_OK;
}

int pyjac_hip_species_rates_dev(pyjac_handle* h, size_t n, const double* param, const double* phi,
                                double* dphi, void* stream) {
    if (!h || (n && (!param || !phi || !dphi))) return fail(PYJAC_ERR_ARG, "null argument");
    int rc = check_dev_inputs(h, n, param, phi, (hipStream_t)stream);
    if (rc) return rc;
    return run_rates(h, (long)n, param, phi, dphi, (hipStream_t)stream);
}

int pyjac_hip_jacobian_dev(pyjac_handle* h, size_t n, const double* param, const double* phi,
                           double* jac, void* stream) {
    if (!h || (n && (!param || !phi || !jac))) return fail(PYJAC_ERR_ARG, "null argument");
    int rc = check_dev_inputs(h, n, param, phi, (hipStream_t)stream);
    if (rc) return rc;
    return run_jac(h, (long)n, param, phi, jac, (hipStream_t)stream);
}

// The Jacobian without any allocation: the chunk buffers come from the
// library's fixed scratch area (pjk_host::noalloc_scratch); the F-order
// transposes run on the caller's stream after each chunk, one chunk buffer
int pyjac_hip_jacobian_dev_noalloc(pyjac_handle* h, size_t n, const double* param,
                                   const double* phi, double* jac, void* stream) {
    if (!h || (n && (!param || !phi || !jac))) return fail(PYJAC_ERR_ARG, "null argument");
    if (!n) return PYJAC_OK;
    if (h->input_check)
        return fail(PYJAC_ERR_UNSUPPORTED, "the input check synchronises and is not available "
                                           "without allocation; turn it off");
    if (h->jac_type == PYJAC_JAC_FINITE_DIFFERENCE)
        return fail(PYJAC_ERR_UNSUPPORTED, "finite-difference Jacobians need work buffers");
    hipStream_t s = (hipStream_t)stream;
    double* buf = nullptr;
    const size_t cap = pjk_host::NOALLOC_SCRATCH_BYTES / sizeof(double);
    // the specialised module's scratch array (a fixed stride of states per
    // chunk) does not fit the fixed area: with a module loaded this entry
    // point runs the block kernels the handle would run without it
    const int v1 = h->variant[1];
    const bool block = use_block(h, 1) ||
                       (use_jac_module(h) && h->block_ok_jac && (v1 == 0 || v1 == 31 || v1 == 32));
    if (block) {
        if (h->order == 'C') {
            HIP_TRY(pjk_host::launch_jacobian_block(h->m, h->d_args[1], h->lds[1],
                                                    block_wps(h, 1), h->conp, 'C', (long)n,
                                                    param, phi, jac, nullptr, 0, s, nullptr,
                                                    h->occ[1]));
            return PYJAC_OK;
        }
        // F order: C-order chunks of whole multiples of 64 states
        const size_t per = (size_t)h->m.jac_nnz;
        const long cs = (long)std::min<size_t>(cap / per, ((size_t)n + 63) / 64 * 64) / 64 * 64;
        if (cs < 64)
            return fail(PYJAC_ERR_UNSUPPORTED,
                        "64 states need %zu bytes of scratch, the fixed area holds %zu",
                        64 * per * sizeof(double), pjk_host::NOALLOC_SCRATCH_BYTES);
        HIP_TRY(pjk_host::noalloc_scratch(&buf));
        HIP_TRY(pjk_host::launch_jacobian_block(h->m, h->d_args[1], h->lds[1], block_wps(h, 1),
                                                h->conp, 'F', (long)n, param, phi, jac, buf, cs,
                                                s, nullptr, h->occ[1]));
        return PYJAC_OK;
    }
    if (use_glob(h)) {
        const size_t per = (size_t)h->bt.n_contrib + (h->order == 'F' ? h->m.jac_nnz : 0);
        const long cs = (long)std::min<size_t>(cap / per, ((size_t)n + 63) / 64 * 64) / 64 * 64;
        if (cs < 64)
            return fail(PYJAC_ERR_UNSUPPORTED,
                        "64 states need %zu bytes of scratch, the fixed area holds %zu",
                        64 * per * sizeof(double), pjk_host::NOALLOC_SCRATCH_BYTES);
        HIP_TRY(pjk_host::noalloc_scratch(&buf));
        HIP_TRY(pjk_host::launch_jacobian_global(h->m, h->d_args_g, h->lds_g, glob_wps(h),
                                                 h->conp, h->order, (long)n, param, phi, jac,
                                                 buf, cs, (size_t)h->bt.n_contrib, s));
        return PYJAC_OK;
    }
    return fail(PYJAC_ERR_UNSUPPORTED,
                "the handle's Jacobian variant keeps a work buffer per wave (lane- and "
                "wave-per-state kernels); the block-per-state variants run without allocation");
}

int pyjac_hip_set_input_check(pyjac_handle* h, int on) {
    if (!h) return fail(PYJAC_ERR_ARG, "null handle");
    h->input_check = on ? 1 : 0;
    return PYJAC_OK;
}

int pyjac_hip_nonfinite_dev(pyjac_handle* h, size_t n, const double* d_arr, int width,
                            int32_t* d_flags, size_t* n_bad, size_t* first_bad, void* stream) {
    if (!h || (n && !d_arr) || width < 1) return fail(PYJAC_ERR_ARG, "bad argument");
    if (n_bad) *n_bad = 0;
    if (first_bad) *first_bad = n;
    if (!n) return PYJAC_OK;
    return scan_nonfinite(h, n, d_arr, width, d_flags, n_bad, first_bad, (hipStream_t)stream);
}

int pyjac_hip_species_rates(pyjac_handle* h, const double* param, const double* phi,
                            double* dphi) {
    if (!h || !param || !phi || !dphi) return fail(PYJAC_ERR_ARG, "null argument");
    int rc = ensure_staging(h);
    if (rc) return rc;
    const size_t N = h->problem_size, NN = h->m.NN, per = h->staged_states;
    if ((rc = check_host_inputs(h, N, param, phi))) return rc;
    for (size_t off = 0; off < N; off += per) {
        size_t cnt = std::min(per, N - off);
        if ((rc = copy_states(h, h->d_param, param, N, off, cnt, 1, true, 0))) return rc;
        if ((rc = copy_states(h, h->d_phi, phi, N, off, cnt, NN, true, 0))) return rc;
        if ((rc = run_rates(h, (long)cnt, h->d_param, h->d_phi, h->d_out, 0))) return rc;
        if ((rc = copy_states(h, dphi, h->d_out, N, off, cnt, NN, false, 0))) return rc;
    }
    HIP_TRY(hipStreamSynchronize(0));
    return PYJAC_OK;
}

int pyjac_hip_jacobian(pyjac_handle* h, const double* param, const double* phi, double* jac) {
    if (!h || !param || !phi || !jac) return fail(PYJAC_ERR_ARG, "null argument");
    int rc = ensure_staging(h);
    if (rc) return rc;
    const size_t N = h->problem_size, NN = h->m.NN, per = h->staged_states;
    const size_t nv = nvalues(h);
    if ((rc = check_host_inputs(h, N, param, phi))) return rc;
    for (size_t off = 0; off < N; off += per) {
        size_t cnt = std::min(per, N - off);
        if ((rc = copy_states(h, h->d_param, param, N, off, cnt, 1, true, 0))) return rc;
        if ((rc = copy_states(h, h->d_phi, phi, N, off, cnt, NN, true, 0))) return rc;
        if ((rc = run_jac(h, (long)cnt, h->d_param, h->d_phi, h->d_out, 0))) return rc;
        if ((rc = copy_states(h, jac, h->d_out, N, off, cnt, nv, false, 0))) return rc;
    }
    HIP_TRY(hipStreamSynchronize(0));
    return PYJAC_OK;
}

int pyjac_hip_chem_utils(pyjac_handle* h, const double* phi, double* b, double* energy,
                         double* heat) {
    if (!h || !phi || !energy || !heat || !b) return fail(PYJAC_ERR_ARG, "null argument");
    int rc = ensure_staging(h);
    if (rc) return rc;
    const size_t N = h->problem_size, NN = h->m.NN, Ns = h->m.Ns, per = h->staged_states;
    for (size_t i = 0; i < N * NN; ++i)
        if (!std::isfinite(phi[i]))
            return fail(PYJAC_ERR_ARG, "phi of state %zu is not finite",
                        h->order == 'C' ? i / NN : i % N);
    for (size_t off = 0; off < N; off += per) {
        size_t cnt = std::min(per, N - off);
        if ((rc = copy_states(h, h->d_phi, phi, N, off, cnt, NN, true, 0))) return rc;
        HIP_TRY(pjk_host::launch_chem_utils(h->m, h->conp, h->order, (long)cnt, h->d_phi,
                                            h->d_out, h->d_out2, h->d_out3, 0));
        if ((rc = copy_states(h, energy, h->d_out, N, off, cnt, Ns, false, 0))) return rc;
        if ((rc = copy_states(h, heat, h->d_out2, N, off, cnt, Ns, false, 0))) return rc;
        if ((rc = copy_states(h, b, h->d_out3, N, off, cnt, Ns, false, 0))) return rc;
    }
    HIP_TRY(hipStreamSynchronize(0));
    return PYJAC_OK;
}

// the size checks of the _sized entry points, then the plain call
static int sized_prepare(pyjac_handle* h, size_t n, size_t got_param, size_t got_phi,
                         size_t got_out, size_t out_width, const char* out_name) {
    if (!h) return fail(PYJAC_ERR_ARG, "null handle");
    const size_t NN = h->m.NN;
    if (got_param < n)
        return fail(PYJAC_ERR_ARG, "param holds %zu values, %zu states need %zu", got_param, n, n);
    if (got_phi < n * NN)
        return fail(PYJAC_ERR_ARG, "phi holds %zu values, %zu states need %zu", got_phi, n,
                    n * NN);
    if (got_out < n * out_width)
        return fail(PYJAC_ERR_ARG, "%s holds %zu values, %zu states need %zu", out_name, got_out,
                    n, n * out_width);
    if (n != h->problem_size) return pyjac_hip_resize(h, n, h->chunk);
    return PYJAC_OK;
}

int pyjac_hip_species_rates_sized(pyjac_handle* h, size_t n, const double* param,
                                  size_t param_count, const double* phi, size_t phi_count,
                                  double* dphi, size_t dphi_count) {
    int rc = sized_prepare(h, n, param_count, phi_count, dphi_count, h ? h->m.NN : 0, "dphi");
    if (rc) return rc;
    if (!n) return PYJAC_OK;
    return pyjac_hip_species_rates(h, param, phi, dphi);
}

int pyjac_hip_jacobian_sized(pyjac_handle* h, size_t n, const double* param, size_t param_count,
                             const double* phi, size_t phi_count, double* jac, size_t jac_count) {
    int rc = sized_prepare(h, n, param_count, phi_count, jac_count, h ? nvalues(h) : 0, "jac");
    if (rc) return rc;
    if (!n) return PYJAC_OK;
    return pyjac_hip_jacobian(h, param, phi, jac);
}

int pyjac_hip_chem_utils_sized(pyjac_handle* h, size_t n, const double* phi, size_t phi_count,
                               double* b, double* energy, double* heat, size_t out_count) {
    int rc = sized_prepare(h, n, n, phi_count, out_count, h ? h->m.Ns : 0, "each output");
    if (rc) return rc;
    if (!n) return PYJAC_OK;
    return pyjac_hip_chem_utils(h, phi, b, energy, heat);
}

int pyjac_hip_num_species(const pyjac_handle* h) { return h ? h->m.Ns : -1; }
int pyjac_hip_num_reactions(const pyjac_handle* h) { return h ? h->m.Nr : -1; }
int pyjac_hip_jacobian_size(const pyjac_handle* h) { return h ? nvalues(h) : -1; }

const char* pyjac_hip_species_name(const pyjac_handle* h, int i) {
    if (!h || i < 0 || i >= (int)h->species_names.size()) return nullptr;
    return h->species_names[i].c_str();
}

const char* pyjac_hip_reaction_string(const pyjac_handle* h, int i) {
    if (!h || i < 0 || i >= (int)h->reaction_strings.size()) return nullptr;
    return h->reaction_strings[i].c_str();
}

char pyjac_hip_order(const pyjac_handle* h) { return h ? h->order : 0; }

int pyjac_hip_vector_width(const pyjac_handle* h) { return h ? 1 : -1; }

size_t pyjac_hip_required_memory(const pyjac_handle* h) {
    if (!h) return 0;
    return pjk_host::ws_doubles_per_state(h->m, true) * sizeof(double);
}

// ------------------------------------------------------------ ragged batch
}  // extern "C"

struct pyjac_batch {
    std::vector<pyjac_handle*> h;
    struct Member {
        size_t n = 0;
        const double *param = nullptr, *phi = nullptr;
        double* out = nullptr;
    };
    std::vector<Member> mem;
    hipStream_t cap = nullptr;            // capture stream
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    bool dirty[2] = {true, true};
    // Jacobian in one launch for members of both variants: the slot flags
    // (zeroed, and left zeroed by every launch) and the gather members'
    // contribution buffers
    int queue = 0;
    int* d_busy = nullptr;
    int nslots = 0;
    double* d_slots = nullptr;
    size_t slot_doubles = 0;
};

namespace {

void drop_graph(pyjac_batch* b, int k) {
    if (b->exec[k]) (void)hipGraphExecDestroy(b->exec[k]);
    if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
    b->exec[k] = nullptr;
    b->graph[k] = nullptr;
    b->dirty[k] = true;
}

// the members the queue kernel takes: exact or approximate analytical
// Jacobians that the dense or the gather variant can run
bool queue_member(const pyjac_handle* h) {
    return h->jac_type != PYJAC_JAC_FINITE_DIFFERENCE && (use_block(h, 1) || use_glob(h));
}

struct QueuePlan {
    std::vector<pjk_host::QueueLaunch> q;
    int wps = 2, nslots = 0;
    size_t lds = 0, slot = 0;
};

// the one-launch plan of batch b: members (the gather members, then the
// others by falling reaction count: the most expensive states are dispatched
// first), waves per state, LDS, contribution slots
int queue_plan(pyjac_batch* b, QueuePlan& P) {
    std::vector<int> idx;
    for (size_t i = 0; i < b->h.size(); ++i)
        if (b->mem[i].n && queue_member(b->h[i])) idx.push_back((int)i);
    if (idx.empty()) return PYJAC_OK;
    if (idx.size() > (size_t)pjk_host::RAGGED_MAX)
        return fail(PYJAC_ERR_UNSUPPORTED, "one launch takes up to %d members",
                    pjk_host::RAGGED_MAX);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int c) {
        const bool ga = use_glob(b->h[a]), gc = use_glob(b->h[c]);
        if (ga != gc) return ga;
        return b->h[a]->m.Nr > b->h[c]->m.Nr;
    });
    for (int i : idx) {
        const pyjac_handle* h = b->h[i];
        const bool g = use_glob(h);
        P.wps = std::max(P.wps, g ? glob_wps(h) : block_wps(h, 1));
        P.lds = std::max(P.lds, (size_t)(g ? h->lds_g.total : h->lds[1].total) * sizeof(double));
        if (g) P.slot = std::max(P.slot, (size_t)h->bt.n_contrib);
    }
    if (P.wps > 4) P.wps = 8;
    else if (P.wps > 2) P.wps = 4;
    const int per_cu = pjk_host::queue_resident(P.wps, P.lds);
    if (per_cu < 1) return fail(PYJAC_ERR_UNSUPPORTED, "batch kernel does not fit on a CU");
    // per XCD, twice the workgroups that can be resident on it: a claim
    // always finds a slot
    P.nslots = P.slot ? 2 * per_cu * ((b->h[idx[0]]->n_cu + 7) / 8) : 1;
    for (int i : idx) {
        const pyjac_handle* h = b->h[i];
        const auto& m = b->mem[i];
        const bool g = use_glob(h);
        P.q.push_back(pjk_host::QueueLaunch{g ? h->d_args_g : h->d_args[1], h->conp, h->m.NN,
                                            g ? 1 : 0, h->order, (long)m.n, h->m.jac_nnz,
                                            m.param, m.phi, m.out});
    }
    return PYJAC_OK;
}

// the one-launch buffers, before any capture
int prepare_queue(pyjac_batch* b) {
    QueuePlan P;
    int rc = queue_plan(b, P);
    if (rc || P.q.empty()) return rc;
    if (P.nslots > b->nslots) {
        (void)hipFree(b->d_busy);
        b->d_busy = nullptr;
        b->nslots = 0;
        HIP_TRY(hipMalloc(&b->d_busy, (size_t)(8 * P.nslots + 1) * sizeof(int)));
        HIP_TRY(hipMemset(b->d_busy, 0, (size_t)(8 * P.nslots + 1) * sizeof(int)));
        b->nslots = P.nslots;
    }
    const size_t need = P.slot * (size_t)P.nslots * 8;
    if (need > b->slot_doubles) {
        (void)hipFree(b->d_slots);
        b->d_slots = nullptr;
        b->slot_doubles = 0;
        HIP_TRY(hipMalloc(&b->d_slots, std::max<size_t>(need, 1) * sizeof(double)));
        b->slot_doubles = need;
    }
    return PYJAC_OK;
}

// every launch of one batch evaluation, on stream s
int batch_launches(pyjac_batch* b, int kernel, hipStream_t s) {
    if (kernel == 1 && b->queue) {
        QueuePlan P;
        int rc = queue_plan(b, P);
        if (rc) return rc;
        if (!P.q.empty()) {
            if (!b->d_busy || P.nslots > b->nslots ||
                P.slot * (size_t)P.nslots * 8 > b->slot_doubles)
                return fail(PYJAC_ERR_STATE, "batch buffers not prepared");
            HIP_TRY(pjk_host::launch_queue(P.wps, (int)P.q.size(), P.q.data(), P.lds, b->d_busy,
                                           P.nslots, b->d_slots, (long)P.slot, s));
        }
        // members the queue kernel does not take (finite differences)
        for (size_t i = 0; i < b->h.size(); ++i) {
            const auto& m = b->mem[i];
            if (!m.n || queue_member(b->h[i])) continue;
            if ((rc = run_jac(b->h[i], (long)m.n, m.param, m.phi, m.out, s))) return rc;
        }
        return PYJAC_OK;
    }
    std::vector<pjk_host::RaggedLaunch> rag;
    size_t lds = 0;
    int wps = 1;
    for (size_t i = 0; i < b->h.size(); ++i) {
        pyjac_handle* h = b->h[i];
        const auto& m = b->mem[i];
        if (!m.n) continue;
        if (use_block(h, kernel) && rag.size() < (size_t)pjk_host::RAGGED_MAX) {
            rag.push_back(pjk_host::RaggedLaunch{h->d_args[kernel], h->conp, h->m.NN, h->order,
                                                 (long)m.n, m.param, m.phi, m.out});
            lds = std::max(lds, (size_t)h->lds[kernel].total * sizeof(double));
            wps = std::max(wps, block_wps(h, kernel));
        } else {
            int rc = kernel == 0 ? run_rates(h, (long)m.n, m.param, m.phi, m.out, s)
                                 : run_jac(h, (long)m.n, m.param, m.phi, m.out, s);
            if (rc) return rc;
        }
    }
    if (!rag.empty())
        HIP_TRY(pjk_host::launch_ragged(kernel, wps, (int)rag.size(), rag.data(), lds, s));
    return PYJAC_OK;
}

}  // namespace

extern "C" {

int pyjac_hip_batch_create(pyjac_handle* const* handles, int count, pyjac_batch** out) {
    if (!handles || !out || count < 1) return fail(PYJAC_ERR_ARG, "null argument or empty batch");
    *out = nullptr;
    for (int i = 0; i < count; ++i)
        if (!handles[i]) return fail(PYJAC_ERR_ARG, "null handle %d", i);
    pyjac_batch* b = new pyjac_batch();
    b->h.assign(handles, handles + count);
    b->mem.resize(count);
    if (hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return fail(PYJAC_ERR_HIP, "stream creation failed");
    }
    *out = b;
    return PYJAC_OK;
}

int pyjac_hip_batch_set(pyjac_batch* b, int i, size_t n_states, const double* d_param,
                        const double* d_phi, double* d_out) {
    if (!b || i < 0 || i >= (int)b->h.size()) return fail(PYJAC_ERR_ARG, "bad batch member");
    if (n_states && (!d_param || !d_phi || !d_out)) return fail(PYJAC_ERR_ARG, "null argument");
    auto& m = b->mem[i];
    if (m.n != n_states || m.param != d_param || m.phi != d_phi || m.out != d_out) {
        m.n = n_states;
        m.param = d_param;
        m.phi = d_phi;
        m.out = d_out;
        drop_graph(b, 0);
        drop_graph(b, 1);
    }
    return PYJAC_OK;
}

int pyjac_hip_batch_slot_failures(pyjac_batch* b, int* out) {
    if (!b || !out) return fail(PYJAC_ERR_ARG, "null argument");
    *out = 0;
    if (!b->d_busy) return PYJAC_OK;
    HIP_TRY(hipMemcpy(out, b->d_busy, sizeof(int), hipMemcpyDeviceToHost));
    return PYJAC_OK;
}

int pyjac_hip_batch_set_queue(pyjac_batch* b, int on) {
    if (!b) return fail(PYJAC_ERR_ARG, "null batch");
    if (b->queue != (on ? 1 : 0)) {
        b->queue = on ? 1 : 0;
        drop_graph(b, 1);
    }
    return PYJAC_OK;
}

int pyjac_hip_batch_run(pyjac_batch* b, int kernel, int use_graph, void* stream) {
    if (!b || (kernel != 0 && kernel != 1)) return fail(PYJAC_ERR_ARG, "bad batch call");
    hipStream_t s = (hipStream_t)stream;
    if (kernel == 1 && b->queue) {
        int rc = prepare_queue(b);
        if (rc) return rc;
    }
    if (!use_graph) return batch_launches(b, kernel, s);
    if (b->dirty[kernel]) {
        // allocations and kernel attributes first: nothing may allocate
        // while the stream is captured
        for (size_t i = 0; i < b->h.size(); ++i) {
            if (!b->mem[i].n) continue;
            int rc = prepare_lane(b->h[i], kernel, (long)b->mem[i].n);
            if (rc) return rc;
        }
        int rc = batch_launches(b, kernel, b->cap);  // eager run: sets kernel attributes
        if (rc) return rc;
        HIP_TRY(hipStreamSynchronize(b->cap));
        HIP_TRY(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
        rc = batch_launches(b, kernel, b->cap);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->cap, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(PYJAC_ERR_HIP, "capture failed: %s", hipGetErrorString(e));
        b->graph[kernel] = g;
        HIP_TRY(hipGraphInstantiate(&b->exec[kernel], g, nullptr, nullptr, 0));
        b->dirty[kernel] = false;
    }
    HIP_TRY(hipGraphLaunch(b->exec[kernel], s));
    return PYJAC_OK;
}

int pyjac_hip_batch_destroy(pyjac_batch* b) {
    if (!b) return PYJAC_OK;
    drop_graph(b, 0);
    drop_graph(b, 1);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    (void)hipFree(b->d_busy);
    (void)hipFree(b->d_slots);
    delete b;
    return PYJAC_OK;
}

// ---------------------------------------------------------------- autotuner
// Times the kernel families that can run this handle's mechanism on n
// resident states and keeps the fastest: block-per-state with 1, 2, 4 or 8
// waves per state, wave-per-state, lane-per-state.
int pyjac_hip_autotune(pyjac_handle* h, int kernel, size_t n, const double* d_param,
                       const double* d_phi, double* d_out, void* stream, int* best_variant,
                       float* best_ms) {
    if (!h || (kernel != 0 && kernel != 1) || !n || !d_param || !d_phi || !d_out)
        return fail(PYJAC_ERR_ARG, "bad autotune call");
    hipStream_t s = (hipStream_t)stream;
    // candidates: (variant, occupancy target); then, for the F-order LDS
    // Jacobian, the chunk size of the winner
    struct Cand { int v, occ; };
    std::vector<Cand> cand;
    bool block = kernel == 0 ? h->block_ok_rates : h->block_ok_jac;
    bool wave = kernel == 0 ? h->wave_ok_rates : h->wave_ok_jac;
    if (block)
        for (int w : {1, 2, 4, 8}) {
            cand.push_back(Cand{10 + w, 0});
            for (int occ : {2, 4, 6}) {
                const bool has = kernel == 0 ? pjk_host::rates_block_has(w, occ)
                                             : pjk_host::jacobian_block_has(w, occ);
                if (has) cand.push_back(Cand{10 + w, occ});
            }
        }
    if (kernel == 1 && h->glob_ok_jac)
        for (int w : {2, 4, 8}) cand.push_back(Cand{20 + w, 0});
    if (wave) cand.push_back(Cand{2, 0});
    if (!(kernel == 1 && h->bt.approx_sums)) cand.push_back(Cand{1, 0});
    if (kernel == 0 && h->lane_fn) cand.push_back(Cand{30, 0});
    if (kernel == 1 && jac_module_ready(h)) cand.push_back(Cand{31, 0});
    if (kernel == 1 && ljac_lines_ready(h)) cand.push_back(Cand{32, 0});
    const int saved_variant = h->variant[kernel], saved_wps = h->wps[kernel];
    const int saved_wps_g = h->wps_g, saved_occ = h->occ[kernel];
    const long saved_chunk = h->fchunk_pref;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    auto run = [&]() {
        return kernel == 0 ? run_rates(h, (long)n, d_param, d_phi, d_out, s)
                           : run_jac(h, (long)n, d_param, d_phi, d_out, s);
    };
    // one warm-up call, then the mean of three
    auto timed = [&](float* ms) {
        int rc = run();
        if (rc) return rc;
        if (hipEventRecord(e0, s) != hipSuccess) return fail(PYJAC_ERR_HIP, "event");
        for (int r = 0; r < 3 && !rc; ++r) rc = run();
        if (rc) return rc;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
            return fail(PYJAC_ERR_HIP, "event");
        (void)hipEventElapsedTime(ms, e0, e1);
        *ms /= 3.0f;
        return PYJAC_OK;
    };
    Cand best{-1, 0};
    float best_t = 0.0f;
    long best_chunk = 0;
    int rc = PYJAC_OK;
    for (const Cand& c : cand) {
        h->variant[kernel] = c.v;
        h->occ[kernel] = c.occ;
        float ms = 0.0f;
        if ((rc = timed(&ms))) break;
        if (best.v < 0 || ms < best_t) { best = c; best_t = ms; }
    }
    // the F-order chunk of the LDS Jacobian: a few sizes around the default
    if (!rc && kernel == 1 && best.v > 10 && best.v < 20 && h->order == 'F' &&
        !getenv("PYJAC_HIP_FCHUNK")) {
        h->variant[kernel] = best.v;
        h->occ[kernel] = best.occ;
        for (long cs : {16384L, 32768L, 131072L}) {
            if (cs >= (long)n) continue;
            h->fchunk_pref = cs;
            float ms = 0.0f;
            if ((rc = timed(&ms))) break;
            if (ms < best_t) { best_t = ms; best_chunk = cs; }
        }
        h->fchunk_pref = saved_chunk;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        h->variant[kernel] = saved_variant;
        h->wps[kernel] = saved_wps;
        h->wps_g = saved_wps_g;
        h->occ[kernel] = saved_occ;
        return rc;
    }
    h->occ[kernel] = best.occ;
    if (kernel == 1) h->fchunk_pref = best_chunk;
    if (best.v == 31 || best.v == 32) {
        h->variant[kernel] = best.v;  // the Jacobian module, whatever the layout
    } else if (best.v == 30) {
        h->variant[kernel] = 0;  // the loaded module is the automatic choice
    } else if (best.v > 20) {
        // automatic choice only when the dense accumulator does not fit
        h->variant[kernel] = h->block_ok_jac ? best.v : 0;
        h->wps_g = best.v - 20;
    } else if (best.v > 10) {
        // with a module loaded the automatic choice would be the module
        const bool mod = kernel == 0 ? h->lane_fn != nullptr
                                     : jac_module_ready(h) || ljac_lines_ready(h);
        h->variant[kernel] = mod ? best.v : 0;
        h->wps[kernel] = best.v - 10;
    } else {
        h->variant[kernel] = best.v;
    }
    if (best_variant) *best_variant = best.v;
    if (best_ms) *best_ms = best_t;
    return PYJAC_OK;
}

int pyjac_hip_set_tuning(pyjac_handle* h, int kernel, int occupancy, long chunk_states) {
    if (!h || (kernel != 0 && kernel != 1) || occupancy < 0 || chunk_states < 0)
        return fail(PYJAC_ERR_ARG, "bad tuning");
    if (occupancy) {
        const int w = block_wps(h, kernel);
        const bool has = kernel == 0 ? pjk_host::rates_block_has(w, occupancy)
                                     : pjk_host::jacobian_block_has(w, occupancy);
        if (!has)
            return fail(PYJAC_ERR_UNSUPPORTED,
                        "no build of the %d-wave block kernel for occupancy %d", w, occupancy);
    }
    if (chunk_states && (kernel != 1 || chunk_states % 64))
        return fail(PYJAC_ERR_ARG, "chunk sizes are multiples of 64, for the Jacobian");
    h->occ[kernel] = occupancy;
    if (kernel == 1) h->fchunk_pref = chunk_states;
    return PYJAC_OK;
}

int pyjac_hip_get_tuning(const pyjac_handle* h, int kernel, int* occupancy, long* chunk_states,
                         int* waves_per_state) {
    if (!h || (kernel != 0 && kernel != 1)) return fail(PYJAC_ERR_ARG, "bad tuning query");
    if (occupancy) *occupancy = h->occ[kernel];
    if (chunk_states) *chunk_states = kernel == 1 ? h->fchunk_pref : 0;
    if (waves_per_state) *waves_per_state = use_glob(h) && kernel == 1 ? glob_wps(h)
                                                                       : block_wps(h, kernel);
    return PYJAC_OK;
}

int pyjac_hip_set_variant(pyjac_handle* h, int kernel, int variant) {
    if (!h || kernel < 0 || kernel > 3 || variant < 0) return fail(PYJAC_ERR_ARG, "bad variant");
    if (kernel <= 1) {
        const bool wave = kernel == 0 ? h->wave_ok_rates : h->wave_ok_jac;
        const bool block = kernel == 0 ? h->block_ok_rates : h->block_ok_jac;
        const int w = variant % 10;
        const bool wok = w == 1 || w == 2 || w == 4 || w == 8;
        const bool lane_ok = !(kernel == 1 && h->bt.approx_sums);
        bool ok = variant == 0 || (variant == 1 && lane_ok) || (variant == 2 && wave) ||
                  (variant > 10 && variant < 20 && wok && block) ||
                  (kernel == 1 && variant > 20 && variant < 30 && wok && h->glob_ok_jac) ||
                  (kernel == 0 && variant == 30 && h->lane_fn) ||
                  (kernel == 1 && variant == 31 && jac_module_ready(h)) ||
                  (kernel == 1 && variant == 32 && ljac_lines_ready(h));
        if (!ok)
            return fail(PYJAC_ERR_UNSUPPORTED, "variant %d cannot run this mechanism's %s", variant,
                        kernel == 0 ? "species rates" : "Jacobian");
    }
    h->variant[kernel] = variant;
    return PYJAC_OK;
}

int pyjac_hip_get_variant(const pyjac_handle* h, int kernel) {
    if (!h || kernel < 0 || kernel > 3) return -1;
    return h->variant[kernel];
}

}  // extern "C"
