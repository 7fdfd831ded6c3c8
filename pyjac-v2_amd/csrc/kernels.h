// kernels.h -- host-side launchers of the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "device_mech.h"

namespace pjk_host {

// ---- lane-per-state kernels (pyjac_kernels.hip): per-wave workspace in
// device memory, any mechanism size
size_t ws_doubles_per_state(const DevMech& m, bool jac);

hipError_t launch_species_rates(const DevMech& m, int conp, char order, long n,
                                const double* param, const double* phi, double* dphi,
                                double* ws, int grid, int block, hipStream_t stream);

hipError_t launch_jacobian(const DevMech& m, int conp, char order, long n, const double* param,
                           const double* phi, double* jac, double* ws, int grid, int block,
                           hipStream_t stream);

hipError_t launch_chem_utils(const DevMech& m, int conp, char order, long n, const double* phi,
                             double* energy, double* heat, double* b, hipStream_t stream);

// ---- finite-difference Jacobian (pyjac_fd.hip): C-order chunk buffers
hipError_t launch_fd_gather(long cnt, int NN, long n, long s0, char order, const double* param,
                            const double* phi, double* param_c, double* phi_c, hipStream_t stream);
hipError_t launch_fd_steps(long cnt, int NN, const double* phi_c, const double* f0, double rtol,
                           double atol, double* r, hipStream_t stream);
hipError_t launch_fd_perturb(long cnt, int NN, int K, const int* xk, const double* param_c,
                             const double* phi_c, const double* r, double* pparam, double* pphi,
                             hipStream_t stream);
hipError_t launch_fd_combine(long cnt, int NN, int K, const double* yk, const double* pf,
                             const double* r, const int* acc_map, int nout, long n, long s0,
                             char order, double* jac, hipStream_t stream);

// ---- wave-per-state kernels (pyjac_wave.hip): per-state arrays in LDS
struct WaveTables {
    const void* rx_i;  // int4 [Nr]: kind | (nu_sum + 128) << 8, rev, thd, n species
    const void* rx_d;  // double4 [Nr]: A, b, Ta, PLOG/Chebyshev rate index
    const void* rx_s;  // int [8 Nr]: species | nu_reac << 16 | nu_prod << 24, padded
};

// species slots of a reaction record
constexpr int WAVE_MAXSP = 6;

size_t wave_lds_bytes(const DevMech& m, bool jac, int jac_vals);

hipError_t launch_rates_wave(const DevMech& m, int conp, char order, long n, const double* param,
                             const double* phi, double* dphi, const WaveTables& wt, int grid,
                             hipStream_t stream);

hipError_t launch_jacobian_wave(const DevMech& m, int conp, char order, long n,
                                const double* param, const double* phi, double* jac,
                                const WaveTables& wt, int grid, hipStream_t stream);

// ---- block-per-state kernels (pyjac_block.hip): WPS waves per state,
// per-state arrays and the dense Jacobian accumulator in LDS
struct BlockTables {
    const void* rec_i;       // int4 [Nr] reaction records sorted by structure
    const void* rec_d;       // double4 [Nr]
    const void* rec_s;       // int [8 Nr] species slots (slot 6: a row species)
    const int* grp_flags;    // [ceil(Nr / 64)] flags of each group of 64 records
    const int* thd_order;    // [n_thd] third-body entries: plain, Lindemann, SRI, Troe
    const double* thd_mod;   // [n_thd] efficiency of the bath gas in the entry's d/dT, d/dE
    const double* effns;     // [n_thd] d(ci)/dn_j through the bath gas
    const int* thd_pairs;    // [n_thd][8] per species slot: coefficient pair or -1
    const void* combo;       // int4 [n_combo]: pair, dense index, efficiency (2 ints)
    const void* out_pos;     // int2 [jac_nnz]: dense index, row of the bath-gas sum or -1
    int n_pairs, n_combo;
    // variant gathering in device memory (dense Jacobian larger than LDS)
    const int* acc_map;      // [NN * NN] dense index -> output entry, -1 if not stored
    const int* rec_pos;      // [Nr][8][WAVE_MAXSP]: per record, where the contribution of
                             // (species column slot, row slot) goes: buffer slot >= 0,
                             // LDS slot -2 - p, or -1 for none
    const int* ent_start;    // [jac_nnz + 1]: buffer slots of each output entry (<= GC)
    const int* ent_long;     // [jac_nnz]: LDS slot of an entry with more contributions, or -1
    const void* gcombo;      // int4 [n_combo]: pair, target (as rec_pos), efficiency
    const int* ent_rc;       // [jac_nnz]: (row << 16) | column of each output entry
    int n_contrib, n_long;
    // approximate Jacobian on a sparse pattern: rows T and V|P sum only the
    // stored species entries, bath-gas sums included
    int approx_sums;
};

// LDS offsets (in doubles) of one workgroup's arrays
struct BlockLds {
    int conc, nmol, b, db, wdot, energy, heat, dheat, U;
    int ci, nfac, dTf, dEf, Cp, mw, ksp, dkT, dkP, jac, total;
    int aux;  // device-memory variant: columns T, V|P and the many-contribution entries
};

// everything the block kernels read besides the states: one copy in device
// memory per handle and kernel, passed by pointer so the tables are fetched
// through the scalar cache when used
struct BlockArgs {
    DevMech m;
    BlockTables bt;
    BlockLds L;
};

// dense: the Jacobian accumulator in LDS (false: the device-memory variant)
BlockLds block_lds(const DevMech& m, bool jac, int n_pairs, bool dense = true, int n_long = 0);
size_t block_lds_bytes(const DevMech& m, bool jac, int n_pairs, bool dense = true,
                       int n_long = 0);
// contributions an entry of the device-memory variant gathers from the buffer
constexpr int GATHER_MAX = 4;

// one workgroup of 64 * wps threads per state
// occ: the occupancy target of the instantiation (0: the default; which
// (wps, occ) pairs exist: rates_block_has / jacobian_block_has)
hipError_t launch_rates_block(const DevMech& m, const BlockArgs* args, const BlockLds& L, int wps,
                              int conp, char order, long n, const double* param,
                              const double* phi, double* dphi, hipStream_t stream, int occ = 0);
bool rates_block_has(int wps, int occ);
bool jacobian_block_has(int wps, int occ);

// Second stream and events of the pipelined F-order path
struct FPipeline {
    hipStream_t aux;
    hipEvent_t computed[2], transposed[2];
};

// F order with chunk_buf: C-order chunks of chunk_states states (a multiple
// of 64) and a tiled transpose per chunk; otherwise written directly.  With
// pipe, chunk_buf holds two chunks and the transposes run on pipe->aux
// beside the next chunk's Jacobians.
hipError_t launch_jacobian_block(const DevMech& m, const BlockArgs* args, const BlockLds& L,
                                 int wps, int conp, char order, long n, const double* param,
                                 const double* phi, double* jac, double* chunk_buf,
                                 long chunk_states, hipStream_t stream,
                                 const FPipeline* pipe = nullptr, int occ = 0);

// count states' entries, F-order lines of a chunk (the entry at F-layout
// position p of state i at src[p * count + i]), into C-order output rows
// s0 .. s0 + count - 1 of nout values each, C-order entry e from position
// fpos[e] (device array of nout; the specialised Jacobian module's C-order
// path)
hipError_t launch_entries_to_c(const double* src, long count, int nout, double* jac, long s0,
                               const int* fpos, hipStream_t stream);

// one workgroup of 64 * wps threads per state, Jacobian gathered in device
// memory, chunk_states states (a multiple of 64) per launch: chunk_buf holds
// their contributions (n_contrib each) and, in F order, their entries, which
// a tiled transpose moves into the output; C order writes entries in place
hipError_t launch_jacobian_global(const DevMech& m, const BlockArgs* args, const BlockLds& L,
                                  int wps, int conp, char order, long n, const double* param,
                                  const double* phi, double* jac, double* chunk_buf,
                                  long chunk_states, size_t n_contrib, hipStream_t stream);

// ---- finiteness scan (pyjac_checks.hip): flags[s] = 1 for every state s of
// a (width values per state, data order `order`) holding a NaN or an
// infinity (flags may be null); counters[0] += their number, counters[1] =
// min(counters[1], smallest such state)
hipError_t launch_nonfinite(long n, int width, const double* a, char order, int32_t* flags,
                            unsigned long long* counters, hipStream_t stream);

// ---- the fixed scratch area of the Jacobian entry point that allocates
// nothing (pyjac_noalloc.hip): 64 MiB per GPU
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
// device address of the scratch area on the current device
hipError_t noalloc_scratch(double** out);

// ---- ragged batch: several handles' states in one block-kernel launch
constexpr int RAGGED_MAX = 8;
struct RaggedLaunch {
    const BlockArgs* args;  // device copy of the member's BlockArgs (its kernel's)
    int conp, NN;
    char order;
    long n;
    const double *param, *phi;
    double* out;
};

// kernel 0: species rates, 1: Jacobian; lds = the largest member's bytes
hipError_t launch_ragged(int kernel, int wps, int count, const RaggedLaunch* members,
                         size_t lds, hipStream_t stream);

// ---- ragged Jacobian batch in one launch, members of both Jacobian variants
struct QueueLaunch {
    const BlockArgs* args;  // the member's dense-kernel or gather-kernel BlockArgs
    int conp, NN, gather;
    char order;
    long n, nout;           // states, stored values per state
    const double *param, *phi;
    double* out;
};
// busy: 8 * nslots + 1 zeroed ints (a count of failed slot claims, then the
// slot flags, left zeroed by every launch); slots: 8 * nslots buffers of
// `slot` doubles for the gather members; nslots per XCD, above the
// workgroups resident on one XCD at once; wps 2, 4 or 8
hipError_t launch_queue(int wps, int count, const QueueLaunch* members, size_t lds, int* busy,
                        int nslots, double* slots, long slot, hipStream_t stream);
// workgroups of the queue kernel that fit on one CU (0 on error)
int queue_resident(int wps, size_t lds);

}  // namespace pjk_host
