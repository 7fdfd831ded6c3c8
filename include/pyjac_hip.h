/*
 * pyjac_hip.h -- C ABI of the MI355X (gfx950) species-rate / analytical
 * Jacobian library.
 *
 * The reference exposes each generated kernel as a C++ class
 * (kernel_utils/common/kernel.hpp.in:86-336): <Name>Kernel(problem_size,
 * work_size), resize(), operator()(double* ...), requiredMemorySize(),
 * speciesNames(), reactionStrings(), numSpecies(), numReactions(), order(),
 * and Python reaches it through the Cython wrapper
 * (pywrap/pyjacob_wrapper.pyx.in:1-92).  This header replaces that class
 * with plain C entry points (plain pointers and sizes) that a ctypes,
 * cgo or JNI binding can call.  Errors are returned as status codes; the
 * reference's cassert/exit(-1) path (kernel_utils/c/error_check.cpp:9-15)
 * has no counterpart here.
 *
 * Data layout is the reference's (docs/data_ordering.rst,
 * core/array_creator.py:2039-2067):
 *   phi   [n_states][NN] (order 'C') or [NN][n_states] (order 'F'),
 *         phi = [T, V (CONP) | P (CONV), n_0 .. n_{Ns-2}], NN = Ns + 1
 *   param [n_states]  P (CONP) or V (CONV)
 *   dphi  same shape as phi
 *   jac   full:   C: [n][NN(row)][NN(col)]   F: [NN(col)][NN(row)][n]
 *         sparse: C: CRS values [n][nnz]     F: CCS values [nnz][n]
 * All floating point is IEEE binary64.
 */
#ifndef PYJAC_HIP_H
#define PYJAC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* version 2 added the species and reaction names at the end of
   pyjac_mech_tables; version 1 tables (no names) are still accepted */
#define PYJAC_MECH_ABI_VERSION 2

/* status codes */
#define PYJAC_OK 0
#define PYJAC_ERR_ARG 1          /* bad argument (null pointer, bad size) */
#define PYJAC_ERR_HIP 2          /* a HIP runtime call failed */
#define PYJAC_ERR_ALLOC 3        /* device or host allocation failed */
#define PYJAC_ERR_UNSUPPORTED 4  /* option combination not supported */
#define PYJAC_ERR_STATE 5        /* handle not ready (e.g. not resized) */
#define PYJAC_ERR_IO 6           /* file missing, unreadable or too short */

/* Arrhenius forms of simple_form / fall_form (rate_subs.py:100-124) */
#define PYJAC_ARR_A 0          /* kf = A */
#define PYJAC_ARR_AT_B 1       /* kf = A * T^b, integer b */
#define PYJAC_ARR_FULL 2       /* kf = exp(logA + b logT - Ta/T) */
#define PYJAC_ARR_NO_TA 3      /* kf = exp(logA + b logT) */
#define PYJAC_ARR_NO_B 4       /* kf = exp(logA - Ta/T) */

/* Jacobian formats / types (core/enum_types.py:85-115) */
#define PYJAC_JAC_FULL 0
#define PYJAC_JAC_SPARSE 1
#define PYJAC_JAC_EXACT 0
#define PYJAC_JAC_APPROXIMATE 1
#define PYJAC_JAC_FINITE_DIFFERENCE 2

/* finite-difference modes (FiniteDifferenceMode, core/enum_types.py) */
#define PYJAC_FD_FORWARD 0
#define PYJAC_FD_CENTRAL 1
#define PYJAC_FD_BACKWARD 2

/*
 * Mechanism tables.  Same content as the dict built by the reference's
 * assign_rates (core/rate_subs.py:74-537) and determine_jac_inds
 * (core/create_jacobian.py:68-282); the caller owns every array.
 * Indices are int32 (core/array_creator.py:554).
 */
typedef struct pyjac_mech_tables {
    int32_t abi_version;
    int32_t n_species;        /* Ns, bath gas last */
    int32_t n_reactions;      /* Nr */
    int32_t n_rxn_spec;       /* number of (reaction, species) pairs */
    int32_t n_simple, n_rev, n_thd, n_thd_spec, n_fall, n_troe, n_sri;
    int32_t n_plog, plog_max_p;
    int32_t n_cheb, cheb_max_t, cheb_max_p;
    int32_t n_rxn_has_ns;
    int32_t jac_nnz;          /* non-zeros of the exact/approximate pattern */

    /* reaction -> participating species (sorted), rate_subs.py:126-160 */
    const int32_t *rxn_off;   /* [Nr+1] */
    const int32_t *rxn_spec;  /* [n_rxn_spec] */
    const int32_t *nu_prod;   /* [n_rxn_spec] */
    const int32_t *nu_reac;   /* [n_rxn_spec] */
    const int32_t *nu_sum;    /* [Nr] net order change, for Kc */

    /* per-reaction indices into the sub-tables, -1 when absent */
    const int32_t *rev_mask;     /* [Nr] -> reverse index   */
    const int32_t *thd_mask;     /* [Nr] -> third-body index */
    const int32_t *simple_mask;  /* [Nr] -> simple index     */
    const int32_t *plog_mask;    /* [Nr] -> PLOG index       */
    const int32_t *cheb_mask;    /* [Nr] -> Chebyshev index  */
    const int32_t *fall_mask;    /* [Nr] -> falloff index    */

    /* simple Arrhenius [n_simple]; A holds log(A) for the exp forms */
    const int32_t *simple_form;
    const double *simple_A, *simple_b, *simple_Ta;

    /* falloff / chemically activated [n_fall] (rate_subs.py:326-385) */
    const int32_t *fall_map;     /* reaction index */
    const int32_t *fall_type;    /* 0 falloff (k0 from LOW), 1 chem. act. */
    const int32_t *fall_form;    /* Arrhenius form of the LOW/HIGH rate */
    const int32_t *fall_blend;   /* 1 Lindemann, 2 Troe, 3 SRI */
    const int32_t *fall_param;   /* index into troe_* or sri_* */
    const int32_t *fall_to_thd;  /* third-body index of the reaction */
    const double *fall_A, *fall_b, *fall_Ta;
    const double *troe_a, *troe_T1inv, *troe_T2, *troe_T3inv;  /* [n_troe] */
    const double *sri_a, *sri_b, *sri_c, *sri_d, *sri_e;        /* [n_sri] */

    /* third bodies [n_thd] (rate_subs.py:387-426) */
    const int32_t *thd_map;   /* reaction index */
    const int32_t *thd_type;  /* 1 mix, 2 single species, 3 unity */
    const int32_t *thd_off;   /* [n_thd+1] into thd_spec / thd_eff */
    const int32_t *thd_spec;  /* [n_thd_spec] */
    const double *thd_eff;    /* [n_thd_spec] */

    /* PLOG [n_plog]; params [4][n_plog][plog_max_p] = log P, log A, b, Ta */
    const int32_t *plog_map, *plog_np;
    const double *plog_params;

    /* Chebyshev [n_cheb]; params [n_cheb][cheb_max_t][cheb_max_p],
       plim [n_cheb][2] = log P limits, tlim [n_cheb][2] = 1/T limits */
    const int32_t *cheb_map, *cheb_nt, *cheb_np;
    const double *cheb_params, *cheb_plim, *cheb_tlim;

    /* NASA-7 [Ns][7] low / high coefficients, T_mid [Ns],
       mw_post [Ns-1] = mw_k / mw_ns */
    const double *a_lo, *a_hi, *T_mid, *mw_post;

    /* reactions in which the bath gas takes part [n_rxn_has_ns] */
    const int32_t *rxn_has_ns;

    /* Jacobian pattern, CCS (F order) and CRS (C order) */
    const int32_t *ccs_col_ptr;  /* [NN+1] */
    const int32_t *ccs_row_ind;  /* [jac_nnz] */
    const int32_t *crs_row_ptr;  /* [NN+1] */
    const int32_t *crs_col_ind;  /* [jac_nnz] */

    /* (abi_version >= 2) species names [Ns] in the tables' order (bath gas
       last) and reaction strings [Nr], NUL-terminated; either may be NULL.
       The generated code holds them as the static speciesNames() /
       reactionStrings() tables (kernel.hpp.in:101-143). */
    const char *const *species_names;
    const char *const *reaction_strings;
} pyjac_mech_tables;

typedef struct pyjac_handle pyjac_handle;

/*
 * Tables from a file (no counterpart in the reference, whose generated code
 * holds them as constants): the binary written by pyJac's
 * MechTables.save_binary -- magic "PJHIPTB1", int32 number of scalar fields
 * and their int32 values in struct order, int32 number of arrays, then per
 * array in struct order: uint8 type (0 int32, 1 float64), int64 length,
 * data; optionally the trailer "PJNAMES1" with the species names and the
 * reaction strings (per list an int32 count, per name an int32 length and
 * its bytes), which fills the struct's name fields (NULL without it).
 * Lets C and C++ callers use the library without Python.  *out is
 * owned by the caller and released with pyjac_hip_tables_free.
 * PYJAC_ERR_IO: missing or short file; PYJAC_ERR_ARG: not such a file, or
 * another table layout version.
 */
int pyjac_hip_tables_load(const char *path, pyjac_mech_tables **out);
int pyjac_hip_tables_free(pyjac_mech_tables *tables);

/* Last error message of the calling thread ("" when none). */
const char *pyjac_hip_last_error(void);

/*
 * Create a handle bound to one mechanism and option set; uploads the tables
 * to the current device.  Replaces the generated <Name>Kernel constructor
 * (kernel.cpp.in:40-76).
 *   conp        1 constant pressure, 0 constant volume (utils.py:972)
 *   order       'C' or 'F' (utils.py:944)
 *   jac_format  PYJAC_JAC_FULL or PYJAC_JAC_SPARSE (utils.py:1014)
 *   jac_type    PYJAC_JAC_EXACT, PYJAC_JAC_APPROXIMATE (tables on the
 *               approximate pattern) or PYJAC_JAC_FINITE_DIFFERENCE
 *               (utils.py:999)
 */
int pyjac_hip_create(const pyjac_mech_tables *mech, int conp, char order,
                     int jac_format, int jac_type, pyjac_handle **out);

/* Release all device memory of the handle (~Kernel, kernel.cpp.in:300). */
int pyjac_hip_destroy(pyjac_handle *h);

/*
 * Size the internal staging buffers for host-pointer calls of up to
 * problem_size states (Kernel::resize, kernel.cpp.in:60-76).  work_size is
 * accepted for signature compatibility and sets the number of states moved
 * per staging chunk when non-zero.
 */
int pyjac_hip_resize(pyjac_handle *h, size_t problem_size, size_t work_size);

/* Host-pointer entry points: copy in, compute, copy out, synchronous.
   Replace <Name>Kernel::operator() (kernel.cpp.in:572-640). */
int pyjac_hip_species_rates(pyjac_handle *h, const double *param,
                            const double *phi, double *dphi);
int pyjac_hip_jacobian(pyjac_handle *h, const double *param,
                       const double *phi, double *jac);
/* chem_utils kernel (rate_subs.py:3391-3452): b = S/R - H/RT - ln T,
   spec_energy = h (CONP) or u (CONV), spec_heat = cp or cv; outputs shaped
   like phi but with Ns values per state.  Argument order of the Cython
   wrapper's PyChem_utilsKernel(phi, b, specific_energy, specific_heat)
   (pywrap/pywrap_gen.py:176-197: inputs, then the outputs in
   kernel_argument_ordering's order, utils.py:203-277); the generated
   device kernel is chem_utils(t, b, phi, h|u, cp|cv, rwk)
   (examples/opencl/ch4/chem_utils.oclh:23), i.e. the same arrays with the
   unused time array and the work buffer dropped and phi moved first. */
int pyjac_hip_chem_utils(pyjac_handle *h, const double *phi, double *b,
                         double *spec_energy, double *spec_heat);

/* Size-checked host entry points: the same calls on n_states states with
   the element count of every array; an array shorter than the layout
   needs (param n_states, phi and dphi n_states * NN, jac n_states *
   pyjac_hip_jacobian_size, chem_utils outputs n_states * Ns) is
   PYJAC_ERR_ARG before anything is read or written.  The handle is resized
   to n_states first when its problem size differs.  No counterpart in the
   reference, whose operator() trusts problem_size
   (kernel_utils/common/kernel.cpp.in:572-640). */
int pyjac_hip_species_rates_sized(pyjac_handle *h, size_t n_states, const double *param,
                                  size_t param_count, const double *phi, size_t phi_count,
                                  double *dphi, size_t dphi_count);
int pyjac_hip_jacobian_sized(pyjac_handle *h, size_t n_states, const double *param,
                             size_t param_count, const double *phi, size_t phi_count,
                             double *jac, size_t jac_count);
int pyjac_hip_chem_utils_sized(pyjac_handle *h, size_t n_states, const double *phi,
                               size_t phi_count, double *b, double *spec_energy,
                               double *spec_heat, size_t out_count);

/* Every host-pointer entry point rejects a NaN or an infinity in param or
   phi with PYJAC_ERR_ARG, naming the first such state in
   pyjac_hip_last_error(); the reference computes on such states unchecked
   (its test harness only rejects non-finite outputs,
   functional_tester/test.py:1153-1156).  The device-pointer entry points
   do the same check on the device when the handle's input check is on
   (off by default: it adds one scan of the inputs and a synchronisation
   of the stream per call). */
int pyjac_hip_set_input_check(pyjac_handle *h, int on);

/* Per-state finiteness report of a device array of n_states states with
   width values each in the handle's data order (phi / dphi: NN, jac:
   pyjac_hip_jacobian_size, param: 1): d_flags[s] (device int32 [n_states],
   may be NULL) = 1 when state s holds a NaN or an infinity, else 0;
   *n_bad = the number of such states and *first_bad the smallest one
   (n_states when none).  Synchronises `stream`. */
int pyjac_hip_nonfinite_dev(pyjac_handle *h, size_t n_states, const double *d_arr, int width,
                            int32_t *d_flags, size_t *n_bad, size_t *first_bad, void *stream);

/* Device-pointer entry points: n_states states already resident in device
   memory in the handle's data order, launched on `stream` (a hipStream_t,
   NULL for the default stream), asynchronous.  A handle's calls must not
   overlap in time (they share its work buffers and, for F-order Jacobians,
   its second stream): serialise them on one stream, or use one handle per
   concurrent stream. */
int pyjac_hip_species_rates_dev(pyjac_handle *h, size_t n_states,
                                const double *d_param, const double *d_phi,
                                double *d_dphi, void *stream);
int pyjac_hip_jacobian_dev(pyjac_handle *h, size_t n_states,
                           const double *d_param, const double *d_phi,
                           double *d_jac, void *stream);

/* The Jacobian device-pointer entry point without any allocation.  No
   counterpart in the reference: for callers that cannot call hipMalloc
   (code under HIP graph capture, tight serving loops, harnesses that forbid
   it).  Same arguments and results as pyjac_hip_jacobian_dev; the device
   memory the F-order path (C-order chunks and their transposes) and the
   device-memory gather variant need beyond the inputs and outputs comes
   from a fixed scratch area of 64 MiB per GPU inside the library, held from
   the library's load on a GPU until the process ends.  Every call on one
   GPU shares that area: calls that overlap in time on one GPU race (issue
   them on one stream).  PYJAC_ERR_UNSUPPORTED when 64 states' buffers
   exceed the area (e.g. a full-format F-order Jacobian of 500 species),
   for the lane- and wave-per-state variants (their per-wave work buffers),
   finite-difference handles and handles with the input check on. */
int pyjac_hip_jacobian_dev_noalloc(pyjac_handle *h, size_t n_states, const double *d_param,
                                   const double *d_phi, double *d_jac, void *stream);

/* Options of a PYJAC_JAC_FINITE_DIFFERENCE handle (finite_difference_jacobian,
   core/create_jacobian.py:4523-4526): mode PYJAC_FD_FORWARD / _BACKWARD
   (order 1-6) or PYJAC_FD_CENTRAL (order 2, 4, 6, 8), and the relative and
   absolute tolerances the steps come from.  Defaults: forward, order 1,
   rtol 1e-8, atol 1e-15.  The Jacobian entry points of such a handle then
   return the finite-difference Jacobian, evaluated with the species-rate
   kernels on C-order chunks of perturbed states. */
int pyjac_hip_set_fd_options(pyjac_handle *h, int mode, int order, double rtol, double atol);

/* Introspection (Kernel::numSpecies etc., kernel.hpp.in:101-268). */
int pyjac_hip_num_species(const pyjac_handle *h);
int pyjac_hip_num_reactions(const pyjac_handle *h);
/* Kernel::speciesNames()[i] / reactionStrings()[i] (kernel.hpp.in:101-143):
   NULL when i is out of range or the tables carried no names (version 1
   tables).  The strings belong to the handle. */
const char *pyjac_hip_species_name(const pyjac_handle *h, int i);
const char *pyjac_hip_reaction_string(const pyjac_handle *h, int i);
/* Kernel::order() (kernel.hpp.in:122): 'C' or 'F' (0 for a null handle). */
char pyjac_hip_order(const pyjac_handle *h);
/* Kernel::vectorWidth() (kernel.hpp.in:139-147): states one kernel
   instance evaluates together.  The block-per-state kernels give each
   state its own workgroup, so 1. */
int pyjac_hip_vector_width(const pyjac_handle *h);
int pyjac_hip_jacobian_size(const pyjac_handle *h); /* NN*NN or nnz */
/* Device workspace bytes per state of the Jacobian / rates kernels
   (Kernel::requiredMemorySize, kernel.hpp.in:220). */
size_t pyjac_hip_required_memory(const pyjac_handle *h);

/* Kernel variant of one handle; no counterpart in the reference, whose
   generated code fixes the layout at generation time (--width/--depth).
   kernel: 0 species rates, 1 Jacobian, 2 chem utils.  variant: 0 automatic
   (the tuned choice), 1 lane-per-state kernels, 2 wave-per-state kernels,
   10 + W block-per-state kernels with W = 1, 2, 4 or 8 waves per state,
   20 + W (Jacobian) block-per-state kernels that gather the Jacobian in
   device memory (the automatic choice when the dense accumulator does not
   fit in LDS), 30 (species rates) the mechanism-specialised module loaded
   by pyjac_hip_load_module (the automatic choice once loaded), 31
   (Jacobian) the mechanism-specialised Jacobian module storing the output
   directly (the automatic choice once loaded in F order), 32 (Jacobian, C
   order) the module's F-layout store kernel writing lines of a chunk
   (entry stride = the chunk's states) that a tiled transpose writes out as
   C-order rows (the automatic choice once loaded in C order).
   PYJAC_ERR_UNSUPPORTED when this mechanism cannot run the variant (its
   arrays exceed the variant's LDS, or no module is loaded). */
int pyjac_hip_set_variant(pyjac_handle *h, int kernel, int variant);
int pyjac_hip_get_variant(const pyjac_handle *h, int kernel);

/*
 * Mechanism-specialised species rates.  The reference generates its kernels
 * per mechanism with the rate parameters as literals (libgen compiles them,
 * pyjac/libgen/libgen.py:24-283; examples/opencl/ch4/species_rates.oclh);
 * pyjac.libgen.lanegen does the same for gfx950: a code object whose kernel
 * evaluates one state per lane in straight-line code.  load_module attaches
 * such a code object to a handle made from the same tables: it checks the
 * module's fingerprint (the tables' 64-bit hash, the species and reaction
 * counts and the module version) and refuses a module generated from other
 * tables (PYJAC_ERR_ARG) or a file it cannot load (PYJAC_ERR_IO).  Once
 * loaded, the handle's species rates use it (variant 30, every data order and
 * both CONP / CONV), the finite-difference Jacobian's perturbed states
 * included.
 *
 * pyjac.libgen.lanejac generates the analytical Jacobian the same way
 * (create_jacobian.py, generated per mechanism by the reference): one state
 * per lane, a kernel that evaluates the rates, the temperature column and
 * the per-reaction values of the other columns into a scratch array, and a
 * kernel per output layout that stores every entry, column by column.  Such
 * a module is five code objects (the first kernel, and the store kernel of
 * sparse C, sparse F, full C and full F); load_module attaches each one
 * after the same fingerprint check and a check that the module's entry list
 * is the handle's sparsity pattern (an approximate-Jacobian handle refuses
 * it).  With the first kernel and the store kernel of the handle's layout
 * loaded, the handle's analytical Jacobian uses the module: variant 31 in F order,
 * variant 32 in C order; its scratch array (about 5.4 KB per GRI state, for a
 * fixed chunk of 131,072 states), and in C order a chunk of entries (15.9 KB
 * per GRI state, sparse), are allocated on the first call.  pyjac_hip_jacobian_dev_noalloc keeps the block
 * kernels.  tables_hash gives the fingerprint of a tables struct.
 */
int pyjac_hip_tables_hash(const pyjac_mech_tables *tables, unsigned long long *out);
int pyjac_hip_load_module(pyjac_handle *h, const char *path);

/*
 * Ragged batches (no counterpart in the reference, which builds one kernel
 * per mechanism): several handles -- different mechanisms or option sets --
 * each with its own resident states, evaluated by one call.  Members that the
 * block-per-state kernels take share one launch (up to 8 per launch); the
 * others get their own launches.  With use_graph the launches are captured
 * into a HIP graph on the first call after a change and replayed afterwards.
 * kernel: 0 species rates, 1 Jacobian.  Calls on one batch are not
 * re-entrant.
 */
typedef struct pyjac_batch pyjac_batch;
int pyjac_hip_batch_create(pyjac_handle *const *handles, int count, pyjac_batch **out);
int pyjac_hip_batch_set(pyjac_batch *b, int member, size_t n_states, const double *d_param,
                        const double *d_phi, double *d_out);
int pyjac_hip_batch_run(pyjac_batch *b, int kernel, int use_graph, void *stream);
int pyjac_hip_batch_destroy(pyjac_batch *b);
/* Jacobian batches in one launch (no counterpart in the reference; its
   queue-based driver, pyjac/core/driver_kernels.py:383, is unimplemented):
   with on = 1 every analytical-Jacobian member, whether its mechanism takes
   the dense LDS accumulator or the device-memory gather, is evaluated by one
   kernel launch (finite-difference members still get their own launches).
   The gather members' workgroups are dispatched first and take their
   contribution buffers from a pool with more slots than workgroups can be
   resident; F-order members are written in place (no chunk transposes).  Up
   to 8 members; the pool is allocated by the first run after a change.
   slot_failures reports claims that found no free slot (always 0 unless the
   pool is too small; such a state's output is not written). */
int pyjac_hip_batch_set_queue(pyjac_batch *b, int on);
int pyjac_hip_batch_slot_failures(pyjac_batch *b, int *out);

/*
 * Autotuner (no counterpart in the reference): times every kernel family that
 * can run the handle's mechanism -- block-per-state with 1, 2, 4 or 8 waves
 * per state and each built occupancy target, wave-per-state, lane-per-state
 * -- on n resident states (results land in d_out), then, for an F-order
 * Jacobian of the LDS block kernel, chunks of 16384, 32768 and 131072 states
 * against the default 65536, and makes the fastest the handle's automatic
 * choice.
 * Reports the winning variant id and its time per call.
 */
int pyjac_hip_autotune(pyjac_handle *h, int kernel, size_t n, const double *d_param,
                       const double *d_phi, double *d_out, void *stream, int *best_variant,
                       float *best_ms);

/* The other knobs the autotuner searches (no counterpart in the reference):
   the occupancy target of the block kernels (waves per SIMD the register
   allocator is held to; 0 = the default build, 3 for the Jacobian and 5 for
   the species rates; 2 and 4 are built for the Jacobian with one or two
   waves per state, 4 and 6 for the one-wave species rates) and the number of
   states per F-order chunk of the Jacobian (0 = the default; multiples of
   64).  pyjac_hip_autotune sets both to its choice; get_tuning also reports
   the waves per state in use. */
int pyjac_hip_set_tuning(pyjac_handle *h, int kernel, int occupancy, long chunk_states);
int pyjac_hip_get_tuning(const pyjac_handle *h, int kernel, int *occupancy, long *chunk_states,
                         int *waves_per_state);

/*
 * Initial conditions on disk (read_initial_conditions,
 * kernel_utils/common/read_initial_conditions.cpp.in:41-116): a binary file of
 * num_states records of NN + 1 doubles, C order, each
 * [T, param, V|P, n_0 .. n_{Ns-2}] (the extra variable is the state's V for
 * CONP and P for CONV; param is the other one).  Fills host param[num_states]
 * and host phi in the given order ('C': [n][NN], 'F': [NN][n]).  Where the
 * reference prints and exits, this returns PYJAC_ERR_IO (missing file, fewer
 * records than asked); a longer file is read up to num_states records, as the
 * reference does.  nn is NN = Ns + 1 (the reference bakes it in).
 */
int pyjac_hip_read_initial_conditions(const char *filename, size_t num_states, int nn,
                                      double *param, double *phi, char order);
/* Raw dump of count doubles (write_data, kernel_utils/common/write_data.hpp:15-32),
   the format the reference's validation outputs use; PYJAC_ERR_IO on failure. */
int pyjac_hip_write_data(const char *filename, const double *arr, size_t count);

#ifdef __cplusplus
}
#endif

#endif /* PYJAC_HIP_H */
