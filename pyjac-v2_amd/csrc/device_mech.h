// device_mech.h -- mechanism tables as seen by the gfx950 kernels.
//
// The host packs every table of a pyjac_mech_tables (include/pyjac_hip.h)
// into one device allocation; DevMech holds the device pointers and the
// sizes and is passed to the kernels by value (it lives in the kernel
// argument segment, so all table reads are wave-uniform scalar loads).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct DevMech {
    int Ns, Nr, NN, ns;
    int n_rev, n_thd, n_fall, n_plog, plog_max_p, n_cheb, cheb_max_t, cheb_max_p;
    int n_rxn_has_ns;
    int jac_nnz;  // values stored per state: NN*NN (full) or nnz (sparse)
    // reaction -> species
    const int *rxn_off, *rxn_spec, *nu_prod, *nu_reac, *nu_sum;
    const int *rev_mask, *thd_mask, *simple_mask, *plog_mask, *cheb_mask, *fall_mask;
    const int *simple_form;
    const double *simple_A, *simple_b, *simple_Ta;
    const int *fall_map, *fall_type, *fall_form, *fall_blend, *fall_param, *fall_to_thd;
    const double *fall_A, *fall_b, *fall_Ta;
    const double *troe_a, *troe_T1inv, *troe_T2, *troe_T3inv;
    const double *sri_a, *sri_b, *sri_c, *sri_d, *sri_e;
    const int *thd_map, *thd_type, *thd_off, *thd_spec;
    const double *thd_eff;
    const int *plog_map, *plog_np;
    const double *plog_params;
    const int *cheb_map, *cheb_nt, *cheb_np;
    const double *cheb_params, *cheb_plim, *cheb_tlim;
    const double *a_lo, *a_hi, *T_mid, *mw_post;
    const int *rxn_has_ns;
    // derived on the host by the runtime
    const int *thd_only;      // [n_thd_only] third-body-only reactions' thd index
    int n_thd_only;
    const int *fall_order;    // [n_fall] falloff indices in Lindemann, SRI, Troe order
    const int *jac_pos;       // [jac_nnz] dense index (row*NN+col) of each stored value
};
