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

}  // namespace pjk_host
