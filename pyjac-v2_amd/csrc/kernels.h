// kernels.h -- host-side launchers of the gfx950 kernels (pyjac_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "device_mech.h"

namespace pjk_host {

// doubles of per-state workspace the rates (jac=false) or Jacobian
// (jac=true) kernel needs; the launchers take a workspace of
// grid * block * this many doubles
size_t ws_doubles_per_state(const DevMech& m, bool jac);

hipError_t launch_species_rates(const DevMech& m, int conp, char order, long n,
                                const double* param, const double* phi, double* dphi,
                                double* ws, int grid, int block, hipStream_t stream);

hipError_t launch_jacobian(const DevMech& m, int conp, char order, long n, const double* param,
                           const double* phi, double* jac, double* ws, int grid, int block,
                           hipStream_t stream);

hipError_t launch_chem_utils(const DevMech& m, int conp, char order, long n, const double* phi,
                             double* energy, double* heat, double* b, hipStream_t stream);

}  // namespace pjk_host
