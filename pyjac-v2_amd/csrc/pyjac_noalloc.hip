// pyjac_noalloc.hip -- the fixed scratch area of the Jacobian entry point
// that allocates nothing (pyjac_hip_jacobian_dev_noalloc).
//
// The F-order Jacobian (C-order chunks of states and a transpose per chunk)
// and the variant that gathers the Jacobian in device memory (a buffer of
// per-entry contributions) need device memory beyond the caller's inputs and
// outputs.  The regular entry points allocate it on first use; callers that
// cannot allocate (code under graph capture, tight serving loops) get it
// from this file-scope array instead, NOALLOC_SCRATCH_BYTES per GPU, held
// from the library's load on that GPU until the process ends.  Every call on
// one GPU shares the array, so calls that overlap in time on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kernels.h"

namespace {

__device__ __align__(256) double scratch[pjk_host::NOALLOC_SCRATCH_BYTES / sizeof(double)];

constexpr int MAX_DEVICES = 64;
std::mutex g_mutex;
double* g_ptr[MAX_DEVICES] = {};

}  // namespace

namespace pjk_host {

// device address of the scratch array on the current device: looked up once
// per device (with that device current) and kept; the array's name is a
// host-side stand-in, never a device address
hipError_t noalloc_scratch(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_ptr[dev];
    return hipSuccess;
}

}  // namespace pjk_host
