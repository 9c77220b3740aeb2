// Host-side launch helpers for the statically compiled kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "device/params.hpp"

// The static kernels are compiled twice (linon_amd/build.py): namespace linon holds the IEEE
// build (correctly rounded division and sqrt, full-precision transcendentals), namespace
// linon_wgslmath the build with WGSL's own accuracy (approximate division and sqrt, hardware
// transcendentals).  kernels.hip defines its launchers in LINON_KERNEL_NS.
#ifndef LINON_KERNEL_NS
#define LINON_KERNEL_NS linon
#endif

#define LINON_DECLARE_KERNEL_LAUNCHERS                                                                  \
    bool field_is_predefined(int id);                                                                   \
    const void* raycast_kernel_for(int id);                                                             \
    hipError_t launch_raycast_static(int field, const RaycastParams& P, unsigned blocks, size_t shmem,  \
                                     hipStream_t s);                                                    \
    hipError_t launch_overlay(const OverlayParams& P, hipStream_t s);                                   \
    const void* batch_kernel_for(int id);                                                               \
    hipError_t launch_batch(int field, const BatchParams& B, unsigned blocks, unsigned threads,         \
                            hipStream_t s);                                                             \
    hipError_t launch_clear_u64(unsigned long long* p, int n, hipStream_t s);                           \
    hipError_t launch_outline_search(const float* exps, int W, int H, unsigned long long* keys, hipStream_t s);

namespace linon {
LINON_DECLARE_KERNEL_LAUNCHERS
// IEEE build only: lw::div3_rn / lw::div_rn against the IEEE division, every mantissa of the
// reciprocal and of numerators against structured denominators, plus random operand pairs
// Measured rate of non-packed FP32 FMA over `cus` compute units (IEEE build)
hipError_t run_fma_rate(int cus, double* tflops);
// lw::sqrt_rn at every float against the compiler's square root (IEEE build)
hipError_t run_sqrt_selftest(unsigned long long* mism, unsigned long long* cases);
hipError_t run_exp_selftest(unsigned long long* over, unsigned long long* cases, float* max_ulp);
hipError_t run_division_selftest(unsigned long long random_pairs, unsigned long long* recip_mism,
                                 unsigned long long* quot_mism, unsigned long long* cases);
}  // namespace linon
namespace linon_wgslmath {
LINON_DECLARE_KERNEL_LAUNCHERS
}  // namespace linon_wgslmath
