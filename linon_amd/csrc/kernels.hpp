// Host-side launch helpers for the statically compiled kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "device/params.hpp"

namespace linon {

bool field_is_predefined(int id);
const void* raycast_kernel_for(int id);
hipError_t launch_raycast_static(int field, const RaycastParams& P, unsigned blocks, size_t shmem,
                                 hipStream_t s);
hipError_t launch_overlay(const OverlayParams& P, hipStream_t s);
const void* batch_kernel_for(int id);
hipError_t launch_batch(int field, const BatchParams& B, unsigned blocks, unsigned threads, hipStream_t s);
hipError_t launch_clear_u64(unsigned long long* p, int n, hipStream_t s);
hipError_t launch_outline_search(const float* exps, int W, int H, unsigned long long* keys, hipStream_t s);

}  // namespace linon
