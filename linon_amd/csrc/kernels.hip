// Statically compiled gfx950 kernels: the ray-cast kernel for every predefined field, the
// Lyapunov overlay and a counter-clearing kernel (a fill kernel rather than hipMemsetAsync,
// so that a caller may capture linon_render into a HIP graph).
#include <hip/hip_runtime.h>

#include "device/batch.hpp"
#include "device/fields.hpp"
#include "device/raycast.hpp"
#include "kernels.hpp"

namespace LINON_KERNEL_NS {

template <class F>
__global__ void __launch_bounds__(64) raycast_kernel(RaycastParams P) {
    lk::raycast_body<F>(P);
}

template <class F>
__global__ void __launch_bounds__(256) batch_kernel(BatchParams B) {
    lk::batch_body<F>(B);
}

__global__ void __launch_bounds__(256) overlay_kernel(OverlayParams P) { lk::overlay_body(P); }

__global__ void __launch_bounds__(256) outline_search_kernel(const float* exps, int W, int H,
                                                             unsigned long long* keys) {
    lk::outline_search_body(exps, W, H, keys);
}

__global__ void __launch_bounds__(64) clear_u64_kernel(unsigned long long* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

bool field_is_predefined(int id) { return (id >= 1 && id <= 9) || (id >= 16 && id <= 20); }

#define LINON_FIELD_TABLE(KERNEL)                                      \
    switch (id) {                                                      \
        case 1: return (const void*)&KERNEL<lw::FieldMirageSpherical>; \
        case 2: return (const void*)&KERNEL<lw::FieldMiragePlane>;     \
        case 3: return (const void*)&KERNEL<lw::FieldMirageSphericalSigmoid>; \
        case 4: return (const void*)&KERNEL<lw::FieldMiragePlaneSigmoid>; \
        case 5: return (const void*)&KERNEL<lw::FieldTranslationX>;    \
        case 6: return (const void*)&KERNEL<lw::FieldTranslationZ>;    \
        case 7: return (const void*)&KERNEL<lw::FieldRotation>;        \
        case 8: return (const void*)&KERNEL<lw::FieldLorenz>;          \
        case 9: return (const void*)&KERNEL<lw::FieldRoessler>;        \
        case 16: return (const void*)&KERNEL<lw::FieldIdentity>;       \
        case 17: return (const void*)&KERNEL<lw::FieldSwirl>;          \
        case 18: return (const void*)&KERNEL<lw::FieldLens>;           \
        case 19: return (const void*)&KERNEL<lw::FieldHotAir>;         \
        case 20: return (const void*)&KERNEL<lw::FieldWarp>;           \
        default: return nullptr;                                       \
    }

const void* batch_kernel_for(int id) { LINON_FIELD_TABLE(batch_kernel) }

hipError_t launch_batch(int field, const BatchParams& B, unsigned blocks, unsigned threads, hipStream_t s) {
    const void* fn = batch_kernel_for(field);
    if (!fn) return hipErrorInvalidDeviceFunction;
    void* args[] = {const_cast<BatchParams*>(&B)};
    return hipLaunchKernel(fn, dim3(blocks), dim3(threads), args, 0, s);
}

const void* raycast_kernel_for(int id) {
    switch (id) {
        case 1: return (const void*)&raycast_kernel<lw::FieldMirageSpherical>;
        case 2: return (const void*)&raycast_kernel<lw::FieldMiragePlane>;
        case 3: return (const void*)&raycast_kernel<lw::FieldMirageSphericalSigmoid>;
        case 4: return (const void*)&raycast_kernel<lw::FieldMiragePlaneSigmoid>;
        case 5: return (const void*)&raycast_kernel<lw::FieldTranslationX>;
        case 6: return (const void*)&raycast_kernel<lw::FieldTranslationZ>;
        case 7: return (const void*)&raycast_kernel<lw::FieldRotation>;
        case 8: return (const void*)&raycast_kernel<lw::FieldLorenz>;
        case 9: return (const void*)&raycast_kernel<lw::FieldRoessler>;
        case 16: return (const void*)&raycast_kernel<lw::FieldIdentity>;
        case 17: return (const void*)&raycast_kernel<lw::FieldSwirl>;
        case 18: return (const void*)&raycast_kernel<lw::FieldLens>;
        case 19: return (const void*)&raycast_kernel<lw::FieldHotAir>;
        case 20: return (const void*)&raycast_kernel<lw::FieldWarp>;
        default: return nullptr;
    }
}

hipError_t launch_raycast_static(int field, const RaycastParams& P, unsigned blocks, size_t shmem,
                                 hipStream_t s) {
    const void* fn = raycast_kernel_for(field);
    if (!fn) return hipErrorInvalidDeviceFunction;
    void* args[] = {const_cast<RaycastParams*>(&P)};
    return hipLaunchKernel(fn, dim3(blocks), dim3(64), args, shmem, s);
}

hipError_t launch_overlay(const OverlayParams& P, hipStream_t s) {
    dim3 grid((P.width + 15) / 16, (P.height + 15) / 16);
    void* args[] = {const_cast<OverlayParams*>(&P)};
    return hipLaunchKernel((const void*)&overlay_kernel, grid, dim3(256), args, 0, s);
}

hipError_t launch_outline_search(const float* exps, int W, int H, unsigned long long* keys, hipStream_t s) {
    const size_t n = (size_t)W * (size_t)(W > H ? W : H);
    void* args[] = {&exps, &W, &H, &keys};
    return hipLaunchKernel((const void*)&outline_search_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), args, 0,
                           s);
}

hipError_t launch_clear_u64(unsigned long long* p, int n, hipStream_t s) {
    void* args[] = {&p, &n};
    return hipLaunchKernel((const void*)&clear_u64_kernel, dim3((n + 63) / 64), dim3(64), args, 0, s);
}

}  // namespace LINON_KERNEL_NS
