// linon_flowmap_noalloc: the flow map of linon_flowmap for callers that cannot allocate device
// memory (tight serving loops, harnesses that forbid hipMalloc).  linon_flowmap stages the
// caller's host states, the exits and the iteration counts in device buffers it allocates and
// frees on every call; this variant stages them in one file-scope device array of fixed size
// and runs the same flow-map kernel on it.
//
// The array is one per device for the whole process (each device holds its own copy from the
// first call on it until the process ends).  Every call on a device reuses that device's copy:
// calls that overlap on one device (from several host threads) race and must be serialised by
// the caller.  The iteration counts are accumulated by the kernel, so they are cleared (a fill
// kernel) on every call; states and exits are written in full before they are read.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "../../include/linon.h"
#include "context.hpp"
#include "surrogate_kernels.hpp"

// states [n][6] | exits [n][6] | iteration counts [n] (int32), n <= LINON_FLOWMAP_NOALLOC_MAX
#define LINON_NOALLOC_FLOATS (13 * LINON_FLOWMAP_NOALLOC_MAX)
static __device__ __align__(256) float scratch[LINON_NOALLOC_FLOATS];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
float* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device, once looked up

// The device address of this device's copy of `scratch`, looked up on the first call on that
// device (with the device current).  In host code the array's own name is a host-side stand-in,
// not a device address, so only this pointer is ever handed to kernels or copies.
hipError_t scratch_on(int device, float** out) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = static_cast<float*>(p);
    }
    *out = g_scratch[device];
    return hipSuccess;
}

linon_status fail(linon_status st, const std::string& msg) {
    linon::set_unbound_error(msg);
    return st;
}

}  // namespace

extern "C" linon_status linon_flowmap_noalloc(int device, int field, int math, const linon_integrator* integrator,
                                              float field_weight, const float* states, int n, float* exits,
                                              int32_t* steps) {
    linon::set_unbound_error("");
    if (!integrator || !states || !exits)
        return fail(LINON_ERR_ARGUMENT, "flowmap_noalloc: null integrator, states or exits");
    if (n < 1 || n > LINON_FLOWMAP_NOALLOC_MAX)
        return fail(LINON_ERR_ARGUMENT, "flowmap_noalloc: n = " + std::to_string(n) + " is outside [1, " +
                                            std::to_string(LINON_FLOWMAP_NOALLOC_MAX) +
                                            "], the capacity of its fixed device staging");
    if (integrator->mode != LINON_MODE_ESCAPE || math < LINON_MATH_AUTO || math > LINON_MATH_WGSL ||
        !(integrator->h_initial > 0.0f) || integrator->max_steps < 1 ||
        (integrator->clip_exit != 0 && integrator->clip_exit != 1))
        return fail(LINON_ERR_ARGUMENT, "flowmap_noalloc: needs an escape-mode integrator, h > 0, max_steps >= 1 "
                                        "and a known math mode");
    if (!linon::sg_field_supported(field))
        return fail(LINON_ERR_UNSUPPORTED, "flowmap_noalloc: field " + std::to_string(field) + " has no flow map");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count || device >= kMaxDevices)
        return fail(LINON_ERR_ARGUMENT, "flowmap_noalloc: no device " + std::to_string(device));
    hipError_t e = hipSetDevice(device);
    float* base = nullptr;
    if (e == hipSuccess) e = scratch_on(device, &base);
    if (e != hipSuccess) return fail(LINON_ERR_HIP, std::string("flowmap_noalloc: ") + hipGetErrorString(e));
    float* d_in = base;
    float* d_out = base + 6 * (size_t)LINON_FLOWMAP_NOALLOC_MAX;
    int* d_steps = reinterpret_cast<int*>(base + 12 * (size_t)LINON_FLOWMAP_NOALLOC_MAX);
    const size_t bytes = (size_t)n * 6 * sizeof(float);
    ls::FlowmapParams F;
    std::memset(&F, 0, sizeof(F));
    linon::fill_step(F.step, *integrator, field_weight);
    F.states = d_in;
    F.n = n;
    F.exits = d_out;
    F.steps = d_steps;
    // all on the null stream, in order: stage in, clear the counts, integrate, copy out
    e = hipMemcpy(d_in, states, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = linon::sg_fill_i32(d_steps, n, 0, nullptr);
    if (e == hipSuccess)
        e = linon::wgsl_math_default(math, field)
                ? linon_wgslmath::sg_flowmap(linon::kernel_field_id(math, field), F, nullptr)
                : linon::sg_flowmap(linon::kernel_field_id(math, field), F, nullptr);
    if (e == hipSuccess) e = hipMemcpy(exits, d_out, bytes, hipMemcpyDeviceToHost);
    if (e == hipSuccess && steps) e = hipMemcpy(steps, d_steps, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(LINON_ERR_HIP, std::string("flowmap_noalloc: ") + hipGetErrorString(e));
    return LINON_OK;
}
