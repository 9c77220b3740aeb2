// Statically compiled gfx950 kernels: the ray-cast kernel for every predefined field, the
// Lyapunov overlay and a counter-clearing kernel (a fill kernel rather than hipMemsetAsync,
// so that a caller may capture linon_render into a HIP graph).
#include <hip/hip_runtime.h>

#include <cstring>

#include "device/batch.hpp"
#include "device/fields.hpp"
#include "device/raycast.hpp"
#include "kernels.hpp"

namespace LINON_KERNEL_NS {

template <class F>
__global__ void __launch_bounds__(64) LK_RAYCAST_OCCUPANCY raycast_kernel(RaycastParams P) {
    lk::raycast_body<F>(P);
}

template <class F>
__global__ void __launch_bounds__(256) LK_BATCH_OCCUPANCY batch_kernel(BatchParams B) {
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
        case LINON_KERNEL_HOT_AIR_FAITHFUL_ROOT:                       \
            return (const void*)&KERNEL<lw::FieldHotAirFaithfulRoot>;  \
        case LINON_KERNEL_LENS_RECIP:                                  \
            return (const void*)&KERNEL<lw::FieldLensRecip>;           \
        default: return nullptr;                                       \
    }

const void* batch_kernel_for(int id) { LINON_FIELD_TABLE(batch_kernel) }

hipError_t launch_batch(int field, const BatchParams& B, unsigned blocks, unsigned threads, hipStream_t s) {
    const void* fn = batch_kernel_for(field);
    if (!fn) return hipErrorInvalidDeviceFunction;
    void* args[] = {const_cast<BatchParams*>(&B)};
    return hipLaunchKernel(fn, dim3(blocks), dim3(threads), args, threads * sizeof(lk::BatchLaneTally), s);
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
        case LINON_KERNEL_HOT_AIR_FAITHFUL_ROOT: return (const void*)&raycast_kernel<lw::FieldHotAirFaithfulRoot>;
        case LINON_KERNEL_LENS_RECIP: return (const void*)&raycast_kernel<lw::FieldLensRecip>;
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

// ---- IEEE build only: self-tests of its arithmetic (the reciprocal-based quotients against
// IEEE division, sqrt_rn against the compiler's square root, exp_wgsl against WGSL's bound
// for exp) and the FMA-rate measurement ----
#ifndef LINON_FAST_MATH
__device__ __forceinline__ unsigned fbits(float x) { return __float_as_uint(x); }

// every mantissa of b at exponent e (both signs): recip_rn(b) against 1.0f / b
__global__ void __launch_bounds__(256) div_selftest_recip_kernel(int e, unsigned long long* mism) {
    const unsigned m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (1u << 23)) return;
    unsigned bad = 0;
    for (int sgn = 0; sgn < 2; ++sgn) {
        const float b = __uint_as_float(((unsigned)sgn << 31) | ((unsigned)(e + 127) << 23) | m);
        if (fbits(lw::recip_rn(b)) != fbits(1.0f / b)) ++bad;
    }
    if (bad) atomicAdd(mism, (unsigned long long)bad);
}

// every float x whose bit pattern has the top byte `hi`: sigmoid_floor(x) equals sigmoid(x)
// where that is >= 2^-125 (2^-125 below; NaN for NaN), and hot air's 20 + 80 s (fused and
// unfused) is the same for both
__device__ __forceinline__ bool same_or_nan(float a, float b) { return fbits(a) == fbits(b) || (a != a && b != b); }
__device__ __forceinline__ float temp_unfused(float s) {
#pragma clang fp contract(off)
    return 20.0f + 80.0f * s;
}
__global__ void __launch_bounds__(256) div_selftest_sigmoid_kernel(unsigned hi, unsigned long long* mism) {
    const unsigned lo = blockIdx.x * blockDim.x + threadIdx.x;
    if (lo >= (1u << 24)) return;
    const float x = __uint_as_float((hi << 24) | lo);
    const float ref = lw::sigmoid(x), s = lw::sigmoid_floor(x);
    unsigned bad = 0;
    bad += !same_or_nan(s, ref >= 0x1p-125f || ref != ref ? ref : 0x1p-125f);
    bad += !same_or_nan(__builtin_fmaf(80.0f, s, 20.0f), __builtin_fmaf(80.0f, ref, 20.0f));
    bad += !same_or_nan(temp_unfused(s), temp_unfused(ref));
    if (bad) atomicAdd(mism, (unsigned long long)bad);
}

// every mantissa of a at exponent ea against one b: quot through div3_rn against a / b
__global__ void __launch_bounds__(256) div_selftest_mantissa_kernel(int ea, float b, unsigned long long* mism) {
    const unsigned m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (1u << 23)) return;
    const float a = __uint_as_float(((unsigned)(ea + 127) << 23) | m);
    const lw::vec3 q = lw::div3_rn(lw::vec3f(a, -a, 0.75f * a), b);
    unsigned bad = 0;
    bad += fbits(q.x) != fbits(a / b);
    bad += fbits(q.y) != fbits(-a / b);
    bad += fbits(q.z) != fbits((0.75f * a) / b);
    if (bad) atomicAdd(mism, (unsigned long long)bad);
}

// The bounded divisions (div3_rn_bounded: normalize; div3_rn_lens: the lens field) with every
// numerator mantissa at exponent ea (ea = -127: the subnormals), negated, and beside a zero of
// each sign and 0.5, against one denominator s, component-wise against IEEE division; the
// numerators stay below 2^61 as their callers guarantee
template <int Which>
__device__ __forceinline__ lw::vec3 bounded_div(lw::vec3 v, float s) {
    return Which == 0 ? lw::div3_rn_bounded(v, s) : lw::div3_rn_lens(v, s);
}
template <int Which>
__device__ __forceinline__ unsigned bounded_mismatches(float a, float s) {
    unsigned bad = 0;
    const lw::vec3 q1 = bounded_div<Which>(lw::vec3f(a, -a, 0.0f), s);
    bad += fbits(q1.x) != fbits(a / s);
    bad += fbits(q1.y) != fbits(-a / s);
    bad += fbits(q1.z) != fbits(0.0f / s);
    const lw::vec3 q2 = bounded_div<Which>(lw::vec3f(-0.0f, 0.5f, a), s);
    bad += fbits(q2.x) != fbits(-0.0f / s);
    bad += fbits(q2.y) != fbits(0.5f / s);
    bad += fbits(q2.z) != fbits(a / s);
    return bad;
}
__global__ void __launch_bounds__(256) div_selftest_bounded_kernel(int ea, float s, unsigned long long* mism) {
    const unsigned m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (1u << 23)) return;
    const float a = __uint_as_float(((unsigned)(ea + 127) << 23) | m);
    const unsigned bad = bounded_mismatches<0>(a, s) + bounded_mismatches<1>(a, s);
    if (bad) atomicAdd(mism, (unsigned long long)bad);
}

// the lens quotient at its floor s = 0.001 for every numerator of magnitude below 2^13 (both
// signs, subnormals included): the captured rays' components
__global__ void __launch_bounds__(256) div_selftest_lens_floor_kernel(int biased_exp, unsigned long long* mism) {
    const unsigned m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (1u << 23)) return;
    const float a = __uint_as_float(((unsigned)biased_exp << 23) | m);
    const float s = 0.001f;
    const lw::vec3 q = lw::div3_rn_lens(lw::vec3f(a, -a, 0.0f), s);
    unsigned bad = 0;
    bad += fbits(q.x) != fbits(a / s);
    bad += fbits(q.y) != fbits(-a / s);
    bad += fbits(q.z) != fbits(0.0f / s);
    bad += fbits(lw::quot_lens_floor(a)) != fbits(a / s);
    if (bad) atomicAdd(mism, (unsigned long long)bad);
}

// random operands (SplitMix64 stream): sign, exponent in [-70, 70] (so the out-of-range and
// zero paths are taken too), mantissa uniform
__device__ __forceinline__ unsigned long long smix(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ float rand_float(unsigned long long r) {
    const unsigned m = (unsigned)(r & 0x7fffffu);
    const int ex = (int)((r >> 23) % 141u) - 70;
    const unsigned sg = (unsigned)((r >> 40) & 1u);
    if (((r >> 41) & 1023u) == 0) return sg ? -0.0f : 0.0f;
    return __uint_as_float((sg << 31) | ((unsigned)(ex + 127) << 23) | m);
}
__global__ void __launch_bounds__(256) div_selftest_random_kernel(unsigned long long seed, unsigned long long n,
                                                                  unsigned long long* mism) {
    const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long r1 = smix(seed + 0x9E3779B97F4A7C15ull * (2 * i + 1));
    const unsigned long long r2 = smix(seed + 0x9E3779B97F4A7C15ull * (2 * i + 2));
    const float a = rand_float(r1), b = rand_float(r2);
    const float c = rand_float(r1 ^ r2);
    const lw::vec3 q = lw::div3_rn(lw::vec3f(a, c, -a), b);
    unsigned bad = 0;
    bad += fbits(q.x) != fbits(a / b);
    bad += fbits(q.y) != fbits(c / b);
    bad += fbits(q.z) != fbits(-a / b);
    bad += fbits(lw::div_rn(a, b)) != fbits(a / b);
    // normalize with its shortened range test against the component-wise IEEE quotients
    const lw::vec3 v = lw::vec3f(a, b, c);
    const float l = lw::length(v);
    const lw::vec3 u = lw::normalize(v);
    bad += !same_or_nan(u.x, a / l);
    bad += !same_or_nan(u.y, b / l);
    bad += !same_or_nan(u.z, c / l);
    // the lens field's quotient 0.3 v / max(|v|^3, 0.001)
    const lw::vec3 w = 0.3f * v;
    const float m = __builtin_fmaxf(l * l * l, 0.001f);
    const lw::vec3 ql = lw::div3_rn_bounded(w, m);
    bad += !same_or_nan(ql.x, w.x / m);
    bad += !same_or_nan(ql.y, w.y / m);
    bad += !same_or_nan(ql.z, w.z / m);
    if (bad) atomicAdd(mism, (unsigned long long)bad);
}

// every float y in [-87, 87] whose bit pattern has the top byte `hi`: lw::exp_wgsl(y) against
// the FP64 exponential, in units of the float spacing at the result (ulp = 2^(exponent - 23),
// subnormal spacing below 2^-126); counts results off by more than WGSL's bound for exp,
// 3 + 2 |y| ulp, and keeps the largest error (in 1/1024 ulp, as an integer for atomicMax)
__global__ void __launch_bounds__(256) exp_selftest_kernel(unsigned hi, unsigned long long* over,
                                                           unsigned* max_milli_ulp) {
    const unsigned lo = blockIdx.x * blockDim.x + threadIdx.x;
    if (lo >= (1u << 24)) return;
    const float y = __uint_as_float((hi << 24) | lo);
    if (!(__builtin_fabsf(y) <= 87.0f)) return;
    const double ref = exp((double)y);
    const float got = lw::exp_wgsl(y);
    int e;
    (void)frexp(ref, &e);  // ref = m 2^e, m in [0.5, 1)
    const double ulp = ldexp(1.0, (e - 1 < -126 ? -126 : e - 1) - 23);
    const double err = fabs((double)got - ref) / ulp;
    if (err > 3.0 + 2.0 * fabs((double)y)) atomicAdd(over, 1ull);
    atomicMax(max_milli_ulp, (unsigned)fmin(err * 1024.0, 4.0e9));
}

// every float s whose bit pattern has the top byte `hi`: the RK4 average's division by six
// (lk::div6) against s / 6, except its documented exceptions: the ties of subnormal quotients
// (s = (6 m + 3) 2^-149 with m odd, below 1.5 2^-124), where div6 returns the other
// neighbour, and s = +-inf, where its residual is inf - inf and the result NaN
__global__ void __launch_bounds__(256) div6_selftest_kernel(unsigned hi, unsigned long long* mism) {
    const unsigned lo = blockIdx.x * blockDim.x + threadIdx.x;
    if (lo >= (1u << 24)) return;
    const float s = __uint_as_float((hi << 24) | lo);
    const float got = lk::div6(s), want = s / 6.0f;
    if (same_or_nan(got, want)) return;
    const float a = __builtin_fabsf(s);
    bool tie = a == __builtin_inff();
    if (a < 0x1.8p-124f) {
        const unsigned long long j = (unsigned long long)((double)a * 0x1p149);  // exact
        tie = j % 6 == 3 && ((j - 3) / 6) % 2 == 1;
    }
    if (!tie) atomicAdd(mism, 1ull);
}

// every float x whose bit pattern has the top byte `hi`: sqrt_rn(x), and sqrt_rn_no_tiny(x)
// where |x| is not in (0, 2^-96), against the compiler's correctly rounded square root
__global__ void __launch_bounds__(256) sqrt_selftest_kernel(unsigned hi, unsigned long long* mism) {
    const unsigned lo = blockIdx.x * blockDim.x + threadIdx.x;
    if (lo >= (1u << 24)) return;
    const float x = __uint_as_float((hi << 24) | lo);
    unsigned bad = !same_or_nan(lw::sqrt_rn(x), __builtin_sqrtf(x));
    // the form without rescaling, on its domain: every float but those of magnitude in (0, 2^-96)
    const float ax = __builtin_fabsf(x);
    if (!(ax > 0.0f && ax < 0x1p-96f)) bad += !same_or_nan(lw::sqrt_rn_no_tiny(x), __builtin_sqrtf(x));
    if (bad) atomicAdd(mism, (unsigned long long)bad);
}

// refraction_in_range against refraction: temperatures uniform in [-250, 1e4] (one in 64 at an
// end of the range, one in 1024 NaN), directions and normals with components in [-1.5, 1.5]
__device__ __forceinline__ float rand_unit(unsigned long long r) { return (float)(r >> 40) * 0x1p-24f; }
__device__ __forceinline__ float rand_temp(unsigned long long r) {
    const unsigned sel = (unsigned)(r & 1023u);
    if (sel == 0) return __uint_as_float(0x7fc00000u);
    if (sel < 16) return (sel & 1) ? -250.0f : 1e4f;
    return -250.0f + 10250.0f * rand_unit(r);
}
template <class F>
__device__ __forceinline__ unsigned first_stage_mismatches(const F& f, lw::vec3 p, lw::vec3 v) {
    const lw::vec3 a = f.first_stage(p, v, v, 0.0f), b = f(p, p, v, v, 0.0f);
    return !same_or_nan(a.x, b.x) + !same_or_nan(a.y, b.y) + !same_or_nan(a.z, b.z);
}
__global__ void __launch_bounds__(256) div_selftest_refraction_kernel(unsigned long long seed, unsigned long long n,
                                                                      unsigned long long* mism) {
    const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long r[8];
    for (int k = 0; k < 8; ++k) r[k] = smix(seed + 0x9E3779B97F4A7C15ull * (8 * i + k + 1));
    const float t_in = rand_temp(r[0]), t_out = rand_temp(r[1]);
    const lw::vec3 v = lw::vec3f(3.0f * rand_unit(r[2]) - 1.5f, 3.0f * rand_unit(r[3]) - 1.5f, 3.0f * rand_unit(r[4]) - 1.5f);
    const lw::vec3 nrm = lw::vec3f(3.0f * rand_unit(r[5]) - 1.5f, 3.0f * rand_unit(r[6]) - 1.5f, 3.0f * rand_unit(r[7]) - 1.5f);
    const lw::vec3 a = lw::refraction(t_in, t_out, v, nrm), b = lw::refraction_in_range(t_in, t_out, v, nrm);
    unsigned bad = !same_or_nan(a.x, b.x) + !same_or_nan(a.y, b.y) + !same_or_nan(a.z, b.z);
    // one temperature on both sides: refraction_in_range_same against the general form
    const lw::vec3 c = lw::refraction_in_range_same(t_in, v, nrm), d = lw::refraction_in_range(t_in, t_in, v, nrm);
    bad += !same_or_nan(c.x, d.x) + !same_or_nan(c.y, d.y) + !same_or_nan(c.z, d.z);
    // the refracting fields' first RK4 stage, f.first_stage(p, v0, v, t) = f(p, p, v0, v, t),
    // with p taken from the random normals (components in [-1.5, 1.5])
    bad += first_stage_mismatches(lw::FieldHotAir{}, nrm, v);
    bad += first_stage_mismatches(lw::FieldHotAirFaithfulRoot{}, nrm, v);
    bad += first_stage_mismatches(lw::FieldMirageSpherical{}, nrm, v);
    bad += first_stage_mismatches(lw::FieldMiragePlane{}, nrm, v);
    bad += first_stage_mismatches(lw::FieldMirageSphericalSigmoid{}, nrm, v);
    bad += first_stage_mismatches(lw::FieldMiragePlaneSigmoid{}, nrm, v);
    // refraction_in_range against the reference's branches restated here (src/main_view.wgsl:
    // 119-135): the side of the normal, and reflection where k < 0 (frequent on these draws)
    {
        using namespace lw;
        const float ei = div_in_range(refraction_index_num(t_in), refraction_index_den(t_in));
        const float eo = div_in_range(refraction_index_num(t_out), refraction_index_den(t_out));
        float cosi = clamp(-1.0f, 1.0f, dot(v, nrm));
        vec3 n_ref = nrm;
        if (cosi < 0.0f) cosi = -cosi;
        else n_ref = -nrm;
        const float eta = div_in_range(ei, eo);
        const float k = __builtin_fmaf(-(eta * eta), __builtin_fmaf(-cosi, cosi, 1.0f), 1.0f);
        const float X = __builtin_fmaf(eta, cosi, -sqrt_rn(k));
        const vec3 refr = {__builtin_fmaf(eta, v.x, mul_rounded(X, n_ref.x)),
                           __builtin_fmaf(eta, v.y, mul_rounded(X, n_ref.y)),
                           __builtin_fmaf(eta, v.z, mul_rounded(X, n_ref.z))};
        const vec3 want = k < 0.0f ? reflect(v, n_ref) : refr;
        bad += !same_or_nan(b.x, want.x) + !same_or_nan(b.y, want.y) + !same_or_nan(b.z, want.z);
    }
    // refraction's k = RN(1 - e t) where e t is near 1 (e in [0.5, 2), t within 8 ulps of 1 / e,
    // at most 1): never in (0, 2^-48), and its root without rescaling equals sqrt_rn's
    const float e2 = 0.5f + 1.5f * rand_unit(r[2] ^ r[5]);
    const int off = (int)(r[6] >> 60) - 8;
    const float tt = __builtin_fminf(__uint_as_float(__float_as_uint(1.0f / e2) + off), 1.0f);
    const float kk = __builtin_fmaf(-e2, tt, 1.0f);
    bad += kk != 0.0f && __builtin_fabsf(kk) < 0x1p-48f;
    if (kk >= 0.0f) bad += !same_or_nan(lw::sqrt_rn_no_tiny(kk), lw::sqrt_rn(kk));
    if (bad) atomicAdd(mism, (unsigned long long)bad);
}
#endif

#ifndef LINON_FAST_MATH
// Rate of non-packed FP32 FMA issue: every lane runs 16 independent chains of v_fma_f32 (the
// build has packed FP32 turned off), so the SIMDs are never short of independent work.
__global__ void __launch_bounds__(256) fma_rate_kernel(float* out, int iters, float b, float c) {
    float a[16];
    for (int k = 0; k < 16; ++k) a[k] = (float)(threadIdx.x + k) * 1e-3f;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 16; ++k) a[k] = __builtin_fmaf(a[k], b, c);
    }
    float sum = 0.0f;
    for (int k = 0; k < 16; ++k) sum += a[k];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = sum;
}

hipError_t run_fma_rate(int cus, double* tflops) {
    const unsigned blocks = (unsigned)(cus > 0 ? cus : 1) * 8u, threads = 256u;
    const int iters = 2048;
    float* d = nullptr;
    hipError_t e = hipMalloc(&d, (size_t)blocks * threads * sizeof(float));
    if (e != hipSuccess) return e;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float ms = 0.0f;
    for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {  // the second run is timed
        e = hipEventRecord(e0, nullptr);
        if (e == hipSuccess)
            hipLaunchKernelGGL(fma_rate_kernel, dim3(blocks), dim3(threads), 0, nullptr, d, iters, 0.999f, 0.001f);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(d);
    if (e == hipSuccess) *tflops = (double)blocks * threads * iters * 16.0 * 2.0 / (ms * 1e-3) / 1e12;
    return e;
}

hipError_t run_sqrt_selftest(unsigned long long* mism, unsigned long long* cases) {
    unsigned long long* d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof(unsigned long long));
    if (e != hipSuccess) return e;
    e = hipMemset(d, 0, sizeof(unsigned long long));
    for (unsigned hi = 0; hi < 256 && e == hipSuccess; ++hi)
        hipLaunchKernelGGL(sqrt_selftest_kernel, dim3((1u << 24) / 256), dim3(256), 0, nullptr, hi, d);
    if (e == hipSuccess) e = hipGetLastError();
    unsigned long long h = 0;
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    *mism = h;
    *cases = 1ull << 32;
    return e;
}

hipError_t run_exp_selftest(unsigned long long* over, unsigned long long* cases, float* max_ulp) {
    unsigned long long* d = nullptr;
    hipError_t e = hipMalloc(&d, 2 * sizeof(unsigned long long));
    if (e != hipSuccess) return e;
    e = hipMemset(d, 0, 2 * sizeof(unsigned long long));
    unsigned* dm = reinterpret_cast<unsigned*>(d + 1);
    unsigned long long n = 0;
    for (unsigned hi = 0; hi < 256 && e == hipSuccess; ++hi) {
        hipLaunchKernelGGL(exp_selftest_kernel, dim3((1u << 24) / 256), dim3(256), 0, nullptr, hi, d, dm);
        n += 1ull << 24;
    }
    if (e == hipSuccess) e = hipGetLastError();
    unsigned long long h[2] = {0, 0};
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    *over = h[0];
    *cases = n;
    unsigned mm = 0;
    std::memcpy(&mm, &h[1], sizeof(mm));
    *max_ulp = (float)mm / 1024.0f;
    return e;
}

hipError_t run_division_selftest(unsigned long long random_pairs, unsigned long long* recip_mism,
                                 unsigned long long* quot_mism, unsigned long long* cases) {
    unsigned long long* d = nullptr;
    hipError_t e = hipMalloc(&d, 2 * sizeof(unsigned long long));
    if (e != hipSuccess) return e;
    e = hipMemset(d, 0, 2 * sizeof(unsigned long long));
    unsigned long long n = 0;
    const unsigned mblocks = (1u << 23) / 256;
    // every mantissa, both signs, at every exponent whose reciprocal is normal
    for (int ex = -126; ex <= 125 && e == hipSuccess; ++ex) {
        hipLaunchKernelGGL(div_selftest_recip_kernel, dim3(mblocks), dim3(256), 0, nullptr, ex, d);
        n += 2ull << 23;
    }
    // every numerator mantissa at three exponents against structured denominators: 1, 1 + ulp,
    // 2 - ulp (all-ones mantissa), thirds, tenths, and the field constants' neighbourhood
    const float denoms[] = {1.0f, 1.00000012f, 1.99999988f, 3.0f, 0.1f, 0.001f, 7.77f, 1.41421354f,
                            0.333333343f, 1e5f, 2.5e-7f, 1.5f, 1.0000024f, 1.9999976f, 12345.678f, 0.0517f};
    for (float b : denoms)
        for (int ea : {-20, 0, 21}) {
            if (e != hipSuccess) break;
            hipLaunchKernelGGL(div_selftest_mantissa_kernel, dim3(mblocks), dim3(256), 0, nullptr, ea, b, d + 1);
            n += 3ull << 23;
        }
    // the bounded division (normalize, the lens field) for zero, tiny and subnormal numerators:
    // every mantissa at exponents from the subnormals up through its fast range's lower end,
    // against denominators across [2^-61, 2^60], the lens field's floor 0.001 and beyond
    const float bdenoms[] = {0.001f, 0.3f, 1.0f, 1.99999988f, 7.77f, 1000.0f, 0x1p-61f, 0x1.8p-61f, 0x1p-40f,
                             0x1p60f, 0x1.fffffep59f, 0x1p61f, 0x1p-62f, 3.0e-20f, 0x1p22f, 0x1.fffffep21f,
                             0x1.000002p22f, 0x1.8p21f, 0x1.fffffep-1f};
    for (float b : bdenoms)
        for (int ea : {-127, -126, -125, -110, -104, -103, -102, -101, -100, -90, -80, -70, -62, -61, -60, -59, -30,
                       0}) {
            if (e != hipSuccess) break;
            hipLaunchKernelGGL(div_selftest_bounded_kernel, dim3(mblocks), dim3(256), 0, nullptr, ea, b, d + 1);
            n += 12ull << 23;
        }
    for (unsigned hi = 0; hi < 256 && e == hipSuccess; ++hi) {  // div6 at every float
        hipLaunchKernelGGL(div6_selftest_kernel, dim3((1u << 24) / 256), dim3(256), 0, nullptr, hi, d + 1);
        n += 1ull << 24;
    }
    for (int be = 0; be < 127 + 13 && e == hipSuccess; ++be) {  // lens floor: |a| < 2^13
        hipLaunchKernelGGL(div_selftest_lens_floor_kernel, dim3(mblocks), dim3(256), 0, nullptr, be, d + 1);
        n += 4ull << 23;
    }
    for (unsigned hi = 0; hi < 256 && e == hipSuccess; ++hi) {  // sigmoid_floor at every float
        hipLaunchKernelGGL(div_selftest_sigmoid_kernel, dim3((1u << 24) / 256), dim3(256), 0, nullptr, hi, d + 1);
        n += 3ull << 24;
    }
    if (random_pairs) {
        const unsigned long long chunk = 1ull << 26;
        for (unsigned long long off = 0; off < random_pairs && e == hipSuccess; off += chunk) {
            const unsigned long long m = random_pairs - off < chunk ? random_pairs - off : chunk;
            hipLaunchKernelGGL(div_selftest_refraction_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0,
                               nullptr, 0x5851F42D4C957F2Dull + off, m, d + 1);
            n += 29 * m;
        }
        for (unsigned long long off = 0; off < random_pairs && e == hipSuccess; off += chunk) {
            const unsigned long long m = random_pairs - off < chunk ? random_pairs - off : chunk;
            hipLaunchKernelGGL(div_selftest_random_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, nullptr,
                               0xD1B54A32D192ED03ull + off, m, d + 1);
            n += 10 * m;
        }
    }
    if (e == hipSuccess) e = hipGetLastError();
    unsigned long long h[2] = {0, 0};
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    *recip_mism = h[0];
    *quot_mism = h[1];
    *cases = n;
    return e;
}
#endif
}  // namespace LINON_KERNEL_NS
