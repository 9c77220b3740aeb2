// Device-side WGSL vocabulary for field functions: vec3<f32> arithmetic, the WGSL built-ins
// a field body may call, and the helper functions linon's shader defines for field bodies
// (src/main_view.wgsl:76-146).  Predefined fields (fields.hpp) and bodies translated from
// WGSL at run time (compiled by hiprtc) are both written against this header.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

namespace lw {

struct vec3 {
    float x, y, z;
};

__device__ __forceinline__ vec3 vec3f(float x, float y, float z) { return vec3{x, y, z}; }
__device__ __forceinline__ vec3 vec3f(float s) { return vec3{s, s, s}; }

__device__ __forceinline__ vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ vec3 operator*(vec3 a, vec3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ vec3 operator/(vec3 a, vec3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ vec3 operator-(vec3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ vec3 operator*(float s, vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ vec3 operator*(vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
// Division of a vec3 by a scalar.  In the IEEE build the three quotients share one correctly
// rounded reciprocal y = RN(1/s) (the hardware approximation, within 1 ulp, and one Newton step
// with FMA), and each quotient takes Markstein's correction: q0 = RN(a y), r = a - s q0 (exact
// with FMA), q = RN(q0 + r y), the correctly rounded a / s when y = RN(1/s) and nothing over-
// or underflows (P. Markstein, IBM J. Res. Dev. 34(1), 1990).  Operands outside
// [2^-60, 2^60] in magnitude, zeros and non-finite values take the ordinary division, so the
// result is a / s bit for bit everywhere (checked exhaustively for the reciprocal and on random
// and structured quotients by linon_selftest_division).  The WGSL-accuracy build divides with
// the hardware approximation directly.
__device__ __forceinline__ bool div_rn_range(float x) {
    const float ax = __builtin_fabsf(x);
    return ax >= 0x1p-60f && ax <= 0x1p60f;  // false for 0, NaN, inf and denormals
}
__device__ __forceinline__ float recip_rn(float b) {
#pragma clang fp contract(off)
    const float y0 = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, y0, 1.0f);
    return __builtin_fmaf(e, y0, y0);
}
__device__ __forceinline__ float quot_rn(float a, float b, float y) {
#pragma clang fp contract(off)
    const float q0 = a * y;
    const float r = __builtin_fmaf(-b, q0, a);
    return __builtin_fmaf(r, y, q0);
}
__device__ __forceinline__ float div_rn(float a, float b) {
    if (div_rn_range(a) && div_rn_range(b)) return quot_rn(a, b, recip_rn(b));
    return a / b;
}
__device__ __forceinline__ vec3 div3_rn(vec3 v, float s) {
    if (div_rn_range(s) && div_rn_range(v.x) && div_rn_range(v.y) && div_rn_range(v.z)) {
        const float y = recip_rn(s);
        return {quot_rn(v.x, s, y), quot_rn(v.y, s, y), quot_rn(v.z, s, y)};
    }
    return {v.x / s, v.y / s, v.z / s};
}
#ifdef LINON_FAST_MATH
__device__ __forceinline__ vec3 operator/(vec3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
#else
__device__ __forceinline__ vec3 operator/(vec3 a, float s) { return div3_rn(a, s); }
#endif
// Scalar quotients whose operands are in [2^-60, 2^60] (or NaN) by construction, so that the
// correctly rounded reciprocal-and-correction needs neither a range test nor the ordinary
// division beside it (each caller states why its operands are in range).  A guarded scalar
// division costs more in tests and branches than it saves, so the general helpers keep `/`.
__device__ __forceinline__ float div_in_range(float a, float b) {
#ifdef LINON_FAST_MATH
    return a / b;
#else
    return quot_rn(a, b, recip_rn(b));
#endif
}
__device__ __forceinline__ vec3 operator/(float s, vec3 a) { return {s / a.x, s / a.y, s / a.z}; }
__device__ __forceinline__ vec3 operator+(vec3 a, float s) { return {a.x + s, a.y + s, a.z + s}; }
__device__ __forceinline__ vec3 operator+(float s, vec3 a) { return {s + a.x, s + a.y, s + a.z}; }
__device__ __forceinline__ vec3 operator-(vec3 a, float s) { return {a.x - s, a.y - s, a.z - s}; }
__device__ __forceinline__ vec3 operator-(float s, vec3 a) { return {s - a.x, s - a.y, s - a.z}; }

// Square root.  The IEEE build forms the correctly rounded root from the hardware one
// (v_sqrt_f32, within 1 ulp) and a one-ulp correction: of s - ulp, s and s + ulp it keeps the
// one that the exact (FMA) residuals x - (s -+ ulp) s select.  Inputs below 2^-96, whose
// residuals could underflow, are scaled by 2^32 first and the root by 2^-16 after.  Zeros,
// +inf and NaN pass through the correction unchanged, so unlike the compiler's expansion this
// one has no separate special-case select.  linon_selftest_sqrt compares it with the
// compiler's square root at every float.  The WGSL-accuracy build takes v_sqrt_f32 alone
// (within 1 ulp for normal inputs): the compiler's approximate sqrt adds a rescaling for
// inputs below 2^-96, five instructions per root, and WGSL bounds sqrt only as
// inverseSqrt does, for normal values.
__device__ __forceinline__ float sqrt_rn(float x) {
#ifdef LINON_FAST_MATH
    return __builtin_amdgcn_sqrtf(x);
#else
    // branch-free: a branch around the rare rescaling would end the basic block, and the step
    // loop then loses the scheduling that hides the root's latency (about 1.5x slower)
#pragma clang fp contract(off)
    const bool tiny = x < 0x1p-96f;
    const float xs = tiny ? x * 0x1p32f : x;
    const float s = __builtin_amdgcn_sqrtf(xs);
    const float dn = __uint_as_float(__float_as_uint(s) - 1u);
    const float up = __uint_as_float(__float_as_uint(s) + 1u);
    const float r_dn = __builtin_fmaf(-dn, s, xs);
    const float r_up = __builtin_fmaf(-up, s, xs);
    float t = r_dn <= 0.0f ? dn : s;
    t = r_up > 0.0f ? up : t;
    return tiny ? t * 0x1p-16f : t;
#endif
}
// sqrt_rn for callers whose operand is never of magnitude in (0, 2^-96): +-0, at least 2^-96 in
// magnitude, infinite or NaN (refraction's k, whose root is taken only for k >= 0).  Without the
// rescaling of tiny inputs; the same result as sqrt_rn there.  (A negative subnormal would give
// the hardware root's -0 instead of NaN.)
__device__ __forceinline__ float sqrt_rn_no_tiny(float x) {
#ifdef LINON_FAST_MATH
    return __builtin_amdgcn_sqrtf(x);
#else
#pragma clang fp contract(off)
    const float s = __builtin_amdgcn_sqrtf(x);
    const float dn = __uint_as_float(__float_as_uint(s) - 1u);
    const float up = __uint_as_float(__float_as_uint(s) + 1u);
    const float r_dn = __builtin_fmaf(-dn, s, x);
    const float r_up = __builtin_fmaf(-up, s, x);
    float t = r_dn <= 0.0f ? dn : s;
    return r_up > 0.0f ? up : t;
#endif
}

__device__ __forceinline__ float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ vec3 cross(vec3 a, vec3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(vec3 a) { return sqrt_rn(dot(a, a)); }
__device__ __forceinline__ float length(float a) { return __builtin_fabsf(a); }
// v / s for callers whose components are bounded through the denominator: every |v_i| is at
// most 2^61 whenever 2^-60 <= s <= 2^60 (normalize: |v_i| <= length(v) to an ulp; the lens
// field: |0.3 p_i| <= 0.3 r <= 0.3 s^(1/3) with s = max(r^3, 0.001)), and s >= 2^-61 wherever
// s is a number and the smallest |v_i| is at least 2^-60 (normalize: length(v) >= |v_i| to an
// ulp; lens: s >= 0.001).  With s in [2^-61, 2^60] and every |v_i| in [2^-60, 2^61] no
// intermediate of the quotients comes near the subnormal or overflow range, so the IEEE
// build's range test needs only s <= 2^60 and the smallest component magnitude; NaN or
// infinite components make s NaN or infinite there, which fails the test as in div3_rn.  Same
// result as v / s.
__device__ __forceinline__ vec3 div3_rn_bounded(vec3 v, float s) {
#ifdef LINON_FAST_MATH
    return v / s;
#else
    const float lo = __builtin_fminf(__builtin_fminf(__builtin_fabsf(v.x), __builtin_fabsf(v.y)), __builtin_fabsf(v.z));
    if (s <= 0x1p60f && lo >= 0x1p-60f) {
        const float y = recip_rn(s);
        return {quot_rn(v.x, s, y), quot_rn(v.y, s, y), quot_rn(v.z, s, y)};
    }
    return {v.x / s, v.y / s, v.z / s};
#endif
}
// Markstein's correction written with the residual's sign flipped, r' = RN(s q0 - a),
// q = RN(q0 - r' y): for a nonzero a, r' is -r exactly and q the same rounding as
// RN(q0 + r y) (quot_rn); for a = +-0 it keeps the zero's sign (q0 = +-0, r' = +0, q = +-0),
// which the unflipped form loses for -0.
__device__ __forceinline__ float quot_rn_signed_zero(float a, float b, float y) {
#pragma clang fp contract(off)
    const float q0 = a * y;
    const float rn = __builtin_fmaf(b, q0, -a);
    return __builtin_fmaf(-rn, y, q0);
}
// The lens field's 0.3 p / s with s = max(|p|^3, 0.001) >= 0.001 (or NaN), where every |0.3 p_i|
// is at most 2^61 while s <= 2^60.  Its captured rays sit at p = 0 or shrink towards it through
// ever smaller coordinates, which div3_rn_bounded sends to the IEEE division.  Here a component
// may be zero (signed-zero form above) or as small as 2^-102 while s <= 2^22.  The quotient
// a / s is then at least 2^-125, a normal number, so q0 = RN(a y) is within an ulp of it as in
// the normal range; a and s q0 are multiples of 2^(e_a - 23) and 2^(e_a - 47) >= 2^-149 (the
// exponents of s and q0 add up to at least e_a - 1), so the exact residual a - s q0, which the
// normal-range argument shows to need at most 24 significant bits, is also a multiple of 2^-149
// and the FMA forms it exactly even where it is subnormal; the corrected quotient is then the
// correctly rounded one, as in the normal range.  For s in (2^22, 2^60] the bound stays 2^-60
// (div3_rn_bounded's); anything else takes the ordinary division.  The test orders magnitudes
// by their bit patterns doubled minus 2, which sends both zeros to the top.  Where s is the
// floor 0.001 itself, every component takes quot_lens_floor instead (the captured rays' tiny
// ones included); a component below 2^-102 above the floor takes the ordinary division.  Same
// result as v / s for every finite p.
//
// quot_lens_floor(a) = a / c correctly rounded, c = RN(0.001) = 0x3a83126f, for |a| <= 2^12 (the
// lens's components while s == c: r^3 <= c gives |0.3 p_i| <= 0.0301): Q = RN(a 2^48 / c) by
// Markstein's correction (a 2^48 is exact and, for a != 0, at least 2^-101 and at most 2^60,
// in quot_rn_signed_zero's range with s = c), then Q 2^-48, one rounding.
//  * |a / c| >= 2^-126: Q 2^-48 is normal, so the product is exact and equals RN(a / c).
//  * otherwise a = j 2^-149 with |j| <= 8388 (a is subnormal: a / c > 999 a), and a / c =
//    X 2^-149 with X = (1000 - 4.7497e-5) j, 1000 |j| - 0.3985 <= |X| <= 1000 |j| < 2^23.  The
//    nearest point of the subnormal grid is 1000 j.  Q 2^-48 carries X to 24 bits: within
//    0.125 + 0.199 of 1000 |j| for |X| < 2^22 (|j| < 4195), and for |X| in [2^22, 2^23) (ulp
//    1/2, offset in [0.199, 0.3985]) exactly 1000 |j| or 1000 |j| - 1/2.  Rounding it to the
//    grid therefore gives 1000 j: a distance below 1/2, or a tie that goes to the even 1000 j.
//  * a = +-0 gives +-0.
// Y_C = RN(1 / c) = 0x4479ffff, the reciprocal recip_rn(c) forms.
__device__ __forceinline__ float quot_lens_floor(float a) {
#pragma clang fp contract(off)
    constexpr float C = 0.001f, Y_C = 999.99993896484375f;
    return quot_rn_signed_zero(a * 0x1p48f, C, Y_C) * 0x1p-48f;
}
// div3_rn_lens(v, 0.001) for |v_i| <= 2^12
__device__ __forceinline__ vec3 div3_rn_lens_floor(vec3 v) {
    return {quot_lens_floor(v.x), quot_lens_floor(v.y), quot_lens_floor(v.z)};
}
__device__ __forceinline__ vec3 div3_rn_lens(vec3 v, float s) {
#ifdef LINON_FAST_MATH
    return v / s;
#else
    // at the floor (r^3 <= 0.001: the lens's centre, where most of its rays are captured) the
    // constant-divisor form takes every component, tiny or not; most waves are entirely there
    if (s == 0.001f) return div3_rn_lens_floor(v);
    const unsigned ux = (__float_as_uint(v.x) << 1) - 2u;
    const unsigned uy = (__float_as_uint(v.y) << 1) - 2u;
    const unsigned uz = (__float_as_uint(v.z) << 1) - 2u;
    const unsigned m = min(min(ux, uy), uz);
    const unsigned lo102 = (0x0c800000u << 1) - 2u;  // 2^-102
    const unsigned lo60 = (0x21800000u << 1) - 2u;   // 2^-60
    if ((s <= 0x1p22f && m >= lo102) || (s <= 0x1p60f && m >= lo60)) {
        const float y = recip_rn(s);
        return {quot_rn_signed_zero(v.x, s, y), quot_rn_signed_zero(v.y, s, y), quot_rn_signed_zero(v.z, s, y)};
    }
    return {v.x / s, v.y / s, v.z / s};
#endif
}
// The lens's 0.3 p / s at WGSL's accuracy above the floor: the floor s = 0.001 keeps the exact
// constant-divisor form (the captured centre, where rounding decides whether a ray stays held);
// above it the three components multiply by the correctly rounded reciprocal of s, RN(v_i RN(1
// / s)), within about one ulp of v_i / s where WGSL bounds the division by 2.5 ulp.  s is at
// least 0.001 or NaN, so the reciprocal needs no range test.
__device__ __forceinline__ vec3 div3_lens_recip(vec3 v, float s) {
#ifdef LINON_FAST_MATH
    return v / s;
#else
    if (s == 0.001f) return div3_rn_lens_floor(v);
    const float y = recip_rn(s);
    return {v.x * y, v.y * y, v.z * y};
#endif
}
__device__ __forceinline__ vec3 normalize(vec3 a) { return div3_rn_bounded(a, length(a)); }
__device__ __forceinline__ float distance(vec3 a, vec3 b) { return length(a - b); }
__device__ __forceinline__ vec3 reflect(vec3 e1, vec3 e2) { return e1 - (2.0f * dot(e2, e1)) * e2; }

// Scalar built-ins (f32) and their component-wise vec3 forms.
#define LW_UNARY(name, expr)                                                        \
    __device__ __forceinline__ float name(float x) { return expr; }                \
    __device__ __forceinline__ vec3 name(vec3 v) { return {name(v.x), name(v.y), name(v.z)}; }
LW_UNARY(sin, ::sinf(x))
LW_UNARY(cos, ::cosf(x))
LW_UNARY(tan, ::tanf(x))
LW_UNARY(asin, ::asinf(x))
LW_UNARY(acos, ::acosf(x))
LW_UNARY(atan, ::atanf(x))
LW_UNARY(sinh, ::sinhf(x))
LW_UNARY(cosh, ::coshf(x))
LW_UNARY(tanh, ::tanhf(x))
LW_UNARY(exp, ::expf(x))
LW_UNARY(exp2, ::exp2f(x))
LW_UNARY(log, ::logf(x))
LW_UNARY(log2, ::log2f(x))
LW_UNARY(sqrt, sqrt_rn(x))
LW_UNARY(inverseSqrt, 1.0f / sqrt_rn(x))
LW_UNARY(abs, __builtin_fabsf(x))
LW_UNARY(floor, __builtin_floorf(x))
LW_UNARY(ceil, __builtin_ceilf(x))
LW_UNARY(round, __builtin_rintf(x))
LW_UNARY(trunc, __builtin_truncf(x))
LW_UNARY(fract, x - __builtin_floorf(x))
LW_UNARY(sign, (x > 0.0f) ? 1.0f : ((x < 0.0f) ? -1.0f : 0.0f))
LW_UNARY(saturate, __builtin_fminf(__builtin_fmaxf(x, 0.0f), 1.0f))
LW_UNARY(radians, x * 0.017453292519943295f)
LW_UNARY(degrees, x * 57.29577951308232f)
#undef LW_UNARY

__device__ __forceinline__ float atan2(float y, float x) { return ::atan2f(y, x); }
__device__ __forceinline__ float pow(float x, float y) { return ::powf(x, y); }
__device__ __forceinline__ vec3 pow(vec3 x, vec3 y) { return {pow(x.x, y.x), pow(x.y, y.y), pow(x.z, y.z)}; }
__device__ __forceinline__ float min(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ float max(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ vec3 min(vec3 a, vec3 b) { return {min(a.x, b.x), min(a.y, b.y), min(a.z, b.z)}; }
__device__ __forceinline__ vec3 max(vec3 a, vec3 b) { return {max(a.x, b.x), max(a.y, b.y), max(a.z, b.z)}; }
__device__ __forceinline__ int min(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int max(int a, int b) { return a > b ? a : b; }
// WGSL clamp(e, low, high) = min(max(e, low), high); linon calls it as clamp(0.0, 1.0, x)
// (src/functions.rs:45-46) which therefore evaluates to min(1, x).
__device__ __forceinline__ float clamp(float e, float lo, float hi) { return min(max(e, lo), hi); }
__device__ __forceinline__ vec3 clamp(vec3 e, vec3 lo, vec3 hi) { return min(max(e, lo), hi); }
__device__ __forceinline__ float mix(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ vec3 mix(vec3 a, vec3 b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ float step(float edge, float x) { return x < edge ? 0.0f : 1.0f; }
__device__ __forceinline__ float smoothstep(float e0, float e1, float x) {
    float t = clamp((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
__device__ __forceinline__ float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ int abs(int x) { return x < 0 ? -x : x; }
__device__ __forceinline__ int sign(int x) { return (x > 0) - (x < 0); }
__device__ __forceinline__ int clamp(int e, int lo, int hi) { return min(max(e, lo), hi); }
__device__ __forceinline__ unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }
__device__ __forceinline__ unsigned max(unsigned a, unsigned b) { return a > b ? a : b; }
__device__ __forceinline__ unsigned clamp(unsigned e, unsigned lo, unsigned hi) { return min(max(e, lo), hi); }
// WGSL % : truncated remainder for floats (a - b * trunc(a / b)).  Integer / and % are total
// in WGSL: where C++ leaves x / 0 and INT_MIN / -1 undefined, WGSL divides by 1 instead, so
// x / 0 = x, INT_MIN / -1 = INT_MIN, and the remainder of both cases is 0.
__device__ __forceinline__ float wmod(float a, float b) { return a - b * __builtin_truncf(a / b); }
__device__ __forceinline__ vec3 wmod(vec3 a, vec3 b) { return {wmod(a.x, b.x), wmod(a.y, b.y), wmod(a.z, b.z)}; }
__device__ __forceinline__ int wdivisor(int a, int b) { return (b == 0 || (a == (-2147483647 - 1) && b == -1)) ? 1 : b; }
__device__ __forceinline__ int wdiv(int a, int b) { return a / wdivisor(a, b); }
__device__ __forceinline__ int wmod(int a, int b) {
    const int d = wdivisor(a, b);
    return a - (a / d) * d;
}
__device__ __forceinline__ unsigned wdiv(unsigned a, unsigned b) { return a / (b == 0u ? 1u : b); }
__device__ __forceinline__ unsigned wmod(unsigned a, unsigned b) {
    const unsigned d = b == 0u ? 1u : b;
    return a - (a / d) * d;
}
template <int X, int Y, int Z>
__device__ __forceinline__ vec3 swizzle3(vec3 v) {
    const float c[3] = {v.x, v.y, v.z};
    return {c[X], c[Y], c[Z]};
}

// ---- helpers of src/main_view.wgsl:76-146 ----------------------------------------------
constexpr float PI = 3.141592653589793f;

// sin and cos of one angle.  The WGSL-accuracy build (LINON_FAST_MATH, compiled with
// -fgpu-approx-transcendentals) uses the hardware sine/cosine, within WGSL's stated bound
// (absolute error 2^-11 on [-pi, pi]); the IEEE build OCML's full-precision sincosf.
__device__ __forceinline__ void sin_cos(float x, float* s, float* c) {
#ifdef LINON_FAST_MATH
    *s = __sinf(x);
    *c = __cosf(x);
#else
    ::sincosf(x, s, c);
#endif
}

// The three rotations multiply by mat3x3 columns as the WGSL text lays them out; rotateZ's
// third column is zero, so its z component is always 0 (src/main_view.wgsl:94-100).
__device__ __forceinline__ vec3 rotateX(vec3 v, float phi) {
    float s, c;
    sin_cos(phi, &s, &c);
    return {v.x, c * v.y - s * v.z, s * v.y + c * v.z};
}
__device__ __forceinline__ vec3 rotateY(vec3 v, float phi) {
    float s, c;
    sin_cos(phi, &s, &c);
    return {c * v.x + s * v.z, v.y, -s * v.x + c * v.z};
}
#ifdef LINON_FAST_MATH
// rotateY(v, 2 pi turns) with the hardware sine and cosine, which take their argument in
// turns: rotateY(v, phi) computes them of RN(phi * RN(1 / (2 pi))), so a caller whose angle is
// a power of two times x (swirl: 0.5 |p|) passes RN(x * (0.5 RN(1 / (2 pi)))), the same value
// wherever that scaling of x is exact, one multiplication fewer.
__device__ __forceinline__ vec3 rotateY_turns(vec3 v, float turns) {
    const float s = __builtin_amdgcn_sinf(turns);
    const float c = __builtin_amdgcn_cosf(turns);
    return {c * v.x + s * v.z, v.y, -s * v.x + c * v.z};
}
constexpr float INV_2PI = 0.15915494f;  // RN(1 / (2 pi)), the factor of the hardware sin/cos
#endif
__device__ __forceinline__ vec3 rotateZ(vec3 v, float phi) {
    float s, c;
    sin_cos(phi, &s, &c);
    return {c * v.x - s * v.y, s * v.x + c * v.y, 0.0f};
}
__device__ __forceinline__ vec3 translate(vec3 v, float dx, float dy, float dz) {
    return {v.x + dx, v.y + dy, v.z + dz};
}
// Air refraction index at temperature t (deg C), Zhao et al.; pow(10.0, -10.0) is a
// constant expression and enters as 1e-10.  Numerator and denominator are separate helpers so
// that both quotient forms below round the same operands.
__device__ __forceinline__ float refraction_index_num(float t) {
    const float air_pressure = 101325.0f;
    const float c1 = 0.0000104f;
    return c1 * air_pressure * (1.0f + air_pressure * (60.1f - 0.972f * t) * 1e-10f);
}
__device__ __forceinline__ float refraction_index_den(float t) {
    const float c2 = 0.00366f;
    return 1.0f + c2 * t;
}
__device__ __forceinline__ float refraction_index(float t) { return refraction_index_num(t) / refraction_index_den(t); }
// InRange: temperatures in [-250, 1e4] deg C (or NaN).  There the index's numerator lies in
// [0.95, 1.06] and its denominator in [0.085, 37.6], so both indices lie in [0.025, 12.5] and
// every quotient has its operands in div_in_range's range.
__device__ __forceinline__ float mul_rounded(float a, float b) {  // a product the compiler may not fuse
#pragma clang fp contract(off)
    return a * b;
}
template <bool InRange>
__device__ __forceinline__ vec3 refraction_impl(float t_in, float t_out, vec3 v_in, vec3 n) {
    auto quot = [](float a, float b) { return InRange ? div_in_range(a, b) : a / b; };
    float eta_in = quot(refraction_index_num(t_in), refraction_index_den(t_in));
    float eta_out = quot(refraction_index_num(t_out), refraction_index_den(t_out));
    float cosi = clamp(-1.0f, 1.0f, dot(v_in, n));
    vec3 n_ref = n;
    if (cosi < 0.0f) {
        cosi = -cosi;
    } else {
        n_ref = -n;
    }
    float eta = quot(eta_in, eta_out);
    // k = RN(1 - e t), e = RN(eta eta), t = RN(1 - cosi cosi) in [0, 1] (the FMA form the
    // contracted expression 1 - eta * eta * (1 - cosi * cosi) compiles to).  k is never in
    // (0, 2^-48): where |1 - e t| < 2^-48 the exact product e t lies in [0.5, 2], so the
    // exponents of e and t add up to at least -2 and e t is a multiple of 2^-48, as is 1 - e t.
    // The root therefore needs no rescaling of tiny operands.
    const float k = __builtin_fmaf(-(eta * eta), __builtin_fmaf(-cosi, cosi, 1.0f), 1.0f);
    // eta v_in + (eta cosi - sqrt(k)) n_ref with its contraction written out, fma(eta, v_i,
    // RN(X n_i)), X = fma(eta, cosi, -sqrt(k)): the form the compiler picks for the expression, so
    // that it no longer depends on that pick.  NaN where k < 0.
    const float X = __builtin_fmaf(eta, cosi, -sqrt_rn_no_tiny(k));
    vec3 out = {__builtin_fmaf(eta, v_in.x, mul_rounded(X, n_ref.x)),
                __builtin_fmaf(eta, v_in.y, mul_rounded(X, n_ref.y)),
                __builtin_fmaf(eta, v_in.z, mul_rounded(X, n_ref.z))};
    // total internal reflection (k < 0) on any lane: a wave-uniform branch, so that waves
    // without it (nearly all) carry no reflection code, not even with an empty lane mask
    if (__any(k < 0.0f)) {
        const vec3 r = reflect(v_in, n_ref);
        const bool tir = k < 0.0f;
        out = {tir ? r.x : out.x, tir ? r.y : out.y, tir ? r.z : out.z};
    }
    return out;
}
__device__ __forceinline__ vec3 refraction(float t_in, float t_out, vec3 v_in, vec3 n) {
    return refraction_impl<false>(t_in, t_out, v_in, n);
}
// refraction() for temperatures in [-250, 1e4] deg C (the predefined fields': hot air's lie in
// [20, 100], the mirages' in [15, 200]); bit for bit the same result.
__device__ __forceinline__ vec3 refraction_in_range(float t_in, float t_out, vec3 v_in, vec3 n) {
    return refraction_impl<true>(t_in, t_out, v_in, n);
}
// refraction_in_range(t, t, v_in, n): one temperature on both sides, as in the first RK4 stage
// of a field that compares T(p_prev) with T(p).  The index ratio n(t) / n(t) is then exactly 1
// (a correctly rounded x / x), so its quotients and every product with it drop out:
//   k = fma(-(eta eta), fma(-cosi, cosi, 1), 1)  ->  1 - fma(-cosi, cosi, 1)
//   X = fma(eta, cosi, -sqrt(k))                 ->  cosi - sqrt(k)
//   fma(v_in, eta, RN(X n))                      ->  v_in + RN(X n)
// which is how refraction_impl's expression compiles, so the result is the same bit for bit.  A
// NaN t gives NaN in every component, as its NaN ratio does there (cosi - (t - t) is cosi for
// a finite t).  The WGSL-accuracy build divides approximately, where n(t) / n(t) need not be
// 1, and keeps the general form.
__device__ __forceinline__ vec3 refraction_in_range_same(float t, vec3 v_in, vec3 n) {
#ifdef LINON_FAST_MATH
    return refraction_impl<true>(t, t, v_in, n);
#else
    float cosi = clamp(-1.0f, 1.0f, dot(v_in, n)) - (t - t);
    vec3 n_ref = n;
    if (cosi < 0.0f) {
        cosi = -cosi;
    } else {
        n_ref = -n;
    }
    // 1 - t with t = RN(1 - cosi cosi) in [0, 1]: zero, or at least 2^-24 (t >= 1/2 is a
    // multiple of 2^-24) -- no tiny operand for the root; never negative, so never the total
    // internal reflection of the general form (a NaN t gives NaN there as here)
    const float k = 1.0f - __builtin_fmaf(-cosi, cosi, 1.0f);
    const float x = cosi - sqrt_rn_no_tiny(k);
    auto axpy = [](float a, float b, float c) {
#pragma clang fp contract(off)
        return a * b + c;
    };
    return {axpy(x, n_ref.x, v_in.x), axpy(x, n_ref.y, v_in.y), axpy(x, n_ref.z, v_in.z)};
#endif
}
__device__ __forceinline__ float point_plane_distance(vec3 p, vec3 n, vec3 p0) {
    float d = dot(p0, n);
    return __builtin_fabsf(dot(p, n) - d);
}
__device__ __forceinline__ float sigmoid(float x) { return 1.0f / (1.0f + ::expf(-x)); }
// sigmoid(x) with its reciprocal taken of min(1 + e^-x, 2^125): the operand is then in
// [1, 2^125] (or NaN), so the correctly rounded reciprocal needs no range test.  Equal to
// sigmoid(x) wherever sigmoid(x) >= 2^-125, and 2^-125 below, for callers to which the two are
// indistinguishable (hot air's 20 + 80 s rounds to 20 for any s below 2^-118).
__device__ __forceinline__ float sigmoid_floor(float x) {
#ifdef LINON_FAST_MATH
    return sigmoid(x);
#else
    const float b = 1.0f + ::expf(-x);
    return recip_rn(b > 0x1p125f ? 0x1p125f : b);
#endif
}
// e^y at WGSL's accuracy: the hardware 2^t of t = RN(y L), L = RN(log2 e).  The product's
// rounding and L's error (together below 7.3e-8 relative) move t by at most 7.3e-8 |y| log2 e,
// so 2^t by at most 7.3e-8 |y| relative, about 1.2 |y| ulp, on top of the instruction's own
// error; WGSL bounds exp by 3 + 2 |y| ulp, which linon_selftest_exp checks at every float in
// [-87, 87].  Two instructions.  Large y gives +inf, small y 0 (or a subnormal), NaN gives NaN.
__device__ __forceinline__ float exp_wgsl(float y) { return __builtin_amdgcn_exp2f(y * 1.44269502f); }
// 1 / (1 + e^-x) at WGSL's accuracy: exp_wgsl and the hardware reciprocal (within 1 ulp, where
// WGSL bounds the division by 2.5 ulp).  e^-x = +inf gives 0, NaN gives NaN.
__device__ __forceinline__ float sigmoid_wgsl(float x) { return __builtin_amdgcn_rcpf(1.0f + exp_wgsl(-x)); }

}  // namespace lw
