// Predefined field functions as device functors.  Bodies 1..9 are the WGSL bodies of
// PredefinedFunction::to_code (src/functions.rs:33-128) written in the lw:: vocabulary;
// Identity is the stub of src/main_view.wgsl:148; Swirl/Lens/HotAir/Warp are the four
// benchmark fields.  Every functor has the field_function signature
//   (p_prev, p, v0, v, t) -> vec3   (src/main_view.wgsl:148, README.md:28-36).
#pragma once
#include "wgsl_lib.hpp"

namespace lw {

// A functor may also define first_stage(p, v0, v, t) = (*this)(p, p, v0, v, t), bit for bit,
// with less work; the integrators call it for the first RK4 stage, where p_prev == p.
template <bool Spherical, bool Sigmoid>
struct FieldMirage {  // src/functions.rs:36-96
    // distance from the hot source: the sphere's centre or the plane
    __device__ __forceinline__ static float distance_to_source(vec3 q) {
        if (Spherical) return length(q - vec3f(-0.5f, 0.5f, -0.5f));
        return point_plane_distance(q, vec3f(0.0f, 1.0f, 0.0f), vec3f(0.0f, 0.1f, 0.0f));
    }
    __device__ __forceinline__ static vec3 normal_at(vec3 p) {
        if (Spherical) return normalize(p - vec3f(-0.5f, 0.5f, -0.5f));
        return vec3f(0.0f, 1.0f, 0.0f);
    }
    __device__ __forceinline__ static float temperature(float dist) {  // in [15, 200]
        const float t_env = 15.0f;
        const float t_src = Spherical ? 200.0f : 30.0f;
        const float max_dist = Spherical ? 0.25f : 0.01f;
        float part;
        if (Sigmoid) {
            // arguments >= -6 (distances are >= 0): sigmoid_floor equals sigmoid here
            part = clamp(0.0f, 1.0f, sigmoid_floor(dist / max_dist * 12.0f - 6.0f));
        } else {
            part = clamp(0.0f, 1.0f, dist / max_dist);
        }
        return part * t_env + (1.0f - part) * t_src;
    }
    __device__ __forceinline__ vec3 operator()(vec3 p_prev, vec3 p, vec3, vec3 v, float) const {
        return refraction_in_range(temperature(distance_to_source(p_prev)), temperature(distance_to_source(p)), v,
                                   normal_at(p));
    }
    __device__ __forceinline__ vec3 first_stage(vec3 p, vec3, vec3 v, float) const {
        return refraction_in_range_same(temperature(distance_to_source(p)), v, normal_at(p));
    }
};
using FieldMirageSpherical = FieldMirage<true, false>;
using FieldMiragePlane = FieldMirage<false, false>;
using FieldMirageSphericalSigmoid = FieldMirage<true, true>;
using FieldMiragePlaneSigmoid = FieldMirage<false, true>;

struct FieldTranslationX {  // src/functions.rs:98-101
    __device__ __forceinline__ vec3 operator()(vec3, vec3, vec3 v0, vec3, float t) const {
        return translate(v0, 0.5f * t, 0.0f, 0.0f);
    }
};
struct FieldTranslationZ {  // src/functions.rs:103-106
    __device__ __forceinline__ vec3 operator()(vec3, vec3, vec3 v0, vec3, float t) const {
        return translate(v0, 0.0f, 0.0f, 0.5f * t);
    }
};
struct FieldRotation {  // src/functions.rs:108
    __device__ __forceinline__ vec3 operator()(vec3, vec3, vec3 v0, vec3, float t) const {
        return normalize(v0 + rotateZ(v0, PI * t));
    }
};
struct FieldLorenz {  // src/functions.rs:109-116
    __device__ __forceinline__ vec3 operator()(vec3, vec3 p, vec3, vec3, float) const {
        const float rho = 28.0f, sigma = 10.0f, beta = 8.0f / 3.0f;
        return vec3f(sigma * (p.y - p.x), p.x * (rho - p.z) - p.y, p.x * p.y - beta * p.z);
    }
};
struct FieldRoessler {  // src/functions.rs:118-125
    __device__ __forceinline__ vec3 operator()(vec3, vec3 p, vec3, vec3, float) const {
        const float a = 0.1f, b = 0.1f, c = 14.0f;
        return vec3f(-p.y - p.z, p.x + a * p.y, b + p.z * (p.x - c));
    }
};
struct FieldIdentity {  // src/main_view.wgsl:148
    __device__ __forceinline__ vec3 operator()(vec3, vec3, vec3, vec3 v, float) const { return v; }
};
struct FieldSwirl {  // rotateY(v0, 0.5 * length(p))
    __device__ __forceinline__ vec3 operator()(vec3, vec3 p, vec3 v0, vec3, float) const {
#ifdef LINON_FAST_MATH
        // the angle's factor 0.5 joins the turns conversion (exact for |p| >= 2^-125)
        return rotateY_turns(v0, length(p) * (0.5f * INV_2PI));
#else
        return rotateY(v0, 0.5f * length(p));
#endif
    }
};
// Lens.  RecipAbove: the variant AUTO picks for lens, whose quotient above the floor is taken
// at WGSL's accuracy (div3_lens_recip); |p|, the floor and its quotient stay exact.
template <bool RecipAbove>
struct FieldLensT {  // v - 0.3 * p / max(|p|^3, 1e-3)
    // |p| enters only through s = max(|p|^3, 0.001).  Where dot(p, p) < 2^-96 the root is below
    // 2^-47 whatever its last bits (or NaN, which max also drops), its cube below the floor, and
    // s the floor 0.001 either way; elsewhere sqrt_rn_no_tiny is sqrt_rn.  So |p| needs no
    // rescaling of tiny operands, and s (hence the field) is bit for bit length()'s.
    __device__ __forceinline__ vec3 operator()(vec3, vec3 p, vec3, vec3 v, float) const {
        const float d2 = dot(p, p);
#ifndef LINON_FAST_MATH
        // The hardware root is within an ulp of |p|: below 0.0999 on every lane of the wave
        // (|p|^3 < 0.000998 however |p| rounds), s is the floor throughout and the wave skips
        // the correctly rounded root and the cube (a wave-uniform branch: the lens's captured
        // centre, where most waves are).  NaN fails the test and takes the general path.
        if (__all(__builtin_amdgcn_sqrtf(d2) < 0.0999f)) return v - div3_rn_lens_floor(0.3f * p);
#endif
        const float r = sqrt_rn_no_tiny(d2);
        const float s = max(r * r * r, 0.001f);
        return v - (RecipAbove ? div3_lens_recip(0.3f * p, s) : div3_rn_lens(0.3f * p, s));  // 0.3 p / s
    }
};
using FieldLens = FieldLensT<false>;
using FieldLensRecip = FieldLensT<true>;
// Hot air: refraction(T(p_prev), T(p), v, normalize(p - c)), T = 20 + 80 sigmoid(8 (0.5 - |p - c|)).
// FaithfulRoot: the variant AUTO picks for hot air.  Its temperature and normal are taken at
// WGSL's accuracy: the distance |p - c| (the temperature's argument and the normal's length)
// with the hardware square root, within 1 ulp (WGSL bounds sqrt only to inverseSqrt's
// accuracy); the sigmoid with the hardware exponential and reciprocal (sigmoid_wgsl); the
// normal's quotients through one correctly rounded reciprocal.  Refraction (the index ratio and
// sqrt(k), where rounding turns into exit error near grazing incidence) stays correctly
// rounded.  Its exit points pass every parity test as the exact variant's do, in about 40%
// fewer instructions per step.  The temperature enters the step only through the refraction
// indices, whose relative slope is at most 3.5e-3 per degree: an ulp of the sigmoid moves them
// by about 1e-8, below the index ratio's own rounding.
template <bool FaithfulRoot>
struct FieldHotAirT {
    __device__ __forceinline__ static vec3 offset(vec3 q) { return q - vec3f(0.0f, -0.5f, 0.0f); }
    __device__ __forceinline__ static float distance(vec3 q) {
        const vec3 d = offset(q);
        return FaithfulRoot ? __builtin_amdgcn_sqrtf(dot(d, d)) : length(d);
    }
    __device__ __forceinline__ static float temperature(vec3 q) {
        // 8 (0.5 - d) = -RN(8 d - 4) exactly: scaling by 8 commutes with the rounding, so one
        // FMA forms the product of the rounded difference
        const float x = -__builtin_fmaf(8.0f, distance(q), -4.0f);
#ifdef LINON_FAST_MATH
        return 20.0f + 80.0f * sigmoid_floor(x);  // in [20, 100]
#else
        return 20.0f + 80.0f * (FaithfulRoot ? sigmoid_wgsl(x) : sigmoid_floor(x));
#endif
    }
    // normalize(p - c).  The faithful variant multiplies by the correctly rounded reciprocal of
    // its faithful |p - c|: RN(d_i RN(1 / s)) is within about one ulp of d_i / s (two roundings
    // of half an ulp), where WGSL bounds normalize by x / length(x), a 2.5-ulp division.  No
    // range test: s = 0 gives NaN components as 0 / 0 does.
    __device__ __forceinline__ static vec3 normal(vec3 p) {
        if (FaithfulRoot) {
            const vec3 d = offset(p);
            const float y = recip_rn(distance(p));
            return {d.x * y, d.y * y, d.z * y};
        }
        return normalize(offset(p));
    }
    __device__ __forceinline__ vec3 operator()(vec3 p_prev, vec3 p, vec3, vec3 v, float) const {
        return refraction_in_range(temperature(p_prev), temperature(p), v, normal(p));
    }
    __device__ __forceinline__ vec3 first_stage(vec3 p, vec3, vec3 v, float) const {
        return refraction_in_range_same(temperature(p), v, normal(p));
    }
};
using FieldHotAir = FieldHotAirT<false>;
using FieldHotAirFaithfulRoot = FieldHotAirT<true>;
struct FieldWarp {  // v0 + 0.3 * (sin 4p.y, 0, sin 4p.x)
    __device__ __forceinline__ vec3 operator()(vec3, vec3 p, vec3 v0, vec3, float) const {
#ifdef LINON_FAST_MATH
        // the hardware sine takes turns: sin(4 x) is v_sin(RN(RN(4 x) RN(1 / (2 pi)))), and
        // RN(4 x) = 4 x exactly (below overflow), so one product with 4 RN(1 / (2 pi)) gives the
        // same argument, one multiplication fewer per sine; frames bit-identical
        return v0 + 0.3f * vec3f(__builtin_amdgcn_sinf(p.y * (4.0f * INV_2PI)), 0.0f,
                                 __builtin_amdgcn_sinf(p.x * (4.0f * INV_2PI)));
#else
        return v0 + 0.3f * vec3f(sin(4.0f * p.y), 0.0f, sin(4.0f * p.x));
#endif
    }
};

}  // namespace lw
