// Learned flow-map surrogate: device side.
//
// The surrogate replaces the field integration between a ray's entry into the bounding sphere
// (radius R) and its exit: (p_entry, q) -> (p_exit, d_exit), plus a confidence.  q is the
// direction that, with p, fixes the rest of the path: the initial direction v0 for fields that
// read v0 (swirl, warp) and the current direction for fields that read v (lens, hot-air).
// Rays start at the camera outside the sphere, so a frame integrates with RK4 up to the entry
// point (approach_body), runs the MLP on the entered rays, and integrates with RK4 again
// (flowmap_body) wherever the surrogate is not trusted.
#pragma once
#include "mlp.hpp"
#include "params.hpp"
#include "raycast.hpp"

#define LS_FREQS 16
#define LS_IN 6
#define LS_FEATS (LS_IN + 2 * LS_IN * LS_FREQS)  // 198
#define LS_FEATS_PAD 224                            // padded to a multiple of the GEMM k-tile
#define LS_HIDDEN 256
#define LS_OUT 7                                    // exit point, exit direction, confidence logit

namespace ls {

using lw::vec3;
using lw::vec3f;

struct SurrogateIO {
    // entry states (p, q) of the rays the MLP sees, and where their results go
    const float* states;   // [n][6]
    int n;
    float radius;
    const float* freqs;    // [LS_FREQS]
    lm::u16* features;     // [n][LS_FEATS_PAD] bf16
};

// Fourier features of an entry state: (p / R, q), then sin and cos of every frequency times
// every coordinate; zero padding up to LS_FEATS_PAD.
__device__ __forceinline__ void features_body(const SurrogateIO& S) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.n) return;
    float x[LS_IN];
    const float inv_r = 1.0f / S.radius;
    for (int k = 0; k < 3; ++k) x[k] = S.states[6 * (size_t)i + k] * inv_r;
    for (int k = 3; k < 6; ++k) x[k] = S.states[6 * (size_t)i + k];
    lm::u16* o = S.features + (size_t)i * LS_FEATS_PAD;
    for (int k = 0; k < LS_IN; ++k) o[k] = lm::f2bf(x[k]);
    int c = LS_IN;
    for (int f = 0; f < LS_FREQS; ++f) {
        const float w = S.freqs[f];
        for (int k = 0; k < LS_IN; ++k) {
            float s, co;
            __sincosf(w * x[k], &s, &co);
            o[c + 2 * (f * LS_IN + k)] = lm::f2bf(s);
            o[c + 2 * (f * LS_IN + k) + 1] = lm::f2bf(co);
        }
    }
    for (int k = LS_FEATS; k < LS_FEATS_PAD; ++k) o[k] = 0;
}

// ---- RK4 legs of a surrogate frame ------------------------------------------------------
struct ApproachParams {
    CameraParams cam;
    StepParams step;        // escape mode, need_mapping
    float radius;
    int tiles_x;
    float* entry_states;    // [W*H][6]: (p_entry, q) of rays that entered, by texel
    int* entered;           // [W*H]: 1 entered, 0 finished outside the sphere
    float* mapping;         // [W*H] float4: final point of rays that never entered
    int* steps;             // [W*H] RK4 iterations spent
    int q_is_v0;            // 1: the field reads v0 (q = v0), 0: it reads v (q = current direction)
};

struct InsideSphere {
    float r2;
    __device__ __forceinline__ bool operator()(const lk::RayState& r) const { return lk::rn::dot(r.p, r.p) < r2; }
};

// Integrates a camera ray until it is inside the sphere (|p| < R) or finished (escape test,
// max_steps).  Fields act outside the sphere too, so this leg is plain RK4.
template <class F>
__device__ void approach_body(const ApproachParams& A) {
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int tiles_y = (A.cam.height + LINON_TILE - 1) / LINON_TILE;
    if (tile >= A.tiles_x * tiles_y) return;
    const int gx = (tile % A.tiles_x) * LINON_TILE + (lane & (LINON_TILE - 1));
    const int gy = (tile / A.tiles_x) * LINON_TILE + (lane / LINON_TILE);
    if (gx >= A.cam.width || gy >= A.cam.height) return;
    const lk::CameraFrame cam = lk::camera_frame(A.cam);
    lk::RayState r;
    lk::ray_begin(r, cam.origin, lk::camera_dir(cam, A.cam, gx, gy), A.step);
    SceneParams none;
    none.accel = LINON_ACCEL_BRUTE_K;
    none.num_faces = 0;
    none.tris = nullptr;
    none.skip_margin = 0.0f;
    const float r2 = A.radius * A.radius;
    bool entered = lk::rn::dot(r.p, r.p) < r2;
    bool done = false;
    lk::QueryStats st{0ull, 0u, 0u, 0u};  // no mesh: nothing is counted
    if (!entered) {
        // steps until the ray is inside the sphere (the loop's own stop test) or finished
        done = lk::ray_advance<F>(r, A.step, none, nullptr, 0x7fffffff, st, InsideSphere{r2});
        entered = lk::rn::dot(r.p, r.p) < r2;
    }
    const size_t texel = (size_t)(A.cam.height - 1 - gy) * A.cam.width + gx;
    A.steps[texel] = r.iters;
    if (entered && !done) {
        const vec3 q = A.q_is_v0 ? r.v0 : r.d;
        float* s = A.entry_states + 6 * texel;
        s[0] = r.p.x; s[1] = r.p.y; s[2] = r.p.z;
        s[3] = q.x;   s[4] = q.y;   s[5] = q.z;
        A.entered[texel] = 1;
    } else {
        A.entered[texel] = 0;
        reinterpret_cast<float4*>(A.mapping)[texel] = make_float4(r.p.x, r.p.y, r.p.z, 1.0f);
    }
}

struct FlowmapParams {
    StepParams step;        // escape mode, need_mapping, no mesh
    const float* states;    // [n][6] entry states (p, q)
    const int* index;       // optional: states[index[i]] (compacted subsets); null = identity
    int n;                  // states (or list entries); with n_dev, the launch's capacity
    const int* n_dev;       // optional: the count on the device (a frame's compacted lists)
    float* exits;           // [n or max index][6] exit point and direction, at the state's slot
    int* steps;             // optional, RK4 iterations per state slot (added to what is there)
    // optional (a frame's fallback): the RK4 iterations the ray already ran before its entry
    // state, iters_at[texel_of[slot]], which count toward max_steps as in one plain RK4 run
    const int* iters_at;
    const int* texel_of;
};

// RK4 from an entry state to the sphere exit (the surrogate's training labels, and the
// fallback of a frame).
template <class F>
__device__ void flowmap_body(const FlowmapParams& P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (P.n_dev ? *P.n_dev : P.n)) return;
    const int slot = P.index ? P.index[i] : i;
    const float* s = P.states + 6 * (size_t)slot;
    lk::RayState r;
    lk::ray_begin(r, vec3f(s[0], s[1], s[2]), vec3f(s[3], s[4], s[5]), P.step);
    const int before = P.iters_at ? P.iters_at[P.texel_of[slot]] : 0;
    r.iters = before;
    SceneParams none;
    none.accel = LINON_ACCEL_BRUTE_K;
    none.num_faces = 0;
    none.tris = nullptr;
    none.skip_margin = 0.0f;
    lk::QueryStats st{0ull, 0u, 0u, 0u};  // no mesh: nothing is counted
    lk::ray_advance<F>(r, P.step, none, nullptr, 0x7fffffff, st);
    float* o = P.exits + 6 * (size_t)slot;
    o[0] = r.p.x; o[1] = r.p.y; o[2] = r.p.z;
    o[3] = r.d.x; o[4] = r.d.y; o[5] = r.d.z;
    if (P.steps) P.steps[slot] += r.iters - before;
}

}  // namespace ls
