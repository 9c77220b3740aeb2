// The ray-cast kernel body, templated on the field functor.
//
// One lane traces one pixel's ray through the field with RK4 and intersects every step's
// segment with the scene until the first hit (src/main_view.wgsl:229-279); one wave64 covers
// an 8x8 pixel tile, like the reference's @workgroup_size(8, 8).  Block 0 of the launch can
// trace the 8 sample rays of src/main_view.wgsl:377-480 so that this long serial work starts
// first and overlaps the pixel tiles.
//
// Scene queries.  Brute force (LINON_ACCEL_BRUTE_K) tests all faces in index order as the
// reference does, with up to LINON_MAX_BRUTE_LDS faces staged in LDS.  The grid path keeps a
// per-lane lower bound on the distance to the mesh (from a Chebyshev distance field over a
// uniform grid) and skips every segment that provably cannot reach a face; the remaining
// segments walk the grid cells they cross (3-D DDA) and test the faces listed there.  Both
// paths use the same Moeller-Trumbore arithmetic and the same (t, face index) lexicographic
// minimum, so they return the same face and t bit for bit.
#pragma once
#include "params.hpp"
#include "wgsl_lib.hpp"

// Waves per SIMD the ray-cast kernels (static and run-time compiled) are built for.  6 caps a
// kernel at 80 VGPRs; the compiler's own choice is 94-103 (4-5 waves).  At 6 a few values
// spill outside the step loop; the four 1080p benchmark fields render faster than at 4 waves,
// as fast as at 5 (within run-to-run noise) and faster than at 7 (2-7%) or 8.  0 leaves the
// choice to the compiler.
#ifndef LK_RAYCAST_WAVES_PER_EU
#define LK_RAYCAST_WAVES_PER_EU 6
#endif
#if LK_RAYCAST_WAVES_PER_EU > 0
#define LK_RAYCAST_OCCUPANCY __attribute__((amdgpu_waves_per_eu(LK_RAYCAST_WAVES_PER_EU)))
#else
#define LK_RAYCAST_OCCUPANCY
#endif
// The same for the persistent batch kernel (256-lane blocks): the compiler's own 121-131 VGPRs
// allow 3-4 waves; at 6 (80 VGPRs) the 65-view batch runs 10% faster, at 7 slower again and at
// 8 twice as slow.
#ifndef LK_BATCH_WAVES_PER_EU
#define LK_BATCH_WAVES_PER_EU 6
#endif
#if LK_BATCH_WAVES_PER_EU > 0
#define LK_BATCH_OCCUPANCY __attribute__((amdgpu_waves_per_eu(LK_BATCH_WAVES_PER_EU)))
#else
#define LK_BATCH_OCCUPANCY
#endif

namespace lk {

using lw::vec3;
using lw::vec3f;
using lw::dot;
using lw::cross;
using lw::length;
using lw::normalize;

__device__ __forceinline__ vec3 ld3(const float* p) { return vec3f(p[0], p[1], p[2]); }

__device__ __forceinline__ vec3 ld_tri_vec(const float* tris, int face, int k) {
    const float4 q = reinterpret_cast<const float4*>(tris)[3 * face + k];
    return vec3f(q.x, q.y, q.z);
}

// WGSL-literal arithmetic: every operation rounded once, no FMA contraction, in the order the
// shader text writes it (and fma, where the integrator contracts on purpose: the RK4 stage
// points).  The integrator, camera and sample-ray statements use it, so every kernel that
// advances a ray (plain, batched, the surrogate's RK4 legs) rounds them alike whatever the
// surrounding code lets the compiler fuse.
namespace rn {
__device__ __forceinline__ float mul(float a, float b) {
#pragma clang fp contract(off)
    return a * b;
}
__device__ __forceinline__ float add(float a, float b) {
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ float sub(float a, float b) {
#pragma clang fp contract(off)
    return a - b;
}
__device__ __forceinline__ vec3 mul(float s, vec3 a) { return vec3f(mul(s, a.x), mul(s, a.y), mul(s, a.z)); }
__device__ __forceinline__ vec3 mul(vec3 a, float s) { return vec3f(mul(a.x, s), mul(a.y, s), mul(a.z, s)); }
__device__ __forceinline__ vec3 add(vec3 a, vec3 b) { return vec3f(add(a.x, b.x), add(a.y, b.y), add(a.z, b.z)); }
__device__ __forceinline__ vec3 sub(vec3 a, vec3 b) { return vec3f(sub(a.x, b.x), sub(a.y, b.y), sub(a.z, b.z)); }
// s a + b with one rounding per component
__device__ __forceinline__ vec3 fma(float s, vec3 a, vec3 b) {
    return vec3f(__builtin_fmaf(s, a.x, b.x), __builtin_fmaf(s, a.y, b.y), __builtin_fmaf(s, a.z, b.z));
}
__device__ __forceinline__ float dot(vec3 a, vec3 b) { return add(add(mul(a.x, b.x), mul(a.y, b.y)), mul(a.z, b.z)); }
__device__ __forceinline__ vec3 cross(vec3 a, vec3 b) {
    return vec3f(sub(mul(a.y, b.z), mul(a.z, b.y)), sub(mul(a.z, b.x), mul(a.x, b.z)),
                 sub(mul(a.x, b.y), mul(a.y, b.x)));
}
__device__ __forceinline__ float length(vec3 a) { return lw::sqrt_rn(rn::dot(a, a)); }
__device__ __forceinline__ vec3 normalize(vec3 a) {
    const float l = rn::length(a);
    return vec3f(a.x / l, a.y / l, a.z / l);
}
}  // namespace rn

// An opaque copy of x: the compiler cannot see that it equals x, so no operation on the copy
// is merged (common-subexpression elimination) with an operation on x.  The integrator passes
// the field its arguments this way.  The field's own arithmetic may be contracted into FMAs
// while the integrator's rn:: arithmetic may not; were a product of the field merged with the
// same product of the integrator (say p.x * p.x of length(p) and of the escape test's
// dot(p, p)), the merged product would lose its contraction, and the field's rounding would
// depend on the kernel around it.  With opaque arguments every kernel that integrates a ray
// rounds the field alike.
__device__ __forceinline__ float opaque(float x) {
    asm("" : "+v"(x));
    return x;
}
__device__ __forceinline__ vec3 opaque(vec3 a) { return vec3f(opaque(a.x), opaque(a.y), opaque(a.z)); }

// Moeller-Trumbore in the operation order of hit_triangle (src/main_view.wgsl:150-175):
// two-sided, |a| < 1e-7 rejected, inclusive edges, t > 1e-7 accepted.  e1 = v1 - v0 and
// e2 = v2 - v0 are precomputed in f32, which gives the same bits as computing them here.
__device__ __forceinline__ float mt_test(vec3 v0, vec3 e1, vec3 e2, vec3 o, vec3 d) {
    const float eps = 0.0000001f;
    vec3 h = rn::cross(d, e2);
    float a = rn::dot(e1, h);
    if (a > -eps && a < eps) return -1.0f;
    float f = 1.0f / a;
    vec3 s = rn::sub(o, v0);
    float u = rn::mul(f, rn::dot(s, h));
    if (u < 0.0f || u > 1.0f) return -1.0f;
    vec3 q = rn::cross(s, e1);
    float v = rn::mul(f, rn::dot(d, q));
    if (v < 0.0f || rn::add(u, v) > 1.0f) return -1.0f;
    float t = rn::mul(f, rn::dot(e2, q));
    return t > eps ? t : -1.0f;
}

// Accept t_new when it is a valid hit on [0, max_dist) that precedes the best so far, with
// the lower face index winning equal t (the reference's strict `<` over ascending faces).
__device__ __forceinline__ void consider(float t_new, int face, float max_dist, float& best_t, int& best_face) {
    if (t_new > 0.0f && t_new < max_dist &&
        (best_t < 0.0f || t_new < best_t || (t_new == best_t && face < best_face))) {
        best_t = t_new;
        best_face = face;
    }
}

// Scene-query counters of one lane, accumulated over the rays it traces.  The kernels keep
// them in LDS (one slot per lane) rather than in registers, which the step loop needs more.
struct QueryStats {
    unsigned long long tests;
    unsigned cells;
    unsigned queries;
    unsigned skips;
};

// Lower bound on the distance from p to any face: 0 inside an occupied cell; else the
// distance to the boundary of the empty (2k-1)^3 block of cells around p's cell, where k is
// the cell's Chebyshev distance to the nearest occupied cell; outside the grid, the distance
// to the grid box (all faces lie inside it).
__device__ __forceinline__ float grid_safe_distance(const SceneParams& P, vec3 p) {
    const float ics = P.inv_cell_size;
    float gx = (p.x - P.grid_lo[0]) * ics;
    float gy = (p.y - P.grid_lo[1]) * ics;
    float gz = (p.z - P.grid_lo[2]) * ics;
    const float nx = P.gdimf[0], ny = P.gdimf[1], nz = P.gdimf[2];
    if (!(gx >= 0.0f && gx < nx && gy >= 0.0f && gy < ny && gz >= 0.0f && gz < nz)) {
        float dx = fmaxf(fmaxf(-gx, gx - nx), 0.0f);
        float dy = fmaxf(fmaxf(-gy, gy - ny), 0.0f);
        float dz = fmaxf(fmaxf(-gz, gz - nz), 0.0f);
        float d2 = dx * dx + dy * dy + dz * dz;
        return (d2 > 0.0f) ? __builtin_sqrtf(d2) * P.cell_size : 0.0f;  // NaN -> 0 -> query
    }
    int ix = (int)gx, iy = (int)gy, iz = (int)gz;
    int k = P.dist[((size_t)iz * P.gdim[1] + iy) * P.gdim[0] + ix];
    if (k == 0) return 0.0f;
    float lo = (float)(k - 1), hi = (float)k;
    float bx = fminf(gx - ((float)ix - lo), ((float)ix + hi) - gx);
    float by = fminf(gy - ((float)iy - lo), ((float)iy + hi) - gy);
    float bz = fminf(gz - ((float)iz - lo), ((float)iz + hi) - gz);
    return fminf(fminf(bx, by), bz) * P.cell_size;
}

__device__ __forceinline__ void test_cell(const SceneParams& P, int cell, vec3 o, vec3 d, float L,
                                          float& best_t, int& best_face, QueryStats& st) {
    unsigned b = P.cell_start[cell], e = P.cell_start[cell + 1];
    st.cells += 1;
    st.tests += (e - b);
    for (unsigned k = b; k < e; ++k) {
        int f = (int)P.cell_faces[k];
        float t = mt_test(ld_tri_vec(P.tris, f, 0), ld_tri_vec(P.tris, f, 1), ld_tri_vec(P.tris, f, 2), o, d);
        consider(t, f, L, best_t, best_face);
    }
}

// Walk the grid cells crossed by the segment o + s*d, s in [0, L] (3-D DDA).  The walk
// takes exactly |dx|+|dy|+|dz| steps from the entry cell to the exit cell and only steps
// along axes that still differ, so it always stays inside the grid and always ends; cells
// grazed within rounding error are covered by the margin the cell lists are built with.
__device__ __forceinline__ void query_grid(const SceneParams& P, vec3 o, vec3 d, float L, float& best_t,
                                           int& best_face, QueryStats& st) {
    const float ics = P.inv_cell_size;
    const float ax = (o.x - P.grid_lo[0]) * ics, ay = (o.y - P.grid_lo[1]) * ics, az = (o.z - P.grid_lo[2]) * ics;
    const float sL = L * ics;
    const float dx = d.x * sL, dy = d.y * sL, dz = d.z * sL;
    const int nx = P.gdim[0], ny = P.gdim[1], nz = P.gdim[2];
    float s0 = 0.0f, s1 = 1.0f;
    // clip the segment to the grid box
#define LK_CLIP(a, dd, n)                                              \
    if (dd == 0.0f) {                                                  \
        if (!(a >= 0.0f && a <= n)) return;                            \
    } else {                                                           \
        float inv = 1.0f / dd;                                         \
        float ta = (0.0f - a) * inv, tb = (n - a) * inv;               \
        s0 = fmaxf(s0, fminf(ta, tb));                                 \
        s1 = fminf(s1, fmaxf(ta, tb));                                 \
    }
    LK_CLIP(ax, dx, P.gdimf[0])
    LK_CLIP(ay, dy, P.gdimf[1])
    LK_CLIP(az, dz, P.gdimf[2])
#undef LK_CLIP
    if (!(s0 <= s1)) return;
    int cx = min(max((int)floorf(ax + dx * s0), 0), nx - 1);
    int cy = min(max((int)floorf(ay + dy * s0), 0), ny - 1);
    int cz = min(max((int)floorf(az + dz * s0), 0), nz - 1);
    const int ex = min(max((int)floorf(ax + dx * s1), 0), nx - 1);
    const int ey = min(max((int)floorf(ay + dy * s1), 0), ny - 1);
    const int ez = min(max((int)floorf(az + dz * s1), 0), nz - 1);
    const float INF = __builtin_inff();
    const int sx = dx > 0.0f ? 1 : -1, sy = dy > 0.0f ? 1 : -1, sz = dz > 0.0f ? 1 : -1;
    const float tdx = dx != 0.0f ? fabsf(1.0f / dx) : INF;
    const float tdy = dy != 0.0f ? fabsf(1.0f / dy) : INF;
    const float tdz = dz != 0.0f ? fabsf(1.0f / dz) : INF;
    float tx = dx != 0.0f ? ((float)(dx > 0.0f ? cx + 1 : cx) - ax) / dx : INF;
    float ty = dy != 0.0f ? ((float)(dy > 0.0f ? cy + 1 : cy) - ay) / dy : INF;
    float tz = dz != 0.0f ? ((float)(dz > 0.0f ? cz + 1 : cz) - az) / dz : INF;
    int remaining = abs(ex - cx) + abs(ey - cy) + abs(ez - cz);
    const size_t nxy = (size_t)nx * ny;
    while (true) {
        test_cell(P, (int)(cz * nxy + (size_t)cy * nx + cx), o, d, L, best_t, best_face, st);
        if (remaining <= 0) break;
        --remaining;
        float mx = (cx != ex) ? tx : INF;
        float my = (cy != ey) ? ty : INF;
        float mz = (cz != ez) ? tz : INF;
        if (cx != ex && mx <= my && mx <= mz) {
            cx += sx;
            tx += tdx;
        } else if (cy != ey && (my <= mz || cz == ez)) {
            cy += sy;
            ty += tdy;
        } else if (cz != ez) {
            cz += sz;
            tz += tdz;
        } else if (cx != ex) {
            cx += sx;
            tx += tdx;
        } else {
            cy += sy;
            ty += tdy;
        }
    }
}

// (k1 + 2 k2 + 2 k3 + k4) / 6.0 with the correctly rounded quotient of the WGSL division:
// q = s * RN(1/6), then one FMA residual correction in the signed-zero form
// (lw::quot_rn_signed_zero, which also keeps -0 / 6 = -0), so that e.g. a uniform field keeps
// V == v bit for bit as the f32 division does.  Equal to s / 6 for every s except subnormal
// quotients at an exact tie (s = 3 (2m + 1) 2^-149 with m odd), where RN(1/6) > 1/6 tips the
// correction to the other neighbour (an ulp of 2^-149 in V), and s = +-inf, where the residual
// is inf - inf and V NaN (a field value the ray does not survive either way).
__device__ __forceinline__ float div6(float s) {
    return lw::quot_rn_signed_zero(s, 6.0f, 1.0f / 6.0f);
}
__device__ __forceinline__ vec3 rk4_average(vec3 k1, vec3 k2, vec3 k3, vec3 k4) {
    vec3 s = ((k1 + 2.0f * k2) + 2.0f * k3) + k4;
    return vec3f(div6(s.x), div6(s.y), div6(s.z));
}

__device__ __forceinline__ unsigned unorm8(float x) {
    x = fminf(fmaxf(x, 0.0f), 1.0f);  // NaN -> 0
    return (unsigned)__builtin_rintf(x * 255.0f);
}
__device__ __forceinline__ unsigned pack_rgba8(float r, float g, float b, float a) {
    return unorm8(r) | (unorm8(g) << 8) | (unorm8(b) << 16) | (unorm8(a) << 24);
}


// Camera ray of invocation (gx, gy), src/main_view.wgsl:347-366.
struct CameraFrame {
    vec3 origin, vd, nh, nv;
    float vw, vh;
};
__device__ __forceinline__ CameraFrame camera_frame(const CameraParams& P) {
    CameraFrame c;
    c.origin = ld3(P.origin);
    c.vd = ld3(P.view_dir);
    vec3 up = ld3(P.up);
    float width = (float)P.width, height = (float)P.height;
    float aspect = width / height;
    c.vh = P.viewport_h;
    c.vw = aspect * c.vh;
    vec3 w = -c.vd;
    vec3 horizontal = rn::cross(up, w);
    vec3 vertical = rn::cross(w, horizontal);
    c.nh = rn::normalize(horizontal);
    c.nv = rn::normalize(vertical);
    return c;
}
__device__ __forceinline__ vec3 camera_dir(const CameraFrame& c, const CameraParams& P, int gx, int gy) {
    float width = (float)P.width, height = (float)P.height;
    float u = rn::sub(rn::mul((float)gx / (width - 1.0f), c.vw), rn::mul(0.5f, c.vw));
    float v = rn::sub(rn::mul((float)gy / (height - 1.0f), c.vh), rn::mul(0.5f, c.vh));
    vec3 s = rn::add(rn::add(rn::mul(u, c.nh), rn::mul(v, c.nv)), rn::mul(0.035f, c.vd));
    return rn::normalize(s);
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long x) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// XCD-aware tile order: the dispatcher hands consecutive workgroups to the 8 XCDs in turn;
// remap so that each XCD walks one contiguous run of tiles (neighbouring tiles touch the
// same grid cells and faces, which then stay in that XCD's L2).
__device__ __forceinline__ int xcd_tile(int b, int n) {
    int q = n / LINON_NUM_XCD, r = n % LINON_NUM_XCD;
    int x = b % LINON_NUM_XCD, i = b / LINON_NUM_XCD;
    return x * q + min(x, r) + i;
}

// The first RK4 stage, f(p, p, v0, v, t): through the field's first_stage() where it defines
// one (the same result with less work), else through the field itself.  The int argument
// prefers the first overload wherever it is well-formed.
template <class F>
__device__ __forceinline__ auto rk4_stage1(const F& f, vec3 p, vec3 v0, vec3 v, float t, int)
    -> decltype(f.first_stage(p, v0, v, t)) {
    return f.first_stage(p, v0, v, t);
}
template <class F>
__device__ __forceinline__ vec3 rk4_stage1(const F& f, vec3 p, vec3 v0, vec3 v, float t, long) {
    return f(p, p, v0, v, t);
}

// The 8 sample rays (src/main_view.wgsl:281-335,470-479): fixed h = 0.001, no scene queries,
// the position before every stride-th step recorded as {(p, 1), (colour, 0.5)}.
template <class F>
__device__ void sample_rays(const RaycastParams& P, const CameraFrame& cam, int r) {
    const F field{};
    float ox, oy, cr, cg, cb;
    switch (r) {
        case 0: ox = -1; oy = -1; cr = 1.0f; cg = 0.0f; cb = 0.0f; break;
        case 1: ox = 0;  oy = -1; cr = 0.5f; cg = 0.5f; cb = 0.0f; break;
        case 2: ox = 1;  oy = -1; cr = 0.0f; cg = 1.0f; cb = 0.0f; break;
        case 3: ox = 1;  oy = 0;  cr = 0.5f; cg = 1.0f; cb = 0.0f; break;
        case 4: ox = 1;  oy = 1;  cr = 1.0f; cg = 1.0f; cb = 0.0f; break;
        case 5: ox = 0;  oy = 1;  cr = 0.5f; cg = 0.5f; cb = 0.5f; break;
        case 6: ox = -1; oy = 1;  cr = 0.0f; cg = 0.0f; cb = 1.0f; break;
        default: ox = -1; oy = 0; cr = 0.5f; cg = 0.0f; cb = 0.5f; break;
    }
    float px, py;
    if (P.outline) {  // src/main_view.wgsl:381-468: pixel -> normalised position
        int x, y;
        outline_decode(P.outline[r], r, P.cam.width, P.cam.height, &x, &y);
        const float width = (float)P.cam.width, height = (float)P.cam.height;
        px = (float)x / (width - 1.0f);
        py = (height - 1.0f - (float)y) / (height - 1.0f);
    } else {
        px = P.mouse_x + 0.01f * ox;
        py = P.mouse_y + 0.01f * oy;
    }
    float u2 = rn::sub(rn::mul(px, cam.vw), rn::mul(0.5f, cam.vw));
    float v2 = rn::sub(rn::mul(py, cam.vh), rn::mul(0.5f, cam.vh));
    vec3 s2 = rn::add(rn::add(rn::mul(u2, cam.nh), rn::mul(v2, cam.nv)), rn::mul(0.035f, cam.vd));
    const vec3 start_dir = rn::normalize(s2);
    vec3 p = cam.origin, d = start_dir;
    const float w = P.step.field_weight;
    const float h = 0.001f, hh = rn::mul(0.5f, h);
    float t = 0.0f;
    const int nsamp = P.samples_per_ray;
    const int stride = P.sample_steps / nsamp;
    float4* out = reinterpret_cast<float4*>(P.ray_samples);
    for (int i = 0; i < P.sample_steps; ++i) {
        vec3 k1 = rk4_stage1(field, p, start_dir, d, t, 0);
        vec3 k2 = field(p, rn::add(p, rn::mul(hh, k1)), start_dir, k1, rn::add(t, hh));
        vec3 k3 = field(p, rn::add(p, rn::mul(hh, k2)), start_dir, k2, rn::add(t, hh));
        vec3 k4 = field(p, rn::add(p, rn::mul(h, k3)), start_dir, k3, rn::add(t, h));
        vec3 V = rk4_average(k1, k2, k3, k4);
        d = rn::add(rn::mul(rn::sub(1.0f, w), d), rn::mul(w, V));
        vec3 step = rn::mul(d, h);
        if (i % stride == 0 && i / stride < nsamp) {
            int idx = r * nsamp + i / stride;
            out[2 * idx] = make_float4(p.x, p.y, p.z, 1.0f);
            out[2 * idx + 1] = make_float4(cr, cg, cb, 0.5f);
        }
        p = rn::add(p, step);
        t = rn::add(t, h);
    }
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

// The 8 outline searches of src/main_view.wgsl:381-468 in one pass: one lane per (x, y) with
// x < W and y < max(W, H) (search 1 runs y from W - 1 down, beyond the H rows when W > H),
// coalesced along x; a wave maximum per search and one 64-bit atomic maximum per wave.
// Entries the reference would read past the end of the W*H buffer read as 0 here (WGSL leaves
// such reads to the implementation).  NaN compares false, as in WGSL.
__device__ __forceinline__ void outline_search_body(const float* exps, int W, int H, unsigned long long* keys) {
    const int Y = W > H ? W : H;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = t < (size_t)W * Y;
    const int x = in ? (int)(t % (size_t)W) : 0, y = in ? (int)(t / (size_t)W) : 0;
    const size_t idx = (size_t)y * H + x;
    const float e = (in && idx < (size_t)W * H) ? exps[idx] : 0.0f;
    const bool q = in && e >= LINON_OUTLINE_MIN_EXP;
    const unsigned long long F = LINON_OUTLINE_FOUND;
    unsigned long long k[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (q && y < H) {
        const unsigned long long s = (unsigned long long)x * H + y;  // x-major scan index
        k[0] = F | ((unsigned long long)((W - x) + y) << 32) | s;
        k[2] = F | ((unsigned long long)(x + y) << 32) | s;
        k[4] = F | ((unsigned long long)(x + (H - y)) << 32) | s;
        k[6] = F | ((unsigned long long)((W - x) + (H - y)) << 32) | s;
        k[3] = F | (0xffffffffull - ((unsigned long long)(W - 1 - x) * H + y));
        k[5] = F | (0xffffffffull - ((unsigned long long)y * W + x));
        k[7] = F | (0xffffffffull - s);
    }
    if (q && y < W) k[1] = F | (0xffffffffull - ((unsigned long long)(W - 1 - y) * W + x));
    const int lane = threadIdx.x & 63;
    for (int r = 0; r < 8; ++r) {
        const unsigned long long m = wave_max_u64(k[r]);
        if (lane == 0 && m) atomicMax(keys + r, m);
    }
}

// End point of a ray whose last step [p0, p0 + s] left the sphere |x|^2 = r2: the crossing
// p0 + u s, u the larger root of |s|^2 u^2 + 2 (p0 . s) u + (|p0|^2 - r2) = 0, taken in the form
// without cancellation for the sign of p0 . s and clamped to [0, 1].  A step that started
// outside the sphere (|p0|^2 > r2) keeps its end point p0 + s.
__device__ __forceinline__ vec3 sphere_exit(vec3 p0, vec3 s, float r2) {
    const float c = rn::sub(rn::dot(p0, p0), r2);
    if (!(c <= 0.0f)) return rn::add(p0, s);
    const float a = rn::dot(s, s);
    const float b = rn::dot(p0, s);
    const float root = lw::sqrt_rn(rn::sub(rn::mul(b, b), rn::mul(a, c)));
    float u = b > 0.0f ? (-c) / rn::add(b, root) : rn::sub(root, b) / a;
    u = fminf(fmaxf(u, 0.0f), 1.0f);
    return rn::add(p0, rn::mul(u, s));
}

// One ray's integration state (src/main_view.wgsl:233-242), kept in registers.
struct RayState {
    vec3 p, d, v0, last_v;
    float t, h, last_diff;
    float safe;  // grid path: lower bound on the distance from p to the mesh
    unsigned color;  // rgba8 of the hit's shading (alpha 1), packed at the hit; black before
    int face, iters;  // face >= 0 once the ray has its colour
};

__device__ __forceinline__ void ray_begin(RayState& r, vec3 origin, vec3 dir, const StepParams& I) {
    r.p = origin;
    r.d = dir;
    r.v0 = dir;
    r.last_v = vec3f(0.0f, 0.0f, 0.0f);
    r.t = 0.0f;
    r.h = I.h_initial;
    r.last_diff = -1.0f;
    r.safe = 0.0f;
    r.color = 0xff000000u;
    r.face = -1;
    r.iters = 0;
}

__device__ __forceinline__ bool ray_finished(const RayState& r, const StepParams& I) {
    return r.iters >= I.max_steps || (I.mode != LINON_MODE_ESCAPE && !(r.t <= I.t_max));
}

// Runs at most `budget` loop iterations of nonlinear_ray_color (src/main_view.wgsl:244-271):
// RK4 with the previous stage passed as `v`, the adaptive redo at h = 0.001, the blend with
// field_weight, the segment query until the first hit, and the escape tests of the escape
// mode.  Returns true once the ray is finished.  `lds` holds the faces when the brute-force
// path staged them, else null; the scene queries are counted into `st`.
// WOne: the field weight is 1 (every benchmark configuration).  The blend (1 - w) d + w V then
// is RN(RN(0 d) + V), which one FMA fma(0, d, V) rounds identically (0 d is exact), so the
// step loop saves two of its three operations per component.
// Fixed: 0 reads the adaptive flag and the mode from I; 1 and 2 are the fixed-step loops of
// the reference mode and of the escape mode, compiled without the other branches, so the step
// loop holds fewer uniform flags (fewer scalar registers to spill and reload per step).
// Stop: a predicate on the ray after each step; when it holds the loop returns early, with
// ray_finished's verdict (the approach leg of a surrogate frame stops at the sphere).
enum : int { LK_STEP_ANY = 0, LK_STEP_FIXED_REFERENCE = 1, LK_STEP_FIXED_ESCAPE = 2 };
struct NeverStop {
    __device__ __forceinline__ bool operator()(const RayState&) const { return false; }
};
template <class F, bool WOne, int Fixed, class Stop = NeverStop>
__device__ __forceinline__ bool ray_advance_impl(RayState& r, const StepParams& I, const SceneParams& S,
                                                 const float4* lds, int budget, QueryStats& st,
                                                 Stop stop = Stop{}) {
    const F field{};
    const bool adaptive = Fixed == LK_STEP_ANY ? I.adaptive != 0 : false;
    const bool escape = Fixed == LK_STEP_ANY ? (I.mode == LINON_MODE_ESCAPE) : Fixed == LK_STEP_FIXED_ESCAPE;
    const bool brute = (S.accel == LINON_ACCEL_BRUTE_K);
    const float margin = S.skip_margin;
    const float w = I.field_weight;
    const vec3 start_dir = r.v0;
    for (int k = 0; k < budget; ++k) {
        if (r.iters >= I.max_steps) return true;
        if (!escape && !(r.t <= I.t_max)) return true;
        ++r.iters;
        const float h = r.h;
        const float hh = rn::mul(0.5f, h);
        const vec3 p = r.p;
        // the field sees opaque copies of the ray state the integrator itself computes with
        // (position and direction; see opaque())
        const vec3 pf = opaque(p);
        vec3 k1 = rk4_stage1(field, pf, start_dir, opaque(r.d), r.t, 0);
        // stage points p + a k as one FMA per component: a contraction a WGSL driver may make of
        // src/main_view.wgsl's p + h / 2 * k (as the oracle's FMA-contracted build lets its
        // compiler do), written out so that every kernel advancing a ray rounds it the same way
        vec3 k2 = field(pf, rn::fma(hh, k1, p), start_dir, k1, rn::add(r.t, hh));
        vec3 k3 = field(pf, rn::fma(hh, k2, p), start_dir, k2, rn::add(r.t, hh));
        vec3 k4 = field(pf, rn::fma(h, k3, p), start_dir, k3, rn::add(r.t, h));
        vec3 V = rk4_average(k1, k2, k3, k4);
        float diff = 0.0f;
        if (adaptive) {  // src/main_view.wgsl:251-255
            diff = rn::length(rn::sub(V, r.last_v));
            if (r.last_diff >= 0.0f && diff > rn::mul(10.0f, r.last_diff) && h > 0.002f) {
                r.h = 0.001f;
                continue;
            }
        }
        if (WOne)
            r.d = vec3f(__builtin_fmaf(0.0f, r.d.x, V.x), __builtin_fmaf(0.0f, r.d.y, V.y),
                        __builtin_fmaf(0.0f, r.d.z, V.z));
        else
            r.d = rn::add(rn::mul(rn::sub(1.0f, w), r.d), rn::mul(w, V));
        const vec3 step = rn::mul(r.d, h);
        if (r.face < 0) {
            const float L2 = rn::dot(step, step);
            // normalize(0) or NaN: no hit (L2 > 0 and finite exactly when |step| rounds so)
            bool query = (L2 > 0.0f && L2 < __builtin_inff());
            if (query && !brute) {
                // The skip test and the distance bound need only an upper bound of the segment
                // length.  The IEEE build takes it from the hardware square root (within a few
                // ulp) scaled up by 2^-20, which is at least the correctly rounded length;
                // the exact length is formed only for a query.
#ifdef LINON_FAST_MATH
                const float L_hi = rn::length(step);
#else
                const float L_hi = __builtin_amdgcn_sqrtf(L2) * (1.0f + 0x1p-20f);
#endif
                if (!(r.safe > L_hi + margin)) r.safe = grid_safe_distance(S, p);
                query = !(r.safe > L_hi + margin);
                st.skips += query ? 0u : 1u;
                r.safe -= L_hi;
            }
            if (query) {
                const float L = rn::length(step);
                st.queries += 1;
                const vec3 nd = step / L;  // normalize(step_dir)
                float best_t = -1.0f;
                int best_face = -1;
                if (brute) {
                    const int nf = S.num_faces;
                    st.tests += (unsigned long long)nf;
                    if (lds) {
                        for (int f = 0; f < nf; ++f) {
                            float4 a = lds[3 * f], e1 = lds[3 * f + 1], e2 = lds[3 * f + 2];
                            float tt = mt_test(vec3f(a.x, a.y, a.z), vec3f(e1.x, e1.y, e1.z),
                                               vec3f(e2.x, e2.y, e2.z), p, nd);
                            consider(tt, f, L, best_t, best_face);
                        }
                    } else {
                        for (int f = 0; f < nf; ++f) {
                            float tt = mt_test(ld_tri_vec(S.tris, f, 0), ld_tri_vec(S.tris, f, 1),
                                               ld_tri_vec(S.tris, f, 2), p, nd);
                            consider(tt, f, L, best_t, best_face);
                        }
                    }
                } else {
                    query_grid(S, p, nd, L, best_t, best_face, st);
                }
                if (best_face >= 0) {  // shading of ray_color, src/main_view.wgsl:199-218
                    r.face = best_face;
                    vec3 e1 = ld_tri_vec(S.tris, best_face, 1), e2 = ld_tri_vec(S.tris, best_face, 2);
                    vec3 n = normalize(cross(e1, e2));
                    float intensity = fmaxf(dot(n, -nd), 0.0f);
                    if (intensity == 0.0f) intensity = fmaxf(dot(-n, -nd), 0.0f);
                    float spec = powf(intensity, 64.0f);
                    float s = (0.01f + intensity) + spec;
                    r.color = pack_rgba8(s * fabsf(n.x), s * fabsf(n.y), s * fabsf(n.z), 1.0f);
                }
            }
        }
        r.p = rn::add(p, step);
        r.t = rn::add(r.t, h);
        if (adaptive) {  // src/main_view.wgsl:266-270
            r.last_v = V;
            if (rn::mul(2.0f, diff) < r.last_diff && h < I.h_initial) r.h = I.h_initial;
            r.last_diff = diff;
        }
        if (escape) {
            if (r.face >= 0 && !I.need_mapping) return true;
            if (rn::dot(r.p, r.p) > I.escape_r2 && rn::dot(r.p, r.d) > 0.0f) {
                if (I.clip_exit) r.p = sphere_exit(p, step, I.escape_r2);
                return true;
            }
        }
        if (stop(r)) return ray_finished(r, I);
    }
    return ray_finished(r, I);
}

template <class F, class Stop = NeverStop>
__device__ __forceinline__ bool ray_advance(RayState& r, const StepParams& I, const SceneParams& S,
                                            const float4* lds, int budget, QueryStats& st, Stop stop = Stop{}) {
    if (I.field_weight == 1.0f) {
        if (I.adaptive == 0 && I.mode == LINON_MODE_REF)
            return ray_advance_impl<F, true, LK_STEP_FIXED_REFERENCE>(r, I, S, lds, budget, st, stop);
        if (I.adaptive == 0 && I.mode == LINON_MODE_ESCAPE)
            return ray_advance_impl<F, true, LK_STEP_FIXED_ESCAPE>(r, I, S, lds, budget, st, stop);
        return ray_advance_impl<F, true, LK_STEP_ANY>(r, I, S, lds, budget, st, stop);
    }
    return ray_advance_impl<F, false, LK_STEP_ANY>(r, I, S, lds, budget, st, stop);
}

// Colour (alpha forced to 1, black when nothing was hit) and end point of a finished ray
// (src/main_view.wgsl:273-277, 372-374).
__device__ __forceinline__ void ray_store(const RayState& r, size_t texel, unsigned* color, float* mapping,
                                          int* hit_face, int* steps) {
    color[texel] = r.color;
    reinterpret_cast<float4*>(mapping)[texel] = make_float4(r.p.x, r.p.y, r.p.z, 1.0f);
    if (hit_face) hit_face[texel] = r.face;
    if (steps) steps[texel] = r.iters;
}

__device__ __forceinline__ void count_wave(unsigned long long* counters, int lane, unsigned long long steps,
                                           const QueryStats& st, unsigned long long rays, unsigned long long refills) {
    unsigned long long c_steps = wave_sum(steps);
    unsigned long long c_tests = wave_sum(st.tests);
    unsigned long long c_cells = wave_sum((unsigned long long)st.cells);
    unsigned long long c_q = wave_sum((unsigned long long)st.queries);
    unsigned long long c_rays = wave_sum(rays);
    unsigned long long c_skip = wave_sum((unsigned long long)st.skips);
    if (counters && lane == 0) {
        atomicAdd(counters + LINON_CNT_STEPS, c_steps);
        atomicAdd(counters + LINON_CNT_TESTS, c_tests);
        atomicAdd(counters + LINON_CNT_CELLS, c_cells);
        atomicAdd(counters + LINON_CNT_QUERIES, c_q);
        atomicAdd(counters + LINON_CNT_RAYS, c_rays);
        atomicAdd(counters + LINON_CNT_SKIPS, c_skip);
        if (refills) atomicAdd(counters + LINON_CNT_REFILLS, refills);
    }
}

template <class F>
__device__ __forceinline__ void raycast_body(const RaycastParams& P) {
    extern __shared__ float4 lds_tris[];
    const CameraFrame cam = camera_frame(P.cam);
    int b = blockIdx.x;
    const int lane = threadIdx.x;
    if (P.sample_block) {
        if (b == 0) {
            if (lane < 8 && P.ray_samples && P.samples_per_ray > 0) sample_rays<F>(P, cam, lane);
            return;
        }
        b -= 1;
    }
    const int ntiles = P.tiles_x * P.tiles_y;
    const int tile = xcd_tile(b, ntiles);
    const int gx = (tile % P.tiles_x) * LINON_TILE + (lane & (LINON_TILE - 1));
    const int gy = (tile / P.tiles_x) * LINON_TILE + (lane / LINON_TILE);

    const bool brute_lds = P.scene.accel == LINON_ACCEL_BRUTE_K && P.scene.num_faces <= LINON_MAX_BRUTE_LDS;
    if (brute_lds) {  // stage the faces once per wave; every lane then reads them from LDS
        const float4* src = reinterpret_cast<const float4*>(P.scene.tris);
        for (int i = lane; i < 3 * P.scene.num_faces; i += blockDim.x) lds_tris[i] = src[i];
        __syncthreads();
    }
    // lanes outside the image stay in the wave (zero work) so the wave reductions below see
    // every lane
    const bool active = gx < P.cam.width && gy < P.cam.height;
    __shared__ QueryStats lds_stats[64];
    lds_stats[lane] = QueryStats{0ull, 0u, 0u, 0u};

    RayState r;
    ray_begin(r, cam.origin, camera_dir(cam, P.cam, gx, gy), P.step);
    if (active) ray_advance<F>(r, P.step, P.scene, brute_lds ? lds_tris : nullptr, 0x7fffffff, lds_stats[lane]);
    // The pixel again, from the hardware lane count rather than the work-item id, so that no
    // index has to stay in a register (or spill) across the step loop.
    const int l = (int)__lane_id();
    const int x = (tile % P.tiles_x) * LINON_TILE + (l & (LINON_TILE - 1));
    const int y = (tile / P.tiles_x) * LINON_TILE + (l / LINON_TILE);
    const bool stored = x < P.cam.width && y < P.cam.height;
    // row-flipped texel (the host keeps W and H <= 16384, so it fits 32 bits)
    if (stored) ray_store(r, (unsigned)(P.cam.height - 1 - y) * (unsigned)P.cam.width + (unsigned)x, P.color,
                          P.mapping, P.hit_face, P.steps);
    count_wave(P.counters, l, (unsigned long long)r.iters, lds_stats[l], stored ? 1ull : 0ull, 0ull);
}

// src/overlay.wgsl:99-146, one lane per texel.
__device__ __forceinline__ float lyapunov_exponent(const float4* m, int W, int x, int y, int delta) {
    float4 xn = m[(size_t)y * W + x + delta], xp = m[(size_t)y * W + x - delta];
    float4 yn = m[(size_t)(y + delta) * W + x], yp = m[(size_t)(y - delta) * W + x];
    vec3 gx = 0.5f * (vec3f(xn.x, xn.y, xn.z) - vec3f(xp.x, xp.y, xp.z));
    vec3 gy = 0.5f * (vec3f(yn.x, yn.y, yn.z) - vec3f(yp.x, yp.y, yp.z));
    float m00 = dot(gx, gx), m01 = dot(gy, gx), m10 = dot(gx, gy), m11 = dot(gy, gy);
    float a0 = m00 * m11 - m01 * m10;
    float a1 = -(m00 + m11);
    float e0, e1;
    float discrim = a1 * a1 - 4.0f * a0;
    if (discrim >= 0.0f) {
        float root = __builtin_sqrtf(discrim);
        e0 = (-a1 - root) * 0.5f;
        e1 = (-a1 + root) * 0.5f;
    } else {
        float root = __builtin_sqrtf(-discrim);
        e0 = -a1 * 0.5f;
        e1 = root * 0.5f;
    }
    return __builtin_sqrtf(fmaxf(e0, e1));
}

__device__ __forceinline__ float srgb_from_linear(float x) {  // src/overlay.wgsl:148-154
    return x < 0.0031308f ? x / 12.92f : 1.055f * powf(x, 1.0f / 2.4f) - 0.055f;
}

__device__ __forceinline__ void overlay_body(const OverlayParams& P) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= P.width || y >= P.height) return;
    const size_t idx = (size_t)y * P.width + x;
    const unsigned c = P.ray_casting[idx];
    // textureLoad of the rgba8unorm ray-cast texture
    float col[4];
    for (int k = 0; k < 4; ++k) col[k] = (float)((c >> (8 * k)) & 255u) * (1.0f / 255.0f);
    bool blended = false;
    float lam = -1.0f;
    const int pad = 2 * P.delta;
    if (P.overlay_mode == 1 && !(x < pad || y < pad || x >= P.width - pad || y >= P.height - pad)) {
        lam = lyapunov_exponent(reinterpret_cast<const float4*>(P.mapping), P.width, x, y, P.delta);
        float s = expf(P.scaling * lam - 5.0f);
        size_t eidx = (size_t)P.height * y + x;  // row stride = height (src/overlay.wgsl:128)
        if (eidx < (size_t)P.width * P.height) P.exponents[eidx] = s;
        if (s >= 0.8f) {
            float alpha = fminf(1.0f, s) * 0.5f;
            for (int k = 0; k < 4; ++k) col[k] = alpha + (1.0f - alpha) * col[k];
            blended = true;
        }
    }
    unsigned outc = c;
    if (P.srgb) {  // overlay_web stores vec4(srgb_from_linear(color.rgb), color.a)
        outc = pack_rgba8(srgb_from_linear(col[0]), srgb_from_linear(col[1]), srgb_from_linear(col[2]), col[3]);
    } else if (blended) {
        outc = pack_rgba8(col[0], col[1], col[2], col[3]);
    }
    P.final_color[idx] = outc;
    if (P.lyapunov) P.lyapunov[idx] = lam;
}

}  // namespace lk
