// Launch parameters of the ray-cast, batch and overlay kernels.  Plain C layout: the host fills
// the structs and passes them by value (or, for per-view data of a batch, in a device array),
// both to the statically compiled kernels and to kernels compiled at run time with hiprtc.
#pragma once

#define LINON_MODE_REF 0
#define LINON_MODE_ESCAPE 1
#define LINON_ACCEL_BRUTE_K 1
#define LINON_ACCEL_GRID_K 2
#define LINON_TILE 8          // 8x8 pixels per wave64, the reference's @workgroup_size(8, 8)
#define LINON_NUM_XCD 8       // MI355X: 256 CUs in 8 accelerator complex dies
#define LINON_MAX_BRUTE_LDS 96  // faces staged in LDS by the brute-force path
// kernel-table id (not a public field id) of hot air with the faithful distance root, which
// the math choice AUTO runs hot air with (fields.hpp FieldHotAirT)
#define LINON_KERNEL_HOT_AIR_FAITHFUL_ROOT 21
// kernel-table id of the lens with its quotient above the floor at WGSL's accuracy, which the
// math choice AUTO runs the lens with (fields.hpp FieldLensT)
#define LINON_KERNEL_LENS_RECIP 22

// counters[] slots
#define LINON_CNT_STEPS 0
#define LINON_CNT_TESTS 1
#define LINON_CNT_CELLS 2
#define LINON_CNT_QUERIES 3
#define LINON_CNT_RAYS 4
#define LINON_CNT_SKIPS 5
#define LINON_CNT_REFILLS 6
#define LINON_NUM_COUNTERS 8

// The mesh and its uniform grid.
struct SceneParams {
    const float* tris;           // per face three float4: v0, v1 - v0, v2 - v0 (w unused)
    const unsigned* cell_start;  // ncells + 1 offsets into cell_faces
    const unsigned* cell_faces;  // face ids, ascending within a cell
    const unsigned char* dist;   // Chebyshev distance (cells) to the nearest occupied cell, capped
    int num_faces;
    int accel;
    float grid_lo[3];
    float grid_hi[3];
    float cell_size, inv_cell_size;
    float skip_margin;
    int gdim[3];
    float gdimf[3];  // gdim as floats, so that kernels compare against scalar-register arguments
};

// The integration loop of one view (src/main_view.wgsl:229-279 plus the escape mode).
struct StepParams {
    float h_initial, t_max, field_weight, escape_r2;
    int adaptive, max_steps, mode, need_mapping;
    int clip_exit;  // escape mode: end a ray that leaves the sphere from inside on the sphere
    int pad;
};

// One pinhole view (src/main_view.wgsl:347-366): CameraUniform's vectors and the viewport.
struct CameraParams {
    int width, height;
    float origin[3];
    float view_dir[3];
    float up[3];
    float viewport_h;  // 0.025 (reference) or 2 * 0.035 * tan(fov / 2)
};

struct RaycastParams {
    CameraParams cam;
    StepParams step;
    SceneParams scene;
    int tiles_x, tiles_y;
    int sample_block;  // 1: block 0 traces the 8 sample rays of src/main_view.wgsl:377-480
    int pad0;
    float mouse_x, mouse_y;
    // outputs
    unsigned* color;      // rgba8 per texel
    float* mapping;       // float4 per texel
    int* hit_face;        // optional
    int* steps;           // optional
    unsigned long long* counters;  // LINON_NUM_COUNTERS
    // sample rays
    float* ray_samples;   // 8 * samples_per_ray * 8 floats
    int samples_per_ray, sample_steps;
    const unsigned long long* outline;  // 8 outline-search keys; null: rays around the mouse
};

// Outline-ray search (src/main_view.wgsl:381-468).  Each of the 8 searches scans the exponents
// buffer (row stride H) for entries >= 0.8 and keeps either the last pixel, in x-major scan
// order, that maximises a corner key (searches 0, 2, 4, 6: `key >= sum` lets later ties win)
// or the first pixel of a side scan (1, 3, 5, 7).  Both become one 64-bit maximum: bit 63 =
// found; corner searches put the key in bits 32..62 and the scan index x * H + y in the low
// word; side searches put 0xffffffff - (scan position) in the low word.
#define LINON_OUTLINE_MIN_EXP 0.8f
#define LINON_OUTLINE_FOUND 0x8000000000000000ull

// Texel (x, y) of search r from its key; (0, 0) when nothing was found, as in the reference.
__host__ __device__ inline bool outline_decode(unsigned long long key, int r, int W, int H, int* x, int* y) {
    *x = 0;
    *y = 0;
    if (!(key & LINON_OUTLINE_FOUND)) return false;
    const unsigned lo = (unsigned)(key & 0xffffffffull);
    const unsigned pos = 0xffffffffu - lo;
    const unsigned uw = (unsigned)W, uh = (unsigned)H;
    switch (r) {
        case 1: *y = W - 1 - (int)(pos / uw); *x = (int)(pos % uw); break;  // y from W-1 down, x up
        case 3: *x = W - 1 - (int)(pos / uh); *y = (int)(pos % uh); break;  // x from W-1 down, y up
        case 5: *y = (int)(pos / uw); *x = (int)(pos % uw); break;          // y up, x up
        case 7: *x = (int)(pos / uh); *y = (int)(pos % uh); break;          // x up, y up
        default: *x = (int)(lo / uh); *y = (int)(lo % uh); break;           // corners: x * H + y
    }
    return true;
}

// One view of a batch.  Its rays occupy [first_ray, first_ray + 64 * tiles) of the batch's
// flattened ray index space, in 8x8-tile order, so consecutive indices are neighbours.
struct BatchView {
    CameraParams cam;
    StepParams step;
    unsigned long long first_ray;
    int tiles_x, tiles;
    unsigned* color;
    float* mapping;
    int* hit_face;
    int* steps;
};

struct BatchParams {
    const BatchView* views;
    unsigned long long* queue;     // next unclaimed flattened ray index (atomic)
    unsigned long long* counters;  // LINON_NUM_COUNTERS
    unsigned long long total_rays;
    int num_views;
    int compact_every;  // RK4 iterations a wave runs between refills
    int refill_min;     // refill once at least this many lanes are idle
    int pad;
    StepParams step;    // the integration loop shared by all views; each view's h_initial and
                        // field_weight (views[v].step) replace those two
    SceneParams scene;
};

struct OverlayParams {
    int width, height;
    int overlay_mode, delta;
    float scaling;
    int srgb;  // 1: linear -> sRGB encode of the result (overlay_web, src/overlay.wgsl:148-165)
    const unsigned* ray_casting;
    const float* mapping;
    unsigned* final_color;
    float* exponents;
    float* lyapunov;  // optional
};
