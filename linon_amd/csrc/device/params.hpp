// Launch parameters of the ray-cast and overlay kernels.  Plain C layout: the host fills
// the struct and passes it by value, both to the statically compiled kernels and to kernels
// compiled at run time with hiprtc (which receive it through hipModuleLaunchKernel).
#pragma once

#define LINON_MODE_REF 0
#define LINON_MODE_ESCAPE 1
#define LINON_ACCEL_BRUTE_K 1
#define LINON_ACCEL_GRID_K 2
#define LINON_TILE 8          // 8x8 pixels per wave64, the reference's @workgroup_size(8, 8)
#define LINON_NUM_XCD 8       // MI355X: 256 CUs in 8 accelerator complex dies
#define LINON_MAX_BRUTE_LDS 96  // faces staged in LDS by the brute-force path

// counters[] slots
#define LINON_CNT_STEPS 0
#define LINON_CNT_TESTS 1
#define LINON_CNT_CELLS 2
#define LINON_CNT_QUERIES 3
#define LINON_CNT_RAYS 4
#define LINON_CNT_SKIPS 5
#define LINON_NUM_COUNTERS 8

struct RaycastParams {
    // image
    int width, height;
    int tiles_x, tiles_y;
    int sample_block;  // 1: block 0 traces the 8 sample rays of src/main_view.wgsl:377-480
    int pad0;
    // camera (per-ray generation as src/main_view.wgsl:347-366)
    float origin[3];
    float view_dir[3];
    float up[3];
    float viewport_h;  // 0.025 (reference) or 2*0.035*tan(fov/2)
    // settings
    float field_weight;
    float mouse_x, mouse_y;
    // integrator
    float h_initial, t_max;
    int adaptive, max_steps, mode, need_mapping;
    float escape_r2;
    // scene: per face three float4 (v0, v1-v0, v2-v0), w unused
    const float* tris;
    int num_faces;
    int accel;
    // uniform grid over the scene
    const unsigned* cell_start;  // ncells + 1 offsets into cell_faces
    const unsigned* cell_faces;  // face ids, ascending within a cell
    const unsigned char* dist;   // Chebyshev distance (cells) to the nearest occupied cell, capped
    float grid_lo[3];
    float grid_hi[3];
    float cell_size, inv_cell_size;
    float skip_margin;
    int gdim[3];
    // outputs
    unsigned* color;      // rgba8 per texel
    float* mapping;       // float4 per texel
    int* hit_face;        // optional
    int* steps;           // optional
    unsigned long long* counters;  // LINON_NUM_COUNTERS
    // sample rays
    float* ray_samples;   // 8 * samples_per_ray * 8 floats
    int samples_per_ray, sample_steps;
};

struct OverlayParams {
    int width, height;
    int overlay_mode, delta;
    float scaling;
    int pad;
    const unsigned* ray_casting;
    const float* mapping;
    unsigned* final_color;
    float* exponents;
    float* lyapunov;  // optional
};
