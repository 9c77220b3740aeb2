// C-ABI context: the device-side state of linon's MainView compute pass (textures, uniforms,
// mesh buffers, the selected field) and the dispatch of the ray-cast and overlay kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/linon.h"
#include "context.hpp"
#include "grid.hpp"
#include "grid_build.hpp"
#include "kernels.hpp"

static_assert(sizeof(linon_camera) == 112, "CameraUniform is 112 bytes (src/main_view.rs:12-19)");
static_assert(sizeof(linon_settings) == 24, "Settings is 24 bytes (src/main_view.rs:43-51)");
static_assert(sizeof(linon_integrator) == 48, "linon_integrator layout");

namespace linon {

#define LN_CHECK_HIP(ctx, call)                                                           \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            return (ctx)->fail(LINON_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
        }                                                                                 \
    } while (0)

linon_status Context::fail(linon_status s, const std::string& msg) {
    err = msg;
    return s;
}

void Context::free_images() {
    for (void** p : {(void**)&d_color, (void**)&d_final, (void**)&d_mapping, (void**)&d_exponents,
                     (void**)&d_lyap, (void**)&d_hit, (void**)&d_steps}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
}

void Context::free_mesh() {
    // the captured batch graph holds these pointers by value: drop it (a successful upload
    // captures it again), so that it never runs on freed buffers
    if (batch.exec) (void)hipGraphExecDestroy(batch.exec);
    if (batch.graph) (void)hipGraphDestroy(batch.graph);
    batch.exec = nullptr;
    batch.graph = nullptr;
    for (void** p : {(void**)&d_tris, (void**)&d_cell_start, (void**)&d_cell_faces, (void**)&d_dist}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    num_faces = 0;
    have_mesh = false;
}

linon_status Context::alloc_images(uint32_t w, uint32_t h) {
    const linon_status st = alloc_images_at(w, h);
    if (st != LINON_OK) {  // no half-sized set of buffers: render and the reads then refuse
        free_images();
        width = height = 0;
    }
    return st;
}

linon_status Context::alloc_images_at(uint32_t w, uint32_t h) {
    free_images();
    width = w;
    height = h;
    const size_t n = (size_t)w * h;
    LN_CHECK_HIP(this, hipMalloc(&d_color, n * 4));
    LN_CHECK_HIP(this, hipMalloc(&d_final, n * 4));
    LN_CHECK_HIP(this, hipMalloc(&d_mapping, n * 16));
    LN_CHECK_HIP(this, hipMalloc(&d_exponents, n * 4));
    LN_CHECK_HIP(this, hipMalloc(&d_lyap, n * 4));
    LN_CHECK_HIP(this, hipMalloc(&d_hit, n * 4));
    LN_CHECK_HIP(this, hipMalloc(&d_steps, n * 4));
    LN_CHECK_HIP(this, hipMemset(d_color, 0, n * 4));
    LN_CHECK_HIP(this, hipMemset(d_final, 0, n * 4));
    LN_CHECK_HIP(this, hipMemset(d_mapping, 0, n * 16));
    LN_CHECK_HIP(this, hipMemset(d_exponents, 0, n * 4));
    LN_CHECK_HIP(this, hipMemset(d_hit, 0xff, n * 4));
    LN_CHECK_HIP(this, hipMemset(d_steps, 0, n * 4));
    if (d_outline) LN_CHECK_HIP(this, hipMemset(d_outline, 0, 8 * 8));  // positions of the old size
    return LINON_OK;
}

linon_status Context::alloc_samples() {
    const size_t bytes = (size_t)8 * std::max(1, integ.samples_per_ray) * 32;
    if (d_samples && samples_bytes == bytes) return LINON_OK;
    if (d_samples) {
        // renders may still run on a caller's stream: wait for the whole device before freeing
        LN_CHECK_HIP(this, hipSetDevice(device));
        LN_CHECK_HIP(this, hipDeviceSynchronize());
        (void)hipFree(d_samples);
    }
    d_samples = nullptr;
    samples_bytes = 0;
    LN_CHECK_HIP(this, hipMalloc(&d_samples, bytes));
    LN_CHECK_HIP(this, hipMemset(d_samples, 0, bytes));
    samples_bytes = bytes;
    return LINON_OK;
}

void Context::fill_scene(SceneParams& S) const {
    std::memset(&S, 0, sizeof(S));
    S.tris = d_tris;
    S.num_faces = (int)num_faces;
    S.accel = (accel_req == LINON_ACCEL_BRUTE) ? LINON_ACCEL_BRUTE_K : LINON_ACCEL_GRID_K;
    S.cell_start = d_cell_start;
    S.cell_faces = d_cell_faces;
    S.dist = d_dist;
    for (int k = 0; k < 3; ++k) {
        S.grid_lo[k] = grid_lo[k];
        S.grid_hi[k] = grid_hi[k];
        S.gdim[k] = gdim[k];
        S.gdimf[k] = (float)gdim[k];
    }
    S.cell_size = cell_size;
    S.inv_cell_size = 1.0f / cell_size;
    S.skip_margin = skip_margin;
}

void fill_step(StepParams& I, const linon_integrator& ig, float field_weight) {
    I.h_initial = ig.h_initial;
    I.t_max = ig.t_max;
    I.field_weight = field_weight;
    I.escape_r2 = ig.escape_radius * ig.escape_radius;
    I.adaptive = ig.adaptive;
    I.max_steps = ig.max_steps;
    I.mode = ig.mode;
    I.need_mapping = ig.need_mapping;
    I.clip_exit = ig.clip_exit;
}

void fill_camera(CameraParams& C, const linon_camera& cam, uint32_t w, uint32_t h) {
    C.width = (int)w;
    C.height = (int)h;
    for (int k = 0; k < 3; ++k) {
        C.origin[k] = cam.origin[k];
        C.view_dir[k] = cam.view_direction[k];
        C.up[k] = cam.up[k];
    }
    // reference: fixed 0.025 at focal length 0.035 (src/main_view.wgsl:355-356)
    C.viewport_h = cam.view_direction[3] > 0.0f
                       ? 2.0f * 0.035f * (float)std::tan((double)cam.view_direction[3] * 3.141592653589793 / 360.0)
                       : 0.025f;
}

void Context::fill_raycast_params(RaycastParams& P, const linon_integrator& ig) const {
    std::memset(&P, 0, sizeof(P));
    fill_camera(P.cam, camera, width, height);
    fill_step(P.step, ig, settings.field_weight);
    fill_scene(P.scene);
    P.tiles_x = (int)((width + LINON_TILE - 1) / LINON_TILE);
    P.tiles_y = (int)((height + LINON_TILE - 1) / LINON_TILE);
    P.sample_block = (ig.samples_per_ray > 0 && ig.sample_steps > 0) ? 1 : 0;
    P.mouse_x = settings.mouse_pos[0];
    P.mouse_y = settings.mouse_pos[1];
    P.color = d_color;
    P.mapping = d_mapping;
    P.hit_face = d_hit;
    P.steps = d_steps;
    P.counters = d_counters;
    P.ray_samples = d_samples;
    P.samples_per_ray = ig.samples_per_ray;
    P.sample_steps = ig.sample_steps;
}

hipEvent_t* Context::next_timing_slot() {
    if (ring.empty()) {
        ring.assign(3 * LINON_TIMING_SLOTS, nullptr);
        for (auto& e : ring)
            if (hipEventCreate(&e) != hipSuccess) e = nullptr;
    }
    hipEvent_t* slot = &ring[3 * (size_t)ring_next];
    ring_next = (ring_next + 1) % LINON_TIMING_SLOTS;
    ring_count = std::min(ring_count + 1, LINON_TIMING_SLOTS);
    return slot;
}

linon_status Context::render_with(const linon_integrator& ig, hipStream_t s) {
    if (!have_mesh) return fail(LINON_ERR_STATE, "render: no mesh set (linon_set_mesh)");
    if (field_id < 0 && !jit_fn) return fail(LINON_ERR_STATE, "render: no field selected");
    if (!d_color) return fail(LINON_ERR_STATE, "render: no image buffers (the last linon_resize failed)");
    if (ig.samples_per_ray > 0 && ig.sample_steps > 0 &&
        (!d_samples || samples_bytes < (size_t)8 * ig.samples_per_ray * 32))
        return fail(LINON_ERR_STATE, "render: no ray-sample buffer of that size (the last linon_set_integrator failed)");
    LN_CHECK_HIP(this, hipSetDevice(device));
    if (!s) s = stream;
    RaycastParams P;
    fill_raycast_params(P, ig);
    const unsigned blocks = (unsigned)(P.tiles_x * P.tiles_y + P.sample_block);
    const size_t shmem = (P.scene.accel == LINON_ACCEL_BRUTE_K && P.scene.num_faces <= LINON_MAX_BRUTE_LDS)
                             ? (size_t)P.scene.num_faces * 48
                             : 0;
    LN_CHECK_HIP(this, launch_clear_u64(d_counters, LINON_NUM_COUNTERS, s));
    hipEvent_t* slot = next_timing_slot();
    for (int k = 0; k < 3; ++k) ev[k] = slot[k];
    LN_CHECK_HIP(this, hipEventRecord(ev[0], s));
    if (ig.outline_rays && P.sample_block) {
        // src/main_view.wgsl:381-468 reads the exponents the previous render's overlay wrote;
        // this render's overlay runs after the ray cast, so searching first keeps that order
        LN_CHECK_HIP(this, launch_clear_u64(d_outline, 8, s));
        LN_CHECK_HIP(this, launch_outline_search(d_exponents, (int)width, (int)height, d_outline, s));
        P.outline = d_outline;
    }
    if (jit_fn) {
        void* args[] = {&P};
        LN_CHECK_HIP(this, hipModuleLaunchKernel(jit_fn, blocks, 1, 1, 64, 1, 1, (unsigned)shmem, s, args, nullptr));
    } else {
        LN_CHECK_HIP(this, wgsl_math_for(field_id)
                               ? linon_wgslmath::launch_raycast_static(kernel_field(field_id), P, blocks, shmem, s)
                               : launch_raycast_static(kernel_field(field_id), P, blocks, shmem, s));
    }
    LN_CHECK_HIP(this, hipEventRecord(ev[1], s));
    OverlayParams O;
    std::memset(&O, 0, sizeof(O));
    O.width = (int)width;
    O.height = (int)height;
    O.overlay_mode = settings.overlay_mode;
    O.delta = settings.central_difference_delta;
    O.scaling = settings.lyapunov_scaling;
    O.srgb = srgb_output ? 1 : 0;
    O.ray_casting = d_color;
    O.mapping = d_mapping;
    O.final_color = d_final;
    O.exponents = d_exponents;
    O.lyapunov = d_lyap;
    LN_CHECK_HIP(this, launch_overlay(O, s));
    LN_CHECK_HIP(this, hipEventRecord(ev[2], s));
    timed = true;
    return LINON_OK;
}

bool Context::wgsl_math_for(int field) const { return wgsl_math_default(math, field); }
int Context::kernel_field(int field) const { return kernel_field_id(math, field); }

namespace {
thread_local std::string g_unbound_error;
}
void set_unbound_error(const std::string& msg) { g_unbound_error = msg; }
const std::string& unbound_error() { return g_unbound_error; }

int kernel_field_id(int math, int field) {
    if (math != LINON_MATH_AUTO) return field;
    if (field == LINON_FIELD_HOT_AIR) return LINON_KERNEL_HOT_AIR_FAITHFUL_ROOT;
    if (field == LINON_FIELD_LENS) return LINON_KERNEL_LENS_RECIP;
    return field;
}

bool wgsl_math_default(int math, int field) {
    if (math == LINON_MATH_IEEE) return false;
    if (math == LINON_MATH_WGSL) return true;
    switch (field) {  // AUTO
        case LINON_FIELD_TRANSLATION_X: case LINON_FIELD_TRANSLATION_Z: case LINON_FIELD_LORENZ:
        case LINON_FIELD_ROESSLER: case LINON_FIELD_IDENTITY: case LINON_FIELD_SWIRL: case LINON_FIELD_WARP:
            return true;
        default:  // mirages and hot air (refraction), rotation (sin/cos of pi t up to 5 pi),
            return false;  // lens (captured rays near the r^3 singularity)
    }
}

linon_status Context::read(const void* src, void* dst, size_t want, size_t bytes, const char* what) {
    if (!dst) return fail(LINON_ERR_ARGUMENT, std::string(what) + ": null destination");
    if (!src) return fail(LINON_ERR_STATE, std::string(what) + ": no buffer (the last linon_resize failed)");
    if (bytes < want)
        return fail(LINON_ERR_ARGUMENT, std::string(what) + ": buffer too small (" + std::to_string(bytes) +
                                            " < " + std::to_string(want) + ")");
    LN_CHECK_HIP(this, hipSetDevice(device));
    // renders may run on a caller's stream: wait for the whole device
    LN_CHECK_HIP(this, hipDeviceSynchronize());
    LN_CHECK_HIP(this, hipMemcpy(dst, src, want, hipMemcpyDeviceToHost));
    return LINON_OK;
}

linon_status Context::upload_mesh(const float* xyz, uint32_t nv, const uint32_t* faces, uint32_t nf) {
    for (size_t i = 0; i < 3 * (size_t)nf; ++i)
        if (faces[i] >= nv) return fail(LINON_ERR_ARGUMENT, "set_mesh: face index out of range");
    int res = grid_res_req;
    if (res == 0) res = (nf <= 65536) ? 128 : 256;
    const bool on_device =
        grid_builder == LINON_GRID_BUILD_DEVICE || (grid_builder == LINON_GRID_BUILD_AUTO && nf > 65536);
    if (on_device) {
        GridGeom geo;
        std::string e;
        if (!grid_geometry(xyz, nv, faces, nf, res, geo, e)) return fail(LINON_ERR_ARGUMENT, "set_mesh: " + e);
        if (h_xyz.data() != xyz) h_xyz.assign(xyz, xyz + 3 * (size_t)nv);
        if (h_faces.data() != faces) h_faces.assign(faces, faces + 3 * (size_t)nf);
        LN_CHECK_HIP(this, hipSetDevice(device));
        LN_CHECK_HIP(this, hipDeviceSynchronize());
        free_mesh();
        DeviceGrid dg;
        if (!build_grid_device(geo, h_xyz.data(), nv, h_faces.data(), nf, stream, dg, e))
            return fail(LINON_ERR_HIP, "set_mesh: " + e);
        d_tris = dg.tris;
        d_cell_start = dg.cell_start;
        d_cell_faces = dg.cell_faces;
        d_dist = dg.dist;
        for (int k = 0; k < 3; ++k) {
            grid_lo[k] = geo.lo[k];
            grid_hi[k] = geo.hi[k];
            gdim[k] = geo.dim[k];
        }
        cell_size = (float)geo.cs;
        skip_margin = 4.0f * (float)geo.margin;
        grid_refs = dg.refs;
        num_faces = nf;
        have_mesh = true;
        grid_on_device = true;
        return batch.d_views ? batch_capture() : LINON_OK;
    }
    GridBuild g;
    std::string e;
    if (!build_grid(xyz, nv, faces, nf, res, g, e)) return fail(LINON_ERR_ARGUMENT, "set_mesh: " + e);
    std::vector<float> packed;
    pack_faces(xyz, faces, nf, packed);
    // keep host copies: a later linon_set_accel may rebuild the grid
    if (h_xyz.data() != xyz) h_xyz.assign(xyz, xyz + 3 * (size_t)nv);
    if (h_faces.data() != faces) h_faces.assign(faces, faces + 3 * (size_t)nf);
    LN_CHECK_HIP(this, hipSetDevice(device));
    LN_CHECK_HIP(this, hipDeviceSynchronize());
    free_mesh();
    const size_t tri_bytes = std::max<size_t>(packed.size() * 4, 48);
    LN_CHECK_HIP(this, hipMalloc(&d_tris, tri_bytes));
    if (!packed.empty()) LN_CHECK_HIP(this, hipMemcpy(d_tris, packed.data(), packed.size() * 4, hipMemcpyHostToDevice));
    LN_CHECK_HIP(this, hipMalloc(&d_cell_start, g.cell_start.size() * 4));
    LN_CHECK_HIP(this, hipMemcpy(d_cell_start, g.cell_start.data(), g.cell_start.size() * 4, hipMemcpyHostToDevice));
    LN_CHECK_HIP(this, hipMalloc(&d_cell_faces, std::max<size_t>(g.cell_faces.size(), 1) * 4));
    if (!g.cell_faces.empty())
        LN_CHECK_HIP(this, hipMemcpy(d_cell_faces, g.cell_faces.data(), g.cell_faces.size() * 4, hipMemcpyHostToDevice));
    LN_CHECK_HIP(this, hipMalloc(&d_dist, g.dist.size()));
    LN_CHECK_HIP(this, hipMemcpy(d_dist, g.dist.data(), g.dist.size(), hipMemcpyHostToDevice));
    for (int k = 0; k < 3; ++k) {
        grid_lo[k] = g.lo[k];
        grid_hi[k] = g.hi[k];
        gdim[k] = g.dim[k];
    }
    cell_size = g.cell_size;
    skip_margin = 4.0f * g.margin;
    grid_refs = g.cell_faces.size();
    num_faces = nf;
    have_mesh = true;
    grid_on_device = false;
    return batch.d_views ? batch_capture() : LINON_OK;  // a prepared batch renders the new mesh
}

}  // namespace linon

using linon::Context;

#define LN_CTX(ctx)                         \
    if (!(ctx)) return LINON_ERR_ARGUMENT;  \
    (ctx)->err.clear();

extern "C" {

int linon_abi_version(void) { return LINON_ABI_VERSION; }

linon_status linon_create(linon_context** out, int device, uint32_t width, uint32_t height) {
    if (!out) return LINON_ERR_ARGUMENT;
    *out = nullptr;
    if (width < 2 || height < 2 || width > 16384 || height > 16384) return LINON_ERR_ARGUMENT;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return LINON_ERR_HIP;
    Context* c = new (std::nothrow) Context();
    if (!c) return LINON_ERR_HIP;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return LINON_ERR_HIP;
    }
    // MainView::new defaults (src/main_view.rs:100,133-139; src/main_view.wgsl:227,241-244,287-293)
    c->settings = linon_settings{1.0f, {0.5f, 0.6f}, 0, 1, 50.0f};
    c->integ = linon_integrator{0.1f, 5.0f, 1, 100000, LINON_MODE_REFERENCE, 2.0f, 100, 5000, 0, 0, 0, 0};
    c->field_id = LINON_FIELD_MIRAGE_SPHERICAL_SIGMOID;
    // arcball default (src/arcball.rs:33-46 + zoom(-1)): eye = centre + (0,0,2) looking down -z
    std::memset(&c->camera, 0, sizeof(c->camera));
    c->camera.origin[2] = 2.0f;
    c->camera.view_direction[2] = -1.0f;
    c->camera.up[1] = 1.0f;
    for (int i = 0; i < 4; ++i) c->camera.view_matrix[5 * i] = 1.0f;
    if (hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        c->num_cus = 256;
    if (hipMalloc(&c->d_counters, 8 * LINON_NUM_COUNTERS) != hipSuccess ||
        hipMalloc(&c->d_outline, 8 * 8) != hipSuccess) {
        if (c->d_counters) (void)hipFree(c->d_counters);
        delete c;
        return LINON_ERR_HIP;
    }
    (void)hipMemset(c->d_counters, 0, 8 * LINON_NUM_COUNTERS);
    (void)hipMemset(c->d_outline, 0, 8 * 8);
    if (c->alloc_images(width, height) != LINON_OK || c->alloc_samples() != LINON_OK) {
        linon_destroy(reinterpret_cast<linon_context*>(c));
        return LINON_ERR_HIP;
    }
    *out = reinterpret_cast<linon_context*>(c);
    return LINON_OK;
}

void linon_destroy(linon_context* ctx) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return;
    (void)hipSetDevice(c->device);
    // renders may still run on a caller's stream: wait for the whole device before freeing
    (void)hipDeviceSynchronize();
    c->free_images();
    c->free_mesh();
    c->release_jit();
    c->batch.release();
    c->sg.release();
    if (c->d_samples) (void)hipFree(c->d_samples);
    if (c->d_counters) (void)hipFree(c->d_counters);
    if (c->d_outline) (void)hipFree(c->d_outline);
    for (auto& e : c->ring)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : c->batch_join)
        if (e) (void)hipEventDestroy(e);
    if (c->batch_fork) (void)hipEventDestroy(c->batch_fork);
    for (auto& s : c->batch_side)
        if (s) (void)hipStreamDestroy(s);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* linon_last_error(const linon_context* ctx) {
    const Context* c = reinterpret_cast<const Context*>(ctx);
    if (!c) return linon::unbound_error().empty() ? "null context" : linon::unbound_error().c_str();
    return c->err.c_str();
}

linon_status linon_resize(linon_context* ctx, uint32_t width, uint32_t height) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    if (width < 2 || height < 2 || width > 16384 || height > 16384)
        return c->fail(LINON_ERR_ARGUMENT, "resize: width and height must be in [2, 16384]");
    LN_CHECK_HIP(c, hipSetDevice(c->device));
    // renders may still run on a caller's stream: wait for the whole device before freeing
    LN_CHECK_HIP(c, hipDeviceSynchronize());
    return c->alloc_images(width, height);
}

linon_status linon_set_camera(linon_context* ctx, const linon_camera* camera) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    if (!camera) return c->fail(LINON_ERR_ARGUMENT, "set_camera: null camera");
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(camera->origin[k]) || !std::isfinite(camera->view_direction[k]) ||
            !std::isfinite(camera->up[k]))
            return c->fail(LINON_ERR_ARGUMENT, "set_camera: non-finite vector");
    if (!(camera->view_direction[3] >= 0.0f && camera->view_direction[3] < 180.0f))
        return c->fail(LINON_ERR_ARGUMENT, "set_camera: view_direction.w (field of view) must be in [0, 180)");
    c->camera = *camera;
    return LINON_OK;
}

linon_status linon_set_output_srgb(linon_context* ctx, int enable) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    if (enable != 0 && enable != 1) return c->fail(LINON_ERR_ARGUMENT, "set_output_srgb: enable must be 0 or 1");
    c->srgb_output = enable == 1;
    return LINON_OK;
}

linon_status linon_set_settings(linon_context* ctx, const linon_settings* s) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    if (!s) return c->fail(LINON_ERR_ARGUMENT, "set_settings: null settings");
    if (s->overlay_mode < 0 || s->overlay_mode > 1)
        return c->fail(LINON_ERR_ARGUMENT, "set_settings: overlay_mode must be 0 or 1");
    if (s->central_difference_delta < 1 || s->central_difference_delta > 1024)
        return c->fail(LINON_ERR_ARGUMENT, "set_settings: central_difference_delta must be >= 1");
    c->settings = *s;
    return LINON_OK;
}

linon_status linon_set_integrator(linon_context* ctx, const linon_integrator* ig) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    if (!ig) return c->fail(LINON_ERR_ARGUMENT, "set_integrator: null");
    if (!(ig->h_initial > 0.0f) || !std::isfinite(ig->h_initial) || !std::isfinite(ig->t_max))
        return c->fail(LINON_ERR_ARGUMENT, "set_integrator: h_initial must be > 0 and t_max finite");
    if (ig->max_steps < 0 || ig->max_steps > 100000000)
        return c->fail(LINON_ERR_ARGUMENT, "set_integrator: max_steps must be in [0, 1e8]");
    if (ig->mode != LINON_MODE_REFERENCE && ig->mode != LINON_MODE_ESCAPE)
        return c->fail(LINON_ERR_ARGUMENT, "set_integrator: unknown mode");
    if (ig->clip_exit != 0 && ig->clip_exit != 1)
        return c->fail(LINON_ERR_ARGUMENT, "set_integrator: clip_exit must be 0 or 1");
    if (ig->samples_per_ray < 0 || ig->samples_per_ray > 100000 || ig->sample_steps < 0 ||
        (ig->samples_per_ray > 0 && ig->sample_steps < ig->samples_per_ray))
        return c->fail(LINON_ERR_ARGUMENT, "set_integrator: need 0 <= samples_per_ray <= sample_steps");
    c->integ = *ig;
    linon_status st = c->alloc_samples();
    if (st == LINON_OK) st = c->batch_set_step();
    return st;
}

linon_status linon_get_integrator(const linon_context* ctx, linon_integrator* ig) {
    const Context* c = reinterpret_cast<const Context*>(ctx);
    if (!c || !ig) return LINON_ERR_ARGUMENT;
    *ig = c->integ;
    return LINON_OK;
}

linon_status linon_set_accel(linon_context* ctx, int accel, int grid_resolution) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    if (accel < LINON_ACCEL_AUTO || accel > LINON_ACCEL_GRID)
        return c->fail(LINON_ERR_ARGUMENT, "set_accel: unknown accel");
    if (grid_resolution != 0 && (grid_resolution < 4 || grid_resolution > 1024))
        return c->fail(LINON_ERR_ARGUMENT, "set_accel: grid_resolution must be 0 (auto) or in [4, 1024]");
    bool rebuild = c->have_mesh && grid_resolution != c->grid_res_req;
    const bool changed = accel != c->accel_req;
    c->accel_req = accel;
    c->grid_res_req = grid_resolution;
    if (rebuild) return c->upload_mesh(c->h_xyz.data(), (uint32_t)(c->h_xyz.size() / 3), c->h_faces.data(),
                                       (uint32_t)(c->h_faces.size() / 3));
    // the captured batch holds the scene parameters (the accel choice among them) by value
    if (changed && c->batch.exec) return c->batch_capture();
    return LINON_OK;
}

linon_status linon_selftest_division(int device, uint64_t random_pairs, uint64_t* recip_mismatches,
                                     uint64_t* quotient_mismatches, uint64_t* comparisons) {
    if (!recip_mismatches || !quotient_mismatches || !comparisons) return LINON_ERR_ARGUMENT;
    if (hipSetDevice(device) != hipSuccess) return LINON_ERR_HIP;
    unsigned long long r = 0, q = 0, n = 0;
    if (linon::run_division_selftest(random_pairs, &r, &q, &n) != hipSuccess) return LINON_ERR_HIP;
    *recip_mismatches = r;
    *quotient_mismatches = q;
    *comparisons = n;
    return LINON_OK;
}

linon_status linon_selftest_sqrt(int device, uint64_t* mismatches, uint64_t* comparisons) {
    if (!mismatches || !comparisons) return LINON_ERR_ARGUMENT;
    if (hipSetDevice(device) != hipSuccess) return LINON_ERR_HIP;
    unsigned long long m = 0, n = 0;
    if (linon::run_sqrt_selftest(&m, &n) != hipSuccess) return LINON_ERR_HIP;
    *mismatches = m;
    *comparisons = n;
    return LINON_OK;
}

linon_status linon_selftest_exp(int device, uint64_t* over_bound, uint64_t* comparisons, float* max_ulp) {
    if (!over_bound || !comparisons || !max_ulp) return LINON_ERR_ARGUMENT;
    if (hipSetDevice(device) != hipSuccess) return LINON_ERR_HIP;
    unsigned long long o = 0, n = 0;
    float m = 0.0f;
    if (linon::run_exp_selftest(&o, &n, &m) != hipSuccess) return LINON_ERR_HIP;
    *over_bound = o;
    *comparisons = n;
    *max_ulp = m;
    return LINON_OK;
}

linon_status linon_measure_fma_rate(int device, double* tflops) {
    if (!tflops) return LINON_ERR_ARGUMENT;
    if (hipSetDevice(device) != hipSuccess) return LINON_ERR_HIP;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return LINON_ERR_HIP;
    if (linon::run_fma_rate(cus, tflops) != hipSuccess) return LINON_ERR_HIP;
    return LINON_OK;
}

linon_status linon_set_math(linon_context* ctx, int math) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    if (math < LINON_MATH_AUTO || math > LINON_MATH_WGSL) return c->fail(LINON_ERR_ARGUMENT, "set_math: unknown mode");
    c->math = math;
    // a WGSL body is compiled for one build: recompile it for the new one
    if (c->jit_fn && c->jit_wgsl_math != (math == LINON_MATH_WGSL)) {
        const std::string body = c->jit_body;
        return linon_set_field_wgsl(ctx, body.c_str());
    }
    if (c->batch.exec) return c->batch_capture();
    return LINON_OK;
}

linon_status linon_set_grid_builder(linon_context* ctx, int builder) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    if (builder < LINON_GRID_BUILD_AUTO || builder > LINON_GRID_BUILD_DEVICE)
        return c->fail(LINON_ERR_ARGUMENT, "set_grid_builder: unknown builder");
    const bool rebuild = c->have_mesh && builder != c->grid_builder;
    c->grid_builder = builder;
    if (rebuild) return c->upload_mesh(c->h_xyz.data(), (uint32_t)(c->h_xyz.size() / 3), c->h_faces.data(),
                                       (uint32_t)(c->h_faces.size() / 3));
    return LINON_OK;
}

linon_status linon_read_grid(linon_context* ctx, int32_t* dims, uint32_t* cell_start, size_t cell_start_bytes,
                             uint32_t* cell_faces, size_t cell_faces_bytes, uint8_t* dist, size_t dist_bytes,
                             float* tris, size_t tris_bytes, uint64_t* refs) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    if (!c->have_mesh) return c->fail(LINON_ERR_STATE, "read_grid: no mesh set");
    const size_t nc = (size_t)c->gdim[0] * c->gdim[1] * c->gdim[2];
    if (dims)
        for (int k = 0; k < 3; ++k) dims[k] = c->gdim[k];
    if (refs) *refs = (uint64_t)c->grid_refs;
    linon_status st = LINON_OK;
    if (cell_start) st = c->read(c->d_cell_start, cell_start, (nc + 1) * 4, cell_start_bytes, "read_grid cell_start");
    if (st == LINON_OK && cell_faces && c->grid_refs)
        st = c->read(c->d_cell_faces, cell_faces, c->grid_refs * 4, cell_faces_bytes, "read_grid cell_faces");
    if (st == LINON_OK && dist) st = c->read(c->d_dist, dist, nc, dist_bytes, "read_grid dist");
    if (st == LINON_OK && tris && c->num_faces)
        st = c->read(c->d_tris, tris, (size_t)c->num_faces * 48, tris_bytes, "read_grid tris");
    return st;
}

linon_status linon_set_mesh(linon_context* ctx, const float* xyz, uint32_t nv, const uint32_t* faces, uint32_t nf) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    if ((nv && !xyz) || (nf && !faces)) return c->fail(LINON_ERR_ARGUMENT, "set_mesh: null array");
    return c->upload_mesh(xyz, nv, faces, nf);
}

linon_status linon_set_field_predefined(linon_context* ctx, int field) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    if (field == LINON_FIELD_CUSTOM)
        return c->fail(LINON_ERR_ARGUMENT, "Custom has no predefined body; use linon_set_field_wgsl");
    if (!linon::field_is_predefined(field)) return c->fail(LINON_ERR_ARGUMENT, "unknown predefined field id");
    c->release_jit();
    c->field_id = field;
    return LINON_OK;
}

linon_status linon_render(linon_context* ctx, void* stream) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    return c->render_with(c->integ, reinterpret_cast<hipStream_t>(stream));
}

linon_status linon_render_high_accuracy(linon_context* ctx, void* stream) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    linon_integrator ig = c->integ;
    ig.h_initial = 0.001f;  // src/main_view.rs:785-787
    return c->render_with(ig, reinterpret_cast<hipStream_t>(stream));
}

linon_status linon_render_outline_rays(linon_context* ctx, void* stream) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    linon_integrator ig = c->integ;
    ig.outline_rays = 1;  // src/main_view.rs:811-814
    return c->render_with(ig, reinterpret_cast<hipStream_t>(stream));
}

linon_status linon_synchronize(linon_context* ctx) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    (void)hipSetDevice(c->device);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return c->fail(LINON_ERR_HIP, std::string("synchronize: ") + hipGetErrorString(e));
    e = hipDeviceSynchronize();
    if (e != hipSuccess) return c->fail(LINON_ERR_HIP, std::string("synchronize: ") + hipGetErrorString(e));
    return LINON_OK;
}

linon_status linon_read_final(linon_context* ctx, uint8_t* rgba, size_t bytes) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    return c->read(c->d_final, rgba, (size_t)c->width * c->height * 4, bytes, "read_final");
}
linon_status linon_read_ray_casting(linon_context* ctx, uint8_t* rgba, size_t bytes) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    return c->read(c->d_color, rgba, (size_t)c->width * c->height * 4, bytes, "read_ray_casting");
}
linon_status linon_read_mapping(linon_context* ctx, float* xyzw, size_t bytes) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    return c->read(c->d_mapping, xyzw, (size_t)c->width * c->height * 16, bytes, "read_mapping");
}
linon_status linon_read_exponents(linon_context* ctx, float* data, size_t bytes) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    return c->read(c->d_exponents, data, (size_t)c->width * c->height * 4, bytes, "read_exponents");
}
linon_status linon_read_lyapunov(linon_context* ctx, float* data, size_t bytes) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    return c->read(c->d_lyap, data, (size_t)c->width * c->height * 4, bytes, "read_lyapunov");
}
linon_status linon_read_hit_faces(linon_context* ctx, int32_t* data, size_t bytes) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    return c->read(c->d_hit, data, (size_t)c->width * c->height * 4, bytes, "read_hit_faces");
}
linon_status linon_read_steps(linon_context* ctx, int32_t* data, size_t bytes) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    return c->read(c->d_steps, data, (size_t)c->width * c->height * 4, bytes, "read_steps");
}
linon_status linon_read_ray_samples(linon_context* ctx, float* data, size_t bytes) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    size_t want = (size_t)8 * std::max(1, c->integ.samples_per_ray) * 32;
    return c->read(c->d_samples, data, want, bytes, "read_ray_samples");
}

linon_status linon_read_outline(linon_context* ctx, int32_t* data, size_t bytes) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    unsigned long long keys[8];
    linon_status st = c->read(c->d_outline, keys, sizeof(keys), sizeof(keys), "read_outline");
    if (st != LINON_OK) return st;
    if (!data || bytes < 8 * 3 * sizeof(int32_t)) return c->fail(LINON_ERR_ARGUMENT, "read_outline: need 96 bytes");
    for (int r = 0; r < 8; ++r) {
        int x, y;
        const bool found = outline_decode(keys[r], r, (int)c->width, (int)c->height, &x, &y);
        data[3 * r] = x;
        data[3 * r + 1] = y;
        data[3 * r + 2] = found ? 1 : 0;
    }
    return LINON_OK;
}

linon_status linon_get_counters(linon_context* ctx, linon_counters* out) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    if (!out) return c->fail(LINON_ERR_ARGUMENT, "get_counters: null");
    unsigned long long h[LINON_NUM_COUNTERS];
    linon_status s = c->read(c->d_counters, h, sizeof(h), sizeof(h), "get_counters");
    if (s != LINON_OK) return s;
    std::memset(out, 0, sizeof(*out));
    out->rk4_steps = h[LINON_CNT_STEPS];
    out->triangle_tests = h[LINON_CNT_TESTS];
    out->cell_visits = h[LINON_CNT_CELLS];
    out->queries = h[LINON_CNT_QUERIES];
    out->skips = h[LINON_CNT_SKIPS];
    out->rays = h[LINON_CNT_RAYS];
    if (c->timed && c->ev[0]) {
        float a = 0, b = 0;
        if (hipEventElapsedTime(&a, c->ev[0], c->ev[2]) == hipSuccess) out->last_render_ms = a;
        if (hipEventElapsedTime(&b, c->ev[0], c->ev[1]) == hipSuccess) out->last_raycast_ms = b;
    }
    return LINON_OK;
}

linon_status linon_get_timings(linon_context* ctx, float* raycast_ms, float* render_ms, int capacity, int* count) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    if (!count || capacity < 0) return c->fail(LINON_ERR_ARGUMENT, "get_timings: null count");
    if (hipSetDevice(c->device) != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return c->fail(LINON_ERR_HIP, "get_timings: synchronize failed");
    const int n = std::min(c->ring_count, capacity);
    const int first = (c->ring_next - c->ring_count + LINON_TIMING_SLOTS) % LINON_TIMING_SLOTS;
    for (int i = 0; i < n; ++i) {
        hipEvent_t* e = &c->ring[3 * (size_t)((first + i) % LINON_TIMING_SLOTS)];
        float a = -1.0f, b = -1.0f;
        if (hipEventElapsedTime(&a, e[0], e[1]) != hipSuccess) a = -1.0f;
        if (hipEventElapsedTime(&b, e[0], e[2]) != hipSuccess) b = -1.0f;
        if (raycast_ms) raycast_ms[i] = a;
        if (render_ms) render_ms[i] = b;
    }
    *count = n;
    return LINON_OK;
}

linon_status linon_reset_timings(linon_context* ctx) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    c->ring_count = 0;
    return LINON_OK;
}

linon_status linon_get_device_buffers(linon_context* ctx, linon_device_buffers* out) {
    Context* c = reinterpret_cast<Context*>(ctx);
    LN_CTX(c);
    if (!out) return c->fail(LINON_ERR_ARGUMENT, "get_device_buffers: null");
    out->final_rgba8 = c->d_final;
    out->ray_casting_rgba8 = c->d_color;
    out->mapping_rgba32f = c->d_mapping;
    out->exponents_f32 = c->d_exponents;
    out->lyapunov_f32 = c->d_lyap;
    out->hit_face_i32 = c->d_hit;
    out->steps_i32 = c->d_steps;
    out->ray_samples = c->d_samples;
    out->width = c->width;
    out->height = c->height;
    return LINON_OK;
}

}  // extern "C"
