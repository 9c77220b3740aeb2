// Device-side state behind one linon_context (the compute pass of linon's MainView).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/linon.h"
#include "device/params.hpp"

namespace linon {

// One persistent launch of a batch: the views of one field share a work queue.
struct BatchGroup {
    int field = 0;
    int first_view = 0;  // index into BatchState::d_views
    int num_views = 0;
    unsigned long long total_rays = 0;
};

struct BatchState {
    std::vector<linon_view> views;  // caller's order
    std::vector<size_t> view_px;    // pixel offset of each view in the arenas (+ total)
    size_t total_px = 0;
    std::vector<BatchGroup> groups;
    std::vector<BatchView> h_views; // host copy of d_views
    BatchView* d_views = nullptr;   // grouped by field
    unsigned long long* d_queues = nullptr;
    unsigned* d_color = nullptr;
    float* d_mapping = nullptr;
    int* d_hit = nullptr;
    int* d_steps = nullptr;
    StepParams step{};              // the integrator at linon_set_batch (views add h and weight)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void release();
};

// Learned flow-map surrogate of one field: parameters, optimiser state, training and frame
// buffers (all device memory, owned by the context).
struct Surrogate {
    linon_surrogate_config cfg{};
    bool trained = false;
    int adam_t = 0;
    unsigned frames = 0;
    float* W[5] = {};  // FP32 master weights [K][N]; layer 4 is the output layer
    float* b[5] = {};
    float* mW[5] = {};
    float* vW[5] = {};
    float* mb[5] = {};
    float* vb[5] = {};
    float* gW[5] = {};
    float* gb[5] = {};
    unsigned short* Wt[4] = {};  // BF16 copies [N][K] of the hidden layers
    unsigned short* d_wpack = nullptr;  // the hidden layers in MFMA fragment order (fused inference)
    float* d_freqs = nullptr;
    // training
    float *d_states = nullptr, *d_targets = nullptr, *d_bstates = nullptr, *d_btargets = nullptr;
    float* d_tscaled = nullptr;  // network targets (exit - coarse exit) * scale
    float* d_bfine = nullptr;    // the mini-batch's RK4 exits
    float scale_p = 0.5f, scale_d = 1.0f;  // output scales of the exit point and direction
    unsigned short* d_feat = nullptr;
    unsigned short* d_act[5] = {};
    float* d_z[4] = {};
    float *d_out = nullptr, *d_dout = nullptr, *d_err = nullptr, *d_part = nullptr, *d_loss = nullptr;
    unsigned short* d_dz[2] = {};
    // frame
    int frame_px = 0;
    float* f_entry = nullptr;
    int* f_entered = nullptr;
    float* f_states = nullptr;
    int* f_texel = nullptr;
    int* f_count = nullptr;
    float* f_pred = nullptr;
    int* f_accepted = nullptr;
    int* f_list = nullptr;
    float* f_exits = nullptr;
    float* f_errs = nullptr;
    unsigned short *f_feat = nullptr, *f_act0 = nullptr, *f_act1 = nullptr;
    float* f_out = nullptr;
    int* f_steps = nullptr;
    float* f_base = nullptr;  // coarse flow map of the entered rays
    unsigned char* f_mask = nullptr;  // per texel: the frame kept the surrogate's exit
    linon_surrogate_stats train_stats{};
    linon_surrogate_stats frame_stats{};
    hipEvent_t ev[8] = {};      // frame phase boundaries
    bool frame_timed = false;   // the last frame run from the host recorded them (it was not captured)
    bool has_frame = false;     // a frame has been rendered or captured
    void release();
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    // side streams and events that let a batch's field groups run concurrently (created on
    // first use; the fork and joins are captured into the batch's graph as edges)
    hipStream_t batch_side[3] = {nullptr, nullptr, nullptr};
    hipEvent_t batch_fork = nullptr;
    hipEvent_t batch_join[3] = {nullptr, nullptr, nullptr};
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    bool timed = false;
    // per-render event triples (start, after ray cast, after overlay), a ring of slots
    std::vector<hipEvent_t> ring;
    int ring_next = 0, ring_count = 0;
    std::string err;

    uint32_t width = 0, height = 0;
    linon_camera camera{};
    linon_settings settings{};
    linon_integrator integ{};

    // field: a predefined id, or a kernel compiled at run time from a WGSL body
    int field_id = -1;
    int math = LINON_MATH_AUTO;
    bool wgsl_math_for(int field) const;  // which kernel build a predefined field runs with
    int kernel_field(int field) const;    // and which of its kernel-table entries
    bool jit_wgsl_math = false;           // the build the current WGSL body was compiled as
    hipModule_t jit_module = nullptr;
    hipFunction_t jit_fn = nullptr;
    std::string jit_body;

    // images (row-major, texel row 0 at the top)
    unsigned* d_color = nullptr;  // ray_casting_texture, rgba8
    unsigned* d_final = nullptr;  // texture (overlay output), rgba8
    float* d_mapping = nullptr;   // mapping_texture, rgba32f
    float* d_exponents = nullptr; // exponents_buffer
    float* d_lyap = nullptr;
    int* d_hit = nullptr;
    int* d_steps = nullptr;
    float* d_samples = nullptr;
    size_t samples_bytes = 0;
    unsigned long long* d_counters = nullptr;
    unsigned long long* d_outline = nullptr;  // 8 outline-search keys

    // mesh + grid
    bool have_mesh = false;
    uint32_t num_faces = 0;
    float* d_tris = nullptr;
    unsigned* d_cell_start = nullptr;
    unsigned* d_cell_faces = nullptr;
    unsigned char* d_dist = nullptr;
    float grid_lo[3] = {0, 0, 0}, grid_hi[3] = {0, 0, 0};
    int gdim[3] = {1, 1, 1};
    float cell_size = 1.0f;
    float skip_margin = 0.0f;
    size_t grid_refs = 0;
    int accel_req = LINON_ACCEL_AUTO;
    int grid_res_req = 0;
    int grid_builder = LINON_GRID_BUILD_AUTO;
    bool grid_on_device = false;  // how the current grid was built
    std::vector<float> h_xyz;
    std::vector<uint32_t> h_faces;

    linon_status fail(linon_status s, const std::string& msg);
    // (Re)allocates the image buffers for w x h; on failure none is left and the size is 0.
    linon_status alloc_images(uint32_t w, uint32_t h);
    linon_status alloc_images_at(uint32_t w, uint32_t h);
    linon_status alloc_samples();
    void free_images();
    void free_mesh();
    void release_jit();
    linon_status upload_mesh(const float* xyz, uint32_t nv, const uint32_t* faces, uint32_t nf);
    bool srgb_output = false;  // overlay_web epilogue
    int num_cus = 0;

    BatchState batch;
    linon_batch_tuning tuning{16, 1, 32, 16};
    linon_status batch_prepare(BatchState& B, const linon_view* views, int n);
    linon_status batch_launch(BatchState& B, const linon_batch_tuning& T, hipStream_t s);
    linon_status batch_capture();
    linon_status batch_set_step();  // a prepared batch takes the context's current integrator

    Surrogate sg;
    linon_status surrogate_alloc_params();
    linon_status surrogate_init_params(unsigned long long seed);
    linon_status surrogate_forward(int n, const unsigned short* feat, unsigned short* const act[5],
                                   float* const z[4], hipStream_t s);
    linon_status surrogate_train(const linon_surrogate_config& cfg, linon_surrogate_stats* stats);
    linon_status surrogate_eval(const float* states, const float* targets, const float* fine, int n,
                                std::vector<float>& err);
    linon_status surrogate_frame(hipStream_t s);
    linon_status surrogate_collect_stats();
    linon_status surrogate_save(void* out, size_t capacity, size_t* needed);
    linon_status surrogate_pack();  // refreshes d_wpack from the master weights
    linon_status surrogate_infer(const float* states, int n, float* out, bool fused);
    linon_status surrogate_load(const void* data, size_t bytes);
    StepParams surrogate_step_params() const;
    StepParams surrogate_coarse_step_params() const;

    void fill_scene(SceneParams& S) const;
    void fill_raycast_params(RaycastParams& P, const linon_integrator& ig) const;
    linon_status render_with(const linon_integrator& ig, hipStream_t s);
    hipEvent_t* next_timing_slot();
    linon_status read(const void* src, void* dst, size_t want, size_t bytes, const char* what);
};

void fill_step(StepParams& I, const linon_integrator& ig, float field_weight);
// The kernel build (WGSL accuracy or IEEE) that linon_math value `math` selects for a field.
bool wgsl_math_default(int math, int field);
// The kernel-table id a predefined field runs with under a math choice: under AUTO hot air takes
// its faithful-root variant (LINON_KERNEL_HOT_AIR_FAITHFUL_ROOT) and the lens its variant with
// the reciprocal above the floor (LINON_KERNEL_LENS_RECIP), every other field its own id.
int kernel_field_id(int math, int field);
// Error message of the calling thread's last failed entry point that takes no context
// (linon_last_error(NULL) returns it).
void set_unbound_error(const std::string& msg);
const std::string& unbound_error();
void fill_camera(CameraParams& C, const linon_camera& cam, uint32_t w, uint32_t h);

}  // namespace linon
