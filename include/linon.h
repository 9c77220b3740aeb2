/*
 * linon.h — C ABI of the MI355X (gfx950) nonlinear ray-cast pass.
 *
 * One context stands in for the compute pass that linon's `MainView` owns
 * (src/main_view.rs:53-79): the ray-cast kernel of src/main_view.wgsl and the Lyapunov
 * overlay of src/overlay.wgsl, with the uniform and storage layouts those shaders bind.
 * Every entry point below names the reference interface it replaces.
 *
 * Conventions
 *  - Plain pointers and sizes only; `stream` arguments are hipStream_t passed as void*
 *    (NULL = the context's own stream).
 *  - Every function returns a linon_status; on error the context keeps a message that
 *    linon_last_error() returns (the reference reports shader errors the same way, as a
 *    string, src/main_view.rs:657-679).  Nothing throws or aborts across the ABI.
 *  - A context is not thread-safe (as MainView is not); rendering is stream-ordered and
 *    asynchronous; the linon_read_* calls synchronise.
 *  - Image buffers are row-major with texel row 0 at the top, exactly as wgpu stores the
 *    reference's textures: the ray of invocation (gx, gy) lands in texel (gx, H-1-gy)
 *    (src/main_view.wgsl:342).
 */
#ifndef LINON_H
#define LINON_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LINON_ABI_VERSION 1

typedef enum linon_status {
    LINON_OK = 0,
    LINON_ERR_ARGUMENT = 1,   /* bad pointer, size or enum value */
    LINON_ERR_COMPILE = 2,    /* field body rejected by the translator or by hiprtc */
    LINON_ERR_HIP = 3,        /* a HIP runtime call failed */
    LINON_ERR_STATE = 4,      /* call order violated (e.g. render before a mesh is set) */
    LINON_ERR_UNSUPPORTED = 5
} linon_status;

/* CameraUniform, src/main_view.rs:12-19 == struct Camera, src/main_view.wgsl:7-12 (112 B).
 * view_direction[3] is 0 in the reference; here a positive value is a vertical field of view
 * in degrees (the pinhole of the benchmark configurations), 0 keeps the reference's fixed
 * 0.025 x aspect viewport at focal length 0.035 (src/main_view.wgsl:355-357). */
typedef struct linon_camera {
    float origin[4];
    float view_direction[4];
    float up[4];
    float view_matrix[16]; /* carried for layout only; the compute pass does not read it */
} linon_camera;

/* Settings, src/main_view.rs:43-51 == struct Settings, src/main_view.wgsl:16-23 (24 B). */
typedef struct linon_settings {
    float field_weight;
    float mouse_pos[2];
    int32_t overlay_mode;             /* 0 = disabled, 1 = Lyapunov exponents */
    int32_t central_difference_delta; /* 1..10 in the reference GUI */
    float lyapunov_scaling;
} linon_settings;

/* Integrator knobs.  The reference fixes them in the WGSL text (h_initial :241, the
 * `t <= 5.0` loop :244, adaptive_sampling :227, the sample-ray trace :287-293) and flips
 * them by string replacement (src/main_view.rs:752-829). */
typedef enum linon_mode {
    LINON_MODE_REFERENCE = 0, /* for (; t <= t_max;) as main_view.wgsl:244, max_steps as a cap */
    LINON_MODE_ESCAPE = 1     /* stop when |p| > escape_radius moving outward, at max_steps,
                                 or at the first hit unless need_mapping is set */
} linon_mode;

typedef struct linon_integrator {
    float h_initial;          /* reference 0.1; "Enhance" renders once at 0.001 */
    float t_max;              /* reference 5.0 */
    int32_t adaptive;         /* reference: always on (App. note: the toggle is a no-op) */
    int32_t max_steps;        /* cap on RK4 iterations per ray (adaptive redos included) */
    int32_t mode;             /* linon_mode */
    float escape_radius;      /* escape mode: bounding-sphere radius */
    int32_t samples_per_ray;  /* ray_samples records per sample ray, reference 100 */
    int32_t sample_steps;     /* RK4 steps of each sample ray, reference 5000 */
    int32_t need_mapping;     /* escape mode: keep integrating after a hit */
    int32_t outline_rays;     /* 1: the sample rays start at the outermost pixels whose exponent
                                 (exponents buffer of the previous render's overlay) is >= 0.8,
                                 one per side/corner (sample_outline_rays, main_view.wgsl:381-468);
                                 0: around the mouse position (:469-472), the reference's build */
    int32_t clip_exit;        /* escape mode, 1: a ray that leaves the sphere from inside ends on it, at the
                                 crossing of its last step with the sphere, instead of at the first point
                                 past it (a smooth function of the ray; the flow map's exit point) */
    int32_t pad;
} linon_integrator;

/* Field ids: 1..9 follow PredefinedFunction (src/functions.rs:2-13); 0 (Custom) has no
 * predefined body.  16 is the shader's stub body `return v;` (src/main_view.wgsl:148).
 * 17..20 are the benchmark fields written with the same helpers. */
typedef enum linon_field {
    LINON_FIELD_CUSTOM = 0,
    LINON_FIELD_MIRAGE_SPHERICAL = 1,
    LINON_FIELD_MIRAGE_PLANE = 2,
    LINON_FIELD_MIRAGE_SPHERICAL_SIGMOID = 3,
    LINON_FIELD_MIRAGE_PLANE_SIGMOID = 4,
    LINON_FIELD_TRANSLATION_X = 5,
    LINON_FIELD_TRANSLATION_Z = 6,
    LINON_FIELD_ROTATION = 7,
    LINON_FIELD_LORENZ = 8,
    LINON_FIELD_ROESSLER = 9,
    LINON_FIELD_IDENTITY = 16,
    LINON_FIELD_SWIRL = 17,    /* rotateY(v0, 0.5 * length(p)) */
    LINON_FIELD_LENS = 18,     /* v - 0.3 * p / max(|p|^3, 1e-3) */
    LINON_FIELD_HOT_AIR = 19,  /* refraction(T(p_prev), T(p), v, normalize(p - (0,-0.5,0))) */
    LINON_FIELD_WARP = 20      /* v0 + 0.3 * (sin 4p.y, 0, sin 4p.x) */
} linon_field;

/* Scene acceleration: the reference tests every face for every step (src/main_view.wgsl:182).
 * AUTO picks the uniform grid (cell lists + distance field) for meshes above a few dozen
 * faces; BRUTE keeps the reference's loop over all faces (faces staged in LDS). */
typedef enum linon_accel {
    LINON_ACCEL_AUTO = 0,
    LINON_ACCEL_BRUTE = 1,
    LINON_ACCEL_GRID = 2
} linon_accel;

typedef struct linon_counters {
    uint64_t rk4_steps;       /* RK4 iterations (4 field evaluations each) */
    uint64_t triangle_tests;  /* Moeller-Trumbore tests */
    uint64_t cell_visits;     /* grid cells walked by segment queries */
    uint64_t queries;         /* segment queries that reached the triangle stage */
    uint64_t skips;           /* pre-hit segments the distance bound proved empty */
    uint64_t rays;
    float last_render_ms;     /* device time of the last linon_render (main + overlay) */
    float last_raycast_ms;    /* device time of the ray-cast kernel alone */
} linon_counters;

/* Device-resident outputs, for zero-copy hand-off (e.g. external-memory interop). */
typedef struct linon_device_buffers {
    void* final_rgba8;     /* W*H*4 B   — MainView::texture (overlay output) */
    void* ray_casting_rgba8; /* W*H*4 B — ray_casting_texture */
    void* mapping_rgba32f; /* W*H*16 B  — mapping_texture */
    void* exponents_f32;   /* W*H*4 B   — exponents_buffer (row stride = H, as overlay.wgsl:128) */
    void* lyapunov_f32;    /* W*H*4 B   — lambda per texel, row stride W (-1 where not computed) */
    void* hit_face_i32;    /* W*H*4 B   — first-hit face index per texel (-1 = none) */
    void* steps_i32;       /* W*H*4 B   — RK4 iterations per texel */
    void* ray_samples;     /* 8*samples_per_ray*32 B — RaySamples (src/main_view.wgsl:55-63) */
    uint32_t width, height;
} linon_device_buffers;

/* Learned flow-map surrogate of one field (no counterpart in the reference).  A Fourier-feature
 * MLP (16 frequencies of the 6-float entry state = 198 features, 4 x 256 SiLU layers, 7 outputs:
 * exit point, exit direction, confidence) maps a ray's entry into the sphere of `radius` to its
 * exit, trained on uniform entry states labelled by the FP32 RK4 kernel (escape mode, no mesh,
 * step h, exits clipped to the sphere).  Hidden layers run on BF16 MFMA with FP32 accumulation;
 * the output layer is FP32.  Errors are max over x, y, z of |exit - exit_RK4| / max(1, |exit_RK4|),
 * as the parity tests measure them. */
typedef struct linon_surrogate_config {
    int32_t field;             /* predefined field id */
    int32_t train_rays;        /* training-set size (hold-out: 1%, at most 65536) */
    int32_t batch;             /* mini-batch rows, a multiple of 64 */
    int32_t steps;             /* Adam steps */
    float lr;                  /* Adam learning rate */
    float h;                   /* RK4 step of the labels and of the frame's RK4 legs */
    int32_t max_steps;         /* RK4 iteration cap per leg */
    float radius;              /* bounding sphere */
    float conf_threshold;      /* frame: rays below it go to RK4 */
    float lyapunov_threshold;  /* frame: lambda (overlay's exponent) above it -> RK4 */
    float tolerance;           /* exit-point error the confidence head certifies */
    float rejudge_fraction;    /* frame: share of accepted rays re-run with RK4 to measure error */
    uint32_t seed;
    float coarse_h;            /* > 0: the network learns the residual of the flow map at step h over
                                  the flow map at this coarser step (a multi-fidelity surrogate: a frame
                                  integrates each entered ray at coarse_h and adds the network's
                                  correction); 0: the network predicts the exit itself */
    int32_t pad[2];
} linon_surrogate_config;

typedef struct linon_surrogate_stats {
    uint64_t train_rays;
    float train_loss, label_seconds, train_seconds;
    float holdout_err_q50, holdout_err_q99, holdout_frac_within_tol;
    uint64_t frame_rays, entered, accepted, conf_fallback, lyapunov_fallback, rejudged;
    float rejudge_err_q50, rejudge_err_q99, rejudge_err_max, rejudge_frac_within_tol;
    /* frame phases: approach RK4; the fused MLP kernel alone (mlp_ms); RK4 fallback; overlay and
     * Lyapunov gate; re-judging (not in frame_ms) */
    float frame_ms, approach_ms, mlp_ms, fallback_ms, overlay_ms, rejudge_ms;
    float coarse_ms;  /* the coarse flow map of the entered rays (frame_ms includes it) */
} linon_surrogate_stats;

typedef struct linon_context linon_context;

/* One view of a batch: a camera, a resolution, a predefined field and a step size.  The
 * reference renders one view per MainView; a batch renders many views (and step sizes, e.g. an
 * "Enhance" pass at h/4) in one persistent launch per field, captured into a hipGraph. */
typedef struct linon_view {
    linon_camera camera;
    uint32_t width, height;
    int32_t field;        /* predefined field id (linon_field) */
    float h_initial;
    float field_weight;
    int32_t pad[3];
} linon_view;

/* Launch shape of the persistent batch kernel, chosen by linon_autotune_batch. */
typedef struct linon_batch_tuning {
    int32_t waves_per_cu;     /* persistent grid = waves_per_cu x compute units */
    int32_t waves_per_block;  /* 1, 2 or 4 wave64 per workgroup */
    int32_t compact_every;    /* RK4 iterations a wave runs between refills */
    int32_t refill_min;       /* idle lanes (of 64) that trigger a refill */
} linon_batch_tuning;

/* MainView::new (src/main_view.rs:82-458): allocates the outputs for a width x height view,
 * sets the reference's initial settings (:133-139), integrator (h 0.1, t 5, adaptive) and
 * field (MirageSphericalSigmoid, :100).  device = HIP device ordinal. */
linon_status linon_create(linon_context** out, int device, uint32_t width, uint32_t height);
void linon_destroy(linon_context* ctx);
/* The context's last error message; with ctx = NULL, the message of the calling thread's last
 * failed call that takes no context (linon_flowmap_noalloc), or "null context". */
const char* linon_last_error(const linon_context* ctx);
int linon_abi_version(void);

/* MainView::resize_texture (src/main_view.rs:552-647).  Waits for all work on the device,
 * then reallocates the image buffers.  If that fails, the context holds no images: renders and
 * reads return LINON_ERR_STATE until a resize succeeds. */
linon_status linon_resize(linon_context* ctx, uint32_t width, uint32_t height);
/* MainView::update_camera (src/main_view.rs:490-494). */
linon_status linon_set_camera(linon_context* ctx, const linon_camera* camera);
/* MainView::update_settings (src/main_view.rs:547-550). */
linon_status linon_set_settings(linon_context* ctx, const linon_settings* settings);
/* The overlay pass's entry point: 0 = overlay_desktop (the default), 1 = overlay_web, which
 * stores the sRGB encoding of the final colour's rgb (src/overlay.wgsl:148-165); the reference
 * picks overlay_web for its wasm32 build (src/main_view.rs:422-426). */
linon_status linon_set_output_srgb(linon_context* ctx, int enable);
/* MainView::update_model (src/main_view.rs:460-488): the vertex/face storage buffers of
 * bind group 1.  Copied to the device; the acceleration structure is built here.  The face
 * index (position in `faces`) is the hit id and the tie-break key. */
linon_status linon_set_mesh(linon_context* ctx, const float* xyz, uint32_t num_vertices,
                            const uint32_t* faces, uint32_t num_faces);
linon_status linon_set_accel(linon_context* ctx, int accel, int grid_resolution);
/* Where linon_set_mesh builds the grid: host threads or device kernels, which produce the same
 * arrays; AUTO builds on the device for meshes over 65536 faces.  Rebuilds a mesh already set.
 * (The reference has no acceleration structure: ray_color loops over all faces,
 * src/main_view.wgsl:177-220.) */
typedef enum linon_grid_builder {
    LINON_GRID_BUILD_AUTO = 0,
    LINON_GRID_BUILD_HOST = 1,
    LINON_GRID_BUILD_DEVICE = 2
} linon_grid_builder;
linon_status linon_set_grid_builder(linon_context* ctx, int builder);
/* The grid of the current mesh (for tests; no counterpart in the reference): dims[3], the
 * ncells + 1 cell offsets, the face ids of every cell, the ncells distance bytes, the packed
 * faces (12 floats each) and the reference count.  Null pointers skip an array; sizes are
 * checked against what the grid holds. */
linon_status linon_read_grid(linon_context* ctx, int32_t* dims, uint32_t* cell_start, size_t cell_start_bytes,
                             uint32_t* cell_faces, size_t cell_faces_bytes, uint8_t* dist, size_t dist_bytes,
                             float* tris, size_t tris_bytes, uint64_t* refs);
/* Arithmetic of the ray-cast kernels.  The reference's precision is the Vulkan driver's, within
 * the bounds WGSL states (division 2.5 ulp, sqrt from inverseSqrt, sin/cos 2^-11 absolute on
 * [-pi, pi], exp 3 + 2|x| ulp).  IEEE: correctly rounded division and sqrt, full-precision
 * transcendentals.  WGSL: hardware division, sqrt and transcendentals within those bounds.
 * AUTO (default): WGSL for the predefined fields whose end points and hit faces measure as
 * accurate under it (translations, Lorenz, Roessler, identity, swirl, warp); IEEE for the
 * refraction fields (the four mirages, hot air), where the eta*cos - sqrt(k) cancellation near
 * grazing incidence amplifies division rounding, for rotation (hardware sin/cos of pi*t up to
 * 5 pi moves hit faces), for lens (rays captured near the r^3 singularity end elsewhere), and
 * for WGSL bodies.  Under AUTO, hot air takes its temperature and normal at WGSL accuracy with
 * refraction correctly rounded, and lens its quotient above the singularity's floor at WGSL
 * accuracy with |p| and the floor exact; IEEE runs both fields exactly.  (No counterpart in
 * the reference; its driver decides.) */
typedef enum linon_math {
    LINON_MATH_AUTO = 0,
    LINON_MATH_IEEE = 1,
    LINON_MATH_WGSL = 2
} linon_math;
linon_status linon_set_math(linon_context* ctx, int math);
/* MainView::reload_shader with a body from PredefinedFunction::to_code (src/functions.rs:33-128). */
linon_status linon_set_field_predefined(linon_context* ctx, int field);
/* MainView::reload_shader(None, body) (src/main_view.rs:649-690): a WGSL field body, spliced
 * into `fn field_function(p_prev, p, v0, v, t) -> vec3<f32>` (with_field_function,
 * :852-861), translated to HIP and compiled with hiprtc.  LINON_ERR_COMPILE + message on a
 * bad body; the previous field stays active. */
linon_status linon_set_field_wgsl(linon_context* ctx, const char* body);
/* The two halves of linon_set_field_wgsl without a device (no counterpart in the reference,
 * where naga validates at create_shader_module, src/main_view.rs:662-665): the HIP C++ a body
 * translates to, and a full hiprtc compile of the ray-cast kernel with it.  On failure the
 * message (translator error or hiprtc log) is written to out / log. */
linon_status linon_wgsl_translate(const char* body, char* out, size_t capacity, size_t* needed);
linon_status linon_wgsl_compile_check(const char* body, char* log, size_t capacity);
linon_status linon_set_integrator(linon_context* ctx, const linon_integrator* integrator);
linon_status linon_get_integrator(const linon_context* ctx, linon_integrator* integrator);

/* MainView::render (src/main_view.rs:831-849): the ray-cast dispatch, then the overlay. */
linon_status linon_render(linon_context* ctx, void* stream);
/* MainView::render_high_accuracy (src/main_view.rs:778-802): one render at h_initial = 0.001. */
linon_status linon_render_high_accuracy(linon_context* ctx, void* stream);
/* MainView::render_outline_rays (src/main_view.rs:804-829): one render whose 8 sample rays
 * start at the outermost pixels with exponent >= 0.8 of the previous render's overlay
 * (main_view.wgsl:381-468), as integrator.outline_rays = 1 does for every render. */
linon_status linon_render_outline_rays(linon_context* ctx, void* stream);
linon_status linon_synchronize(linon_context* ctx);

/* Batches (no counterpart in the reference, which renders one view per MainView::render).
 * linon_set_batch uploads the views, allocates their outputs and captures the batch (queue
 * resets + one persistent launch per field) into a hipGraph; linon_render_batch replays it.
 * The integration loop (mode, t_max, max_steps, adaptive, escape radius, clip_exit) is the
 * context's integrator, and the scene is the context's mesh and acceleration choice, each as
 * set at the time of linon_render_batch: linon_set_integrator, linon_set_mesh, linon_set_accel
 * and linon_set_math re-capture a prepared batch (after waiting for the device).  Sample rays
 * and the overlay are not part of a batch. */
linon_status linon_set_batch(linon_context* ctx, const linon_view* views, int num_views);
linon_status linon_render_batch(linon_context* ctx, void* stream);
linon_status linon_read_batch_view(linon_context* ctx, int view, uint8_t* rgba, size_t rgba_bytes, float* mapping,
                                   size_t mapping_bytes, int32_t* hit_faces, size_t hit_bytes, int32_t* steps,
                                   size_t steps_bytes);
linon_status linon_set_batch_tuning(linon_context* ctx, const linon_batch_tuning* tuning);
linon_status linon_get_batch_tuning(const linon_context* ctx, linon_batch_tuning* tuning);
/* Times candidate launch shapes on a copy of the current batch with every view scaled to
 * 1/`downscale` of its width and height (coordinate descent over waves per CU, waves per
 * block, steps per compaction round and refill threshold), times the five fastest once more
 * on the batch itself, sweeps waves per CU (16 and up) around the fastest there, keeps the
 * best and re-captures the batch with it.  best_ms: the batch's own time with the chosen
 * shape. */
linon_status linon_autotune_batch(linon_context* ctx, int downscale, linon_batch_tuning* best, float* best_ms);

/* Surrogate (no counterpart in the reference).  linon_surrogate_train generates the data set,
 * trains, and reports hold-out errors in `stats`.  linon_render_surrogate renders the current
 * view without the mesh (mapping texture + Lyapunov overlay; the ray-cast colour is black):
 * RK4 to the sphere, MLP for the rays that entered, RK4 for unconfident and high-lambda rays,
 * and RK4 on a sample of the accepted rays to measure the surrogate's error.  The comparable
 * plain frame is linon_render with an escape-mode integrator (need_mapping) and an empty mesh.
 * The frame never waits for the host (its compacted counts stay on the device), so a caller
 * may capture it on `stream` into a HIP graph, with other views' frames, once one frame at this
 * size has run uncaptured.  linon_get_surrogate_stats reads the counts and re-judged errors of
 * the last frame that ran on the device (eager or replayed) and the phase timings of the last
 * frame enqueued outside a capture (zero when the last one was captured). */
linon_status linon_surrogate_train(linon_context* ctx, const linon_surrogate_config* cfg,
                                   linon_surrogate_stats* stats);
linon_status linon_render_surrogate(linon_context* ctx, void* stream);
linon_status linon_get_surrogate_stats(linon_context* ctx, linon_surrogate_stats* out);
/* Per texel (W * H bytes, row 0 at the top): 1 where the last surrogate frame kept the
 * surrogate's exit point, 0 where RK4 provided it or the ray never entered the sphere. */
linon_status linon_read_surrogate_accepted(linon_context* ctx, uint8_t* mask, size_t bytes);
/* Diagnostics: the trained network's raw outputs (exit point and direction in target units, the
 * confidence logit) for n host entry states [n][6] into out [n][7], through the fused inference
 * kernel a frame uses (fused = 1) or the training path's kernels (fused = 0). */
linon_status linon_surrogate_infer(linon_context* ctx, const float* states, int n, float* out, int fused);
/* Surrogate weights as a byte blob (no counterpart in the reference): the configuration, the
 * Fourier frequencies, the FP32 master weights and biases of all five layers and the training
 * statistics (not the optimiser state).  linon_surrogate_save writes it to `out` when `capacity`
 * suffices and always stores its size in *needed (out = NULL only asks for the size);
 * linon_surrogate_load checks the blob (magic, version, layer sizes, field) and installs it in
 * place of any trained surrogate, ready for linon_render_surrogate. */
linon_status linon_surrogate_save(linon_context* ctx, void* out, size_t capacity, size_t* needed);
linon_status linon_surrogate_load(linon_context* ctx, const void* data, size_t bytes);

/* Diagnostics (no counterpart in the reference): the flow map through the bounding sphere that
 * the surrogate learns, from host entry states.  Each state (p, q) of `states` [n][6] starts a ray
 * at p with v0 = current direction = q and integrates it without a mesh under `integrator`, which
 * must be in escape mode (clip_exit as given), with the kernel build `math` selects for the field
 * (linon_math); exits [n][6] receives the end point and the final direction, steps[n] (optional)
 * the RK4 iterations.  Fields: swirl, lens, hot air, warp, identity, spherical sigmoid mirage. */
linon_status linon_flowmap(int device, int field, int math, const linon_integrator* integrator, float field_weight,
                           const float* states, int n, float* exits, int32_t* steps);
/* linon_flowmap without any device allocation (no counterpart in the reference): for callers
 * that cannot allocate device memory, such as tight serving loops and harnesses that forbid
 * hipMalloc.  Same arguments, kernels and results as linon_flowmap, but the states, exits and
 * iteration counts are staged in a fixed device array the library holds per device from the
 * first call on it until the process ends, so n is at most LINON_FLOWMAP_NOALLOC_MAX; larger n
 * is rejected with LINON_ERR_ARGUMENT and a message (linon_last_error(NULL)).  Calls that
 * overlap on one device share that array and race: serialise them. */
#define LINON_FLOWMAP_NOALLOC_MAX 65536
linon_status linon_flowmap_noalloc(int device, int field, int math, const linon_integrator* integrator,
                                   float field_weight, const float* states, int n, float* exits, int32_t* steps);
/* Diagnostics: the reciprocal-based divisions of the IEEE build against the IEEE division, bit
 * for bit: the reciprocal for every mantissa at exponents -126..125 (both signs); the
 * shared-denominator division (lw::div3_rn) for every numerator mantissa at three exponents
 * against 16 structured denominators and on random_pairs random operand pairs (zeros and
 * out-of-range magnitudes included); lw::sigmoid_floor at every float, and hot air's
 * temperature from it; lw::refraction_in_range against lw::refraction on random_pairs random
 * temperatures in its range, directions and normals, and on the same draws the one-temperature
 * form lw::refraction_in_range_same and the refracting fields' first RK4 stage (hot air and the
 * four mirages: first_stage(p, ...) against the field at p_prev = p); the bounded divisions of
 * normalize and the lens field (lw::div3_rn_bounded, lw::div3_rn_lens) at every numerator
 * mantissa of 18 exponents from the subnormals up against 19 denominators, and the lens
 * quotient at its floor s = 0.001 (lw::quot_lens_floor) at every numerator below 2^13 in
 * magnitude; the RK4 average's division by six at every float, apart from its ties of subnormal
 * quotients s = (6 m + 3) 2^-149 with m odd, where it returns the other neighbour, and s = +-inf,
 * where it returns NaN.
 * quotient_mismatches counts all but the reciprocal's.  Writes the mismatch counts and the number of
 * comparisons. */
linon_status linon_selftest_division(int device, uint64_t random_pairs, uint64_t* recip_mismatches,
                                     uint64_t* quotient_mismatches, uint64_t* comparisons);
/* Diagnostics: the IEEE build's square root (lw::sqrt_rn: the hardware root plus a one-ulp
 * correction) against the compiler's correctly rounded square root at every one of the 2^32
 * floats, bit for bit (any NaN matches any NaN), and its form without the rescaling of tiny
 * operands (lw::sqrt_rn_no_tiny) at every float whose magnitude is not in (0, 2^-96). */
linon_status linon_selftest_sqrt(int device, uint64_t* mismatches, uint64_t* comparisons);
/* Diagnostics (no counterpart in the reference): the exponential of hot air's default variant
 * (lw::exp_wgsl, the hardware 2^t of RN(y log2 e)) at every float y in [-87, 87] against the
 * FP64 exponential, in units of the float spacing at the exact result: how many results are off
 * by more than WGSL's bound for exp, 3 + 2 |y| of them, how many floats were compared (every
 * float bit pattern is visited; those outside the range are skipped), and the largest error. */
linon_status linon_selftest_exp(int device, uint64_t* over_bound, uint64_t* comparisons, float* max_ulp);
/* Diagnostics (no counterpart in the reference): the device's measured rate of non-packed FP32
 * FMA (v_fma_f32, 2 FLOP each), every lane running 16 independent chains over all compute
 * units; the attainable FP32 rate of kernels built without packed FP32. */
linon_status linon_measure_fma_rate(int device, double* tflops);
/* Diagnostics: the surrogate's BF16-MFMA GEMM on host buffers (bf16 bit patterns in, fp32
 * out): C = A * B (+ bias; with silu != 0: C = z = A*B + bias and C_bf16 = bf16(silu(z))).
 * A is [M][K] (a_trans = 0) or [K][M]; B is [K][N] (b_trans = 0) or [N][K]; the K range may
 * be split in `splits` chunks (multiples of 32) that are reduced afterwards. */
linon_status linon_selftest_gemm(int device, const uint16_t* A, const uint16_t* B, int M, int N, int K, int a_trans,
                                 int b_trans, int splits, const float* bias, int silu, float* C, uint16_t* C_bf16);

/* Read-back (the reference keeps these on the GPU; sizes are checked against W*H). */
linon_status linon_read_final(linon_context* ctx, uint8_t* rgba, size_t bytes);
linon_status linon_read_ray_casting(linon_context* ctx, uint8_t* rgba, size_t bytes);
linon_status linon_read_mapping(linon_context* ctx, float* xyzw, size_t bytes);
linon_status linon_read_exponents(linon_context* ctx, float* data, size_t bytes);
linon_status linon_read_lyapunov(linon_context* ctx, float* data, size_t bytes);
linon_status linon_read_hit_faces(linon_context* ctx, int32_t* data, size_t bytes);
linon_status linon_read_steps(linon_context* ctx, int32_t* data, size_t bytes);
linon_status linon_read_ray_samples(linon_context* ctx, float* data, size_t bytes);
/* Texel positions the last outline-ray search found (main_view.wgsl:381-468): 8 x {x, y, found}
 * int32, x = column, y = row with row 0 at the top (the overlay's coordinates).  Not found
 * gives {0, 0, 0}; the reference then starts that ray at pixel (0, 0) as well.  The reference
 * keeps the positions internal; this read-back exists for tests. */
linon_status linon_read_outline(linon_context* ctx, int32_t* data, size_t bytes);
linon_status linon_get_counters(linon_context* ctx, linon_counters* out);
/* Device time of every render since the last reset (the newest LINON_TIMING_SLOTS kept), from
 * HIP events recorded on the render's stream: per render the ray-cast kernel and the whole
 * pass (ray cast + overlay), in ms.  Synchronises.  (The reference has no timing:
 * timestamp_writes is None, src/main_view.rs:832-835.) */
#define LINON_TIMING_SLOTS 256
linon_status linon_get_timings(linon_context* ctx, float* raycast_ms, float* render_ms, int capacity,
                               int* count);
linon_status linon_reset_timings(linon_context* ctx);
linon_status linon_get_device_buffers(linon_context* ctx, linon_device_buffers* out);

#ifdef __cplusplus
}
#endif

#endif /* LINON_H */
