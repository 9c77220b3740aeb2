// Batched multi-view rendering: per-view output arenas, one persistent launch per field group
// (device/batch.hpp), the whole batch captured into a hipGraph and replayed, and a small
// autotuner for the persistent launch shape.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/linon.h"
#include "context.hpp"
#include "kernels.hpp"

namespace linon {

namespace {

#define LB_CHECK(ctx, call)                                                                      \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return (ctx)->fail(LINON_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

void BatchState::release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    for (void** p : {(void**)&d_color, (void**)&d_mapping, (void**)&d_hit, (void**)&d_steps, (void**)&d_views,
                     (void**)&d_queues}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    groups.clear();
    views.clear();
    h_views.clear();
    view_px.clear();
    total_px = 0;
}

linon_status Context::batch_prepare(BatchState& B, const linon_view* views, int n) {
    if (!views || n <= 0 || n > 4096) return fail(LINON_ERR_ARGUMENT, "set_batch: need 1..4096 views");
    if (!have_mesh) return fail(LINON_ERR_STATE, "set_batch: no mesh set (linon_set_mesh)");
    for (int i = 0; i < n; ++i) {
        const linon_view& v = views[i];
        if (v.width < 2 || v.height < 2 || v.width > 16384 || v.height > 16384)
            return fail(LINON_ERR_ARGUMENT, "set_batch: view " + std::to_string(i) + " size out of [2, 16384]");
        if (!field_is_predefined(v.field))
            return fail(LINON_ERR_ARGUMENT, "set_batch: view " + std::to_string(i) + " has no predefined field");
        if (!(v.h_initial > 0.0f) || !std::isfinite(v.h_initial))
            return fail(LINON_ERR_ARGUMENT, "set_batch: view " + std::to_string(i) + " needs h_initial > 0");
        if (!(v.camera.view_direction[3] >= 0.0f && v.camera.view_direction[3] < 180.0f))
            return fail(LINON_ERR_ARGUMENT, "set_batch: view " + std::to_string(i) + " field of view out of range");
    }
    B.release();
    B.views.assign(views, views + n);
    // pixel arenas, views in the caller's order
    B.view_px.resize(n + 1, 0);
    for (int i = 0; i < n; ++i) B.view_px[i + 1] = B.view_px[i] + (size_t)views[i].width * views[i].height;
    B.total_px = B.view_px[n];
    LB_CHECK(this, hipSetDevice(device));
    LB_CHECK(this, hipMalloc(&B.d_color, B.total_px * 4));
    LB_CHECK(this, hipMalloc(&B.d_mapping, B.total_px * 16));
    LB_CHECK(this, hipMalloc(&B.d_hit, B.total_px * 4));
    LB_CHECK(this, hipMalloc(&B.d_steps, B.total_px * 4));
    fill_step(B.step, integ, 1.0f);
    // one group (work queue + persistent launch) per field, views keep their relative order
    std::map<int, std::vector<int>> by_field;
    for (int i = 0; i < n; ++i) by_field[views[i].field].push_back(i);
    std::vector<BatchView> hv;
    for (auto& kv : by_field) {
        BatchGroup g;
        g.field = kv.first;
        g.first_view = (int)hv.size();
        unsigned long long ray = 0;
        for (int vi : kv.second) {
            const linon_view& v = views[vi];
            BatchView bv;
            std::memset(&bv, 0, sizeof(bv));
            fill_camera(bv.cam, v.camera, v.width, v.height);
            fill_step(bv.step, integ, v.field_weight);
            bv.step.h_initial = v.h_initial;
            bv.tiles_x = (int)((v.width + LINON_TILE - 1) / LINON_TILE);
            bv.tiles = bv.tiles_x * (int)((v.height + LINON_TILE - 1) / LINON_TILE);
            bv.first_ray = ray;
            ray += 64ull * (unsigned long long)bv.tiles;
            bv.color = B.d_color + B.view_px[vi];
            bv.mapping = B.d_mapping + 4 * B.view_px[vi];
            bv.hit_face = B.d_hit + B.view_px[vi];
            bv.steps = B.d_steps + B.view_px[vi];
            hv.push_back(bv);
        }
        g.num_views = (int)kv.second.size();
        g.total_rays = ray;
        B.groups.push_back(g);
    }
    B.h_views = hv;
    LB_CHECK(this, hipMalloc(&B.d_views, hv.size() * sizeof(BatchView)));
    LB_CHECK(this, hipMemcpy(B.d_views, hv.data(), hv.size() * sizeof(BatchView), hipMemcpyHostToDevice));
    LB_CHECK(this, hipMalloc(&B.d_queues, B.groups.size() * sizeof(unsigned long long)));
    LB_CHECK(this, hipMemset(B.d_queues, 0, B.groups.size() * sizeof(unsigned long long)));
    return LINON_OK;
}

// Queue resets (fill kernels, so the sequence can be captured) on stream s, then one persistent
// launch per field group: the first on s, the others on side streams forked from s and joined
// back into it, so the groups run concurrently and one group's tail (its last, longest rays)
// overlaps the next group's work instead of idling the device.  The groups share nothing but
// the counters (atomics) and their own queues.
linon_status Context::batch_launch(BatchState& B, const linon_batch_tuning& T, hipStream_t s) {
    const int cus = std::max(1, num_cus);
    const int wpb = std::max(1, std::min(4, T.waves_per_block));
    const unsigned blocks = (unsigned)std::max(1, cus * std::max(1, T.waves_per_cu) / wpb);
    LB_CHECK(this, launch_clear_u64(B.d_queues, (int)B.groups.size(), s));
    LB_CHECK(this, launch_clear_u64(d_counters, LINON_NUM_COUNTERS, s));
    const int nside = (int)std::min<size_t>(3, B.groups.size() > 0 ? B.groups.size() - 1 : 0);
    for (int k = 0; k < nside; ++k) {
        if (!batch_side[k]) LB_CHECK(this, hipStreamCreateWithFlags(&batch_side[k], hipStreamNonBlocking));
        if (!batch_join[k]) LB_CHECK(this, hipEventCreateWithFlags(&batch_join[k], hipEventDisableTiming));
    }
    if (nside > 0) {
        if (!batch_fork) LB_CHECK(this, hipEventCreateWithFlags(&batch_fork, hipEventDisableTiming));
        LB_CHECK(this, hipEventRecord(batch_fork, s));
        for (int k = 0; k < nside; ++k) LB_CHECK(this, hipStreamWaitEvent(batch_side[k], batch_fork, 0));
    }
    SceneParams S;
    fill_scene(S);
    for (size_t gi = 0; gi < B.groups.size(); ++gi) {
        const BatchGroup& g = B.groups[gi];
        BatchParams P;
        std::memset(&P, 0, sizeof(P));
        P.views = B.d_views + g.first_view;
        P.queue = B.d_queues + gi;
        P.counters = d_counters;
        P.total_rays = g.total_rays;
        P.num_views = g.num_views;
        P.compact_every = std::max(1, T.compact_every);
        P.refill_min = std::max(1, std::min(64, T.refill_min));
        P.step = B.step;
        P.scene = S;
        // groups 1, 2, 3, 4, ... go to side streams 0, 1, 2, 0, ... (the first stays on s)
        hipStream_t gs = gi == 0 ? s : batch_side[(gi - 1) % (size_t)nside];
        const int kf = kernel_field(g.field);
        LB_CHECK(this, wgsl_math_for(g.field) ? linon_wgslmath::launch_batch(kf, P, blocks, 64u * (unsigned)wpb, gs)
                                               : launch_batch(kf, P, blocks, 64u * (unsigned)wpb, gs));
    }
    for (int k = 0; k < nside; ++k) {
        LB_CHECK(this, hipEventRecord(batch_join[k], batch_side[k]));
        LB_CHECK(this, hipStreamWaitEvent(s, batch_join[k], 0));
    }
    return LINON_OK;
}

// The integration loop of a prepared batch is the context's integrator: after
// linon_set_integrator, every view's step parameters (keeping its own h and field weight) and
// the batch's shared ones are rewritten, and the graph, which holds the latter by value, is
// captured again.  Waits for the device first: a replay in flight reads the views.
linon_status Context::batch_set_step() {
    if (!batch.d_views) return LINON_OK;
    LB_CHECK(this, hipSetDevice(device));
    LB_CHECK(this, hipDeviceSynchronize());
    fill_step(batch.step, integ, 1.0f);
    for (BatchView& bv : batch.h_views) {
        const float h = bv.step.h_initial, w = bv.step.field_weight;
        fill_step(bv.step, integ, w);
        bv.step.h_initial = h;
    }
    LB_CHECK(this, hipMemcpy(batch.d_views, batch.h_views.data(), batch.h_views.size() * sizeof(BatchView),
                             hipMemcpyHostToDevice));
    return batch_capture();
}

linon_status Context::batch_capture() {
    if (batch.exec) (void)hipGraphExecDestroy(batch.exec);
    if (batch.graph) (void)hipGraphDestroy(batch.graph);
    batch.exec = nullptr;
    batch.graph = nullptr;
    LB_CHECK(this, hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    linon_status st = batch_launch(batch, tuning, stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(stream, &g);
    if (st != LINON_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return fail(LINON_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    batch.graph = g;
    LB_CHECK(this, hipGraphInstantiate(&batch.exec, batch.graph, nullptr, nullptr, 0));
    return LINON_OK;
}

}  // namespace linon

using linon::Context;

extern "C" {

linon_status linon_set_batch(linon_context* ctx, const linon_view* views, int num_views) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return LINON_ERR_ARGUMENT;
    c->err.clear();
    if (hipSetDevice(c->device) != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return c->fail(LINON_ERR_HIP, "set_batch: device not ready");
    linon_status st = c->batch_prepare(c->batch, views, num_views);
    if (st == LINON_OK) st = c->batch_capture();
    // argument and state errors leave the previous batch as it was; a HIP error comes after it
    // was released, so nothing half-prepared is left to render or read
    if (st == LINON_ERR_HIP) c->batch.release();
    return st;
}

linon_status linon_render_batch(linon_context* ctx, void* stream) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return LINON_ERR_ARGUMENT;
    c->err.clear();
    if (!c->batch.exec) return c->fail(LINON_ERR_STATE, "render_batch: no batch set (linon_set_batch)");
    hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : c->stream;
    if (hipSetDevice(c->device) != hipSuccess) return c->fail(LINON_ERR_HIP, "render_batch: hipSetDevice");
    hipEvent_t* slot = c->next_timing_slot();
    for (int k = 0; k < 3; ++k) c->ev[k] = slot[k];
    hipError_t e = hipEventRecord(c->ev[0], s);
    if (e == hipSuccess) e = hipGraphLaunch(c->batch.exec, s);
    if (e == hipSuccess) e = hipEventRecord(c->ev[1], s);
    if (e == hipSuccess) e = hipEventRecord(c->ev[2], s);
    if (e != hipSuccess) return c->fail(LINON_ERR_HIP, std::string("render_batch: ") + hipGetErrorString(e));
    c->timed = true;
    return LINON_OK;
}

linon_status linon_read_batch_view(linon_context* ctx, int view, uint8_t* rgba, size_t rgba_bytes, float* mapping,
                                   size_t mapping_bytes, int32_t* hit_faces, size_t hit_bytes, int32_t* steps,
                                   size_t steps_bytes) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return LINON_ERR_ARGUMENT;
    c->err.clear();
    if (view < 0 || view >= (int)c->batch.views.size())
        return c->fail(LINON_ERR_ARGUMENT, "read_batch_view: view index out of range");
    const size_t off = c->batch.view_px[view];
    const size_t n = c->batch.view_px[view + 1] - off;
    linon_status st = LINON_OK;
    if (rgba) st = c->read(c->batch.d_color + off, rgba, n * 4, rgba_bytes, "read_batch_view rgba");
    if (st == LINON_OK && mapping)
        st = c->read(c->batch.d_mapping + 4 * off, mapping, n * 16, mapping_bytes, "read_batch_view mapping");
    if (st == LINON_OK && hit_faces)
        st = c->read(c->batch.d_hit + off, hit_faces, n * 4, hit_bytes, "read_batch_view hit_faces");
    if (st == LINON_OK && steps) st = c->read(c->batch.d_steps + off, steps, n * 4, steps_bytes, "read_batch_view steps");
    return st;
}

linon_status linon_set_batch_tuning(linon_context* ctx, const linon_batch_tuning* t) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return LINON_ERR_ARGUMENT;
    c->err.clear();
    if (!t || t->waves_per_cu < 1 || t->waves_per_cu > 64 || (t->waves_per_block != 1 && t->waves_per_block != 2 &&
                                                              t->waves_per_block != 4) ||
        t->compact_every < 1 || t->compact_every > 100000 || t->refill_min < 1 || t->refill_min > 64)
        return c->fail(LINON_ERR_ARGUMENT, "set_batch_tuning: out of range");
    c->tuning = *t;
    if (c->batch.exec) return c->batch_capture();
    return LINON_OK;
}

linon_status linon_get_batch_tuning(const linon_context* ctx, linon_batch_tuning* t) {
    const Context* c = reinterpret_cast<const Context*>(ctx);
    if (!c || !t) return LINON_ERR_ARGUMENT;
    *t = c->tuning;
    return LINON_OK;
}

linon_status linon_autotune_batch(linon_context* ctx, int downscale, linon_batch_tuning* best, float* best_ms) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return LINON_ERR_ARGUMENT;
    c->err.clear();
    if (c->batch.views.empty()) return c->fail(LINON_ERR_STATE, "autotune_batch: no batch set (linon_set_batch)");
    if (downscale < 1 || downscale > 64) return c->fail(LINON_ERR_ARGUMENT, "autotune_batch: downscale in [1, 64]");
    // proxy batch: same cameras, fields and steps, 1/downscale of the resolution
    std::vector<linon_view> proxy = c->batch.views;
    for (auto& v : proxy) {
        v.width = std::max<uint32_t>(8, v.width / (uint32_t)downscale);
        v.height = std::max<uint32_t>(8, v.height / (uint32_t)downscale);
    }
    linon::BatchState P;
    linon_status st = c->batch_prepare(P, proxy.data(), (int)proxy.size());
    if (st != LINON_OK) return st;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        P.release();
        return c->fail(LINON_ERR_HIP, "autotune_batch: hipEventCreate");
    }
    auto time_on = [&](linon::BatchState& B, const linon_batch_tuning& T, int reps, float& ms) -> linon_status {
        // keep the last of `reps` runs (the first warms caches and the code object)
        for (int rep = 0; rep < reps; ++rep) {
            if (hipEventRecord(e0, c->stream) != hipSuccess) return c->fail(LINON_ERR_HIP, "autotune: event");
            linon_status s2 = c->batch_launch(B, T, c->stream);
            if (s2 != LINON_OK) return s2;
            if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
                return c->fail(LINON_ERR_HIP, "autotune: synchronize");
            if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return c->fail(LINON_ERR_HIP, "autotune: elapsed");
        }
        return LINON_OK;
    };
    std::vector<std::pair<float, linon_batch_tuning>> seen;  // every shape timed on the proxy
    auto time_it = [&](const linon_batch_tuning& T, float& ms) -> linon_status {
        linon_status s2 = time_on(P, T, 2, ms);
        if (s2 == LINON_OK) seen.emplace_back(ms, T);
        return s2;
    };
    linon_batch_tuning cur = c->tuning;
    float cur_ms = 0.0f;
    st = time_it(cur, cur_ms);
    // coordinate descent: each knob in turn over its candidates, keeping the best so far
    const std::vector<int> wpc = {4, 8, 12, 16, 24, 32};
    const std::vector<int> wpb = {1, 2, 4};
    const std::vector<int> ce = {4, 16, 64, 256};
    const std::vector<int> rm = {1, 16, 32, 48};
    for (int pass = 0; pass < 2 && st == LINON_OK; ++pass) {
        for (int knob = 0; knob < 4 && st == LINON_OK; ++knob) {
            const std::vector<int>& cand = knob == 0 ? wpc : knob == 1 ? wpb : knob == 2 ? ce : rm;
            for (int v : cand) {
                linon_batch_tuning T = cur;
                int* field = knob == 0 ? &T.waves_per_cu : knob == 1 ? &T.waves_per_block
                                                         : knob == 2 ? &T.compact_every : &T.refill_min;
                if (*field == v) continue;
                *field = v;
                float ms = 0.0f;
                st = time_it(T, ms);
                if (st != LINON_OK) break;
                if (ms < cur_ms) {
                    cur = T;
                    cur_ms = ms;
                }
            }
        }
    }
    P.release();
    // the proxy ranks shapes only roughly at full size: the five fastest on the proxy are
    // timed once more on the batch itself, and the fastest there is kept
    if (st == LINON_OK && !seen.empty()) {
        std::stable_sort(seen.begin(), seen.end(),
                         [](const auto& a, const auto& b) { return a.first < b.first; });
        auto same = [](const linon_batch_tuning& a, const linon_batch_tuning& b) {
            return a.waves_per_cu == b.waves_per_cu && a.waves_per_block == b.waves_per_block &&
                   a.compact_every == b.compact_every && a.refill_min == b.refill_min;
        };
        std::vector<linon_batch_tuning> finalists;
        for (const auto& e : seen) {
            bool dup = false;
            for (const auto& f : finalists) dup = dup || same(f, e.second);
            if (!dup) finalists.push_back(e.second);
            if (finalists.size() == 5) break;
        }
        float best_full = 0.0f;
        for (size_t i = 0; i < finalists.size() && st == LINON_OK; ++i) {
            float ms = 0.0f;
            st = time_on(c->batch, finalists[i], i == 0 ? 2 : 1, ms);
            if (st == LINON_OK && (i == 0 || ms < best_full)) {
                best_full = ms;
                cur = finalists[i];
                cur_ms = ms;
            }
        }
        // waves per CU depends on the amount of work: the proxy has 1/downscale^2 of the rays,
        // where its tail favours fewer persistent waves than the batch wants, so that knob is
        // also swept on the batch itself around the pick
        const linon_batch_tuning pick = cur;
        for (int v : wpc) {
            if (st != LINON_OK) break;
            if (v == pick.waves_per_cu || v < 16) continue;
            linon_batch_tuning T = pick;
            T.waves_per_cu = v;
            float ms = 0.0f;
            st = time_on(c->batch, T, 1, ms);
            if (st == LINON_OK && ms < cur_ms) {
                cur = T;
                cur_ms = ms;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipDeviceSynchronize();
    if (st != LINON_OK) return st;
    c->tuning = cur;
    if (best) *best = cur;
    if (best_ms) *best_ms = cur_ms;
    return c->batch_capture();
}

}  // extern "C"
