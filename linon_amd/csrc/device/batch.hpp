// Persistent multi-view ray casting with wave-level ray compaction.
//
// All rays of a batch of views form one flattened index space (views one after another, each
// in 8x8-tile order).  A fixed grid of waves claims rays from an atomic queue: a wave runs its
// lanes' rays for `compact_every` RK4 iterations, then counts its idle lanes with a ballot and,
// once at least `refill_min` lanes are idle, claims that many consecutive rays with a single
// atomic and hands them to the idle lanes by their rank among the idle lanes (mbcnt-style
// prefix of the ballot).  Uneven path lengths (escape mode, many step sizes and resolutions)
// therefore keep lanes busy instead of idling until the slowest ray of a tile finishes.  The
// per-ray arithmetic is the plain kernel's (ray_begin / ray_advance / ray_store), so a view
// gives the same result through either kernel.
#pragma once
#include "params.hpp"
#include "raycast.hpp"

namespace lk {

// One lane's tallies over the rays it traced (LDS; the launch reserves blockDim.x of them).
struct BatchLaneTally {
    QueryStats q;
    unsigned long long steps;
    unsigned rays, pad;
};

// largest v with views[v].first_ray <= ray
__device__ __forceinline__ int find_view(const BatchView* views, int n, unsigned long long ray) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (views[mid].first_ray <= ray)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

template <class F>
__device__ void batch_body(const BatchParams& B) {
    const int lane = threadIdx.x & 63;
    RayState r;
    ray_begin(r, vec3f(0.0f, 0.0f, 0.0f), vec3f(0.0f, 0.0f, 1.0f), B.step);
    // per lane only the view's step size and field weight; the rest of the loop's parameters
    // are the batch's (wave-uniform, so they stay in scalar registers)
    float lane_h = 1.0f, lane_w = 1.0f;
    bool busy = false;
    int view = 0;
    unsigned texel = 0;  // views are at most 16384 x 16384
    bool exhausted = false;
    unsigned long long refills = 0;
    // per-lane tallies in LDS (blockDim.x entries, sized by the launch), out of the registers
    // the step loop needs
    extern __shared__ BatchLaneTally batch_tally[];
    BatchLaneTally& tally = batch_tally[threadIdx.x];
    QueryStats& st = tally.q;
    tally = BatchLaneTally{QueryStats{0ull, 0u, 0u, 0u}, 0ull, 0u, 0u};
    while (true) {
        // ---- refill idle lanes (wave-uniform decision) ----
        const unsigned long long idle = __ballot(!busy);
        const int n_idle = __popcll(idle);
        if (!exhausted && n_idle > 0 && (n_idle >= B.refill_min || n_idle == 64)) {
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(B.queue, (unsigned long long)n_idle);
            base = __shfl(base, 0, 64);
            ++refills;
            if (base >= B.total_rays) exhausted = true;
            if (!busy) {
                const unsigned long long ray = base + (unsigned long long)__popcll(idle & ((1ull << lane) - 1ull));
                if (ray < B.total_rays) {
                    view = find_view(B.views, B.num_views, ray);
                    const BatchView& V = B.views[view];
                    const unsigned long long local = ray - V.first_ray;
                    const int tile = (int)(local >> 6), l = (int)(local & 63ull);
                    const int gx = (tile % V.tiles_x) * LINON_TILE + (l & (LINON_TILE - 1));
                    const int gy = (tile / V.tiles_x) * LINON_TILE + (l / LINON_TILE);
                    if (gx < V.cam.width && gy < V.cam.height) {
                        const CameraFrame cam = camera_frame(V.cam);
                        lane_h = V.step.h_initial;
                        lane_w = V.step.field_weight;
                        StepParams sp = B.step;
                        sp.h_initial = lane_h;
                        sp.field_weight = lane_w;
                        ray_begin(r, cam.origin, camera_dir(cam, V.cam, gx, gy), sp);
                        texel = (unsigned)(V.cam.height - 1 - gy) * (unsigned)V.cam.width + (unsigned)gx;
                        busy = true;
                    }
                }
            }
        }
        if (__ballot(busy) == 0ull) {
            if (exhausted) break;
            continue;  // every lane idle: the next pass refills all 64
        }
        // ---- advance the lanes' rays ----
        if (busy) {
            StepParams sp = B.step;
            sp.h_initial = lane_h;
            sp.field_weight = lane_w;
            if (ray_advance<F>(r, sp, B.scene, nullptr, B.compact_every, st)) {
                const BatchView& V = B.views[view];
                ray_store(r, texel, V.color, V.mapping, V.hit_face, V.steps);
                tally.steps += (unsigned long long)r.iters;
                tally.rays += 1u;
                busy = false;
            }
        }
    }
    count_wave(B.counters, lane, tally.steps, st, (unsigned long long)tally.rays, lane == 0 ? refills : 0ull);
}

}  // namespace lk
