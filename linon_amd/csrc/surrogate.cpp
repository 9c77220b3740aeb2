// Learned flow-map surrogate: host side (training loop, frame pipeline, statistics).
//
// Training follows the usual data-parallel recipe on one device: uniform entry states ->
// labels from the FP32 RK4 flow-map kernel -> mini-batches gathered by a counter-based hash ->
// forward (four BF16-MFMA GEMMs with fused bias+SiLU, FP32 output layer) -> loss (MSE of the
// scaled exit state + logistic loss of the confidence head against "error < tolerance") ->
// backward (FP32 output layer, BF16-MFMA GEMMs for dX with the SiLU' gate fused, split-K GEMMs
// for dW) -> Adam on FP32 master weights, which also refreshes the BF16 [N][K] weight copies.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/linon.h"
#include "context.hpp"
#include "kernels.hpp"
#include "surrogate_kernels.hpp"

namespace linon {

namespace {

#define LS_CHECK(ctx, call)                                                                      \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return (ctx)->fail(LINON_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

const int kLayerK[4] = {LS_FEATS_PAD, LS_HIDDEN, LS_HIDDEN, LS_HIDDEN};

template <class T>
hipError_t dalloc(T*& p, size_t n) {
    if (p) (void)hipFree(p);
    p = nullptr;
    return hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T));
}

float host_bf16_round(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

}  // namespace

void Surrogate::release() {
    for (int l = 0; l < 5; ++l)
        for (float** p : {&W[l], &b[l], &mW[l], &vW[l], &mb[l], &vb[l], &gW[l], &gb[l]}) {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
        }
    for (int l = 0; l < 4; ++l) {
        if (Wt[l]) (void)hipFree(Wt[l]);
        Wt[l] = nullptr;
    }
    void* bufs[] = {d_wpack, d_freqs, d_states, d_targets, d_bstates, d_btargets, d_tscaled, d_bfine, d_feat, d_act[0],
                    d_act[1], d_act[2], d_act[3], d_act[4], d_z[0], d_z[1], d_z[2], d_z[3], d_out, d_dout, d_err,
                    d_dz[0], d_dz[1], d_part, d_loss, f_entry, f_entered, f_states, f_texel, f_count, f_pred,
                    f_accepted, f_list, f_exits, f_errs, f_feat, f_act0, f_act1, f_out, f_steps, f_base, f_mask};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    *this = Surrogate();
}

linon_status Context::surrogate_alloc_params() {
    Surrogate& S = sg;
    for (int l = 0; l < 5; ++l) {
        const int K = l < 4 ? kLayerK[l] : LS_HIDDEN;
        const int N = l < 4 ? LS_HIDDEN : LS_OUT;
        LS_CHECK(this, dalloc(S.W[l], (size_t)K * N));
        LS_CHECK(this, dalloc(S.b[l], N));
        LS_CHECK(this, dalloc(S.mW[l], (size_t)K * N));
        LS_CHECK(this, dalloc(S.vW[l], (size_t)K * N));
        LS_CHECK(this, dalloc(S.mb[l], N));
        LS_CHECK(this, dalloc(S.vb[l], N));
        LS_CHECK(this, dalloc(S.gW[l], (size_t)K * N));
        LS_CHECK(this, dalloc(S.gb[l], N));
        LS_CHECK(this, hipMemset(S.mW[l], 0, (size_t)K * N * 4));
        LS_CHECK(this, hipMemset(S.vW[l], 0, (size_t)K * N * 4));
        LS_CHECK(this, hipMemset(S.mb[l], 0, N * 4));
        LS_CHECK(this, hipMemset(S.vb[l], 0, N * 4));
        if (l < 4) LS_CHECK(this, dalloc(S.Wt[l], (size_t)K * N));
    }
    LS_CHECK(this, dalloc(S.d_freqs, LS_FREQS));
    std::vector<float> fr(LS_FREQS);
    // geometric frequencies from pi/2 to 16 pi over the unit-scaled coordinates
    for (int f = 0; f < LS_FREQS; ++f) fr[f] = (float)(1.5707963267948966 * std::pow(32.0, f / (double)(LS_FREQS - 1)));
    LS_CHECK(this, hipMemcpy(S.d_freqs, fr.data(), LS_FREQS * 4, hipMemcpyHostToDevice));
    return LINON_OK;
}

// Kaiming-uniform initialisation (SplitMix64 stream from the seed) of the master weights and
// the BF16 [N][K] copies; zero biases.
linon_status Context::surrogate_init_params(unsigned long long seed) {
    Surrogate& S = sg;
    unsigned long long st = seed;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        unsigned long long z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int l = 0; l < 5; ++l) {
        const int K = l < 4 ? kLayerK[l] : LS_HIDDEN;
        const int N = l < 4 ? LS_HIDDEN : LS_OUT;
        const int fan_in = l == 0 ? LS_FEATS : K;
        const double bound = (l < 4 ? std::sqrt(6.0 / fan_in) : std::sqrt(1.0 / fan_in));
        std::vector<float> w((size_t)K * N), wt((size_t)K * N, 0.0f);
        std::vector<uint16_t> wtb((size_t)K * N, 0);
        for (int k = 0; k < K; ++k)
            for (int n = 0; n < N; ++n) {
                float v = (l == 0 && k >= LS_FEATS) ? 0.0f : (float)((2.0 * next() - 1.0) * bound);
                w[(size_t)k * N + n] = v;
                if (l < 4) {
                    float r = host_bf16_round(v);
                    uint32_t u;
                    std::memcpy(&u, &r, 4);
                    wtb[(size_t)n * K + k] = (uint16_t)(u >> 16);
                }
            }
        LS_CHECK(this, hipMemcpy(S.W[l], w.data(), w.size() * 4, hipMemcpyHostToDevice));
        LS_CHECK(this, hipMemset(S.b[l], 0, N * 4));
        if (l < 4) LS_CHECK(this, hipMemcpy(S.Wt[l], wtb.data(), wtb.size() * 2, hipMemcpyHostToDevice));
    }
    S.adam_t = 0;
    return LINON_OK;
}

// Forward of the hidden layers for n rows of features: act[0] = features, act[l+1] = silu(z).
// With z != null the pre-activations are kept (training).
linon_status Context::surrogate_forward(int n, const lm::u16* feat, lm::u16* const act[5], float* const z[4],
                                        hipStream_t s) {
    Surrogate& S = sg;
    const lm::u16* x = feat;
    for (int l = 0; l < 4; ++l) {
        lm::GemmParams P;
        std::memset(&P, 0, sizeof(P));
        P.A = x;
        P.B = S.Wt[l];
        P.M = n;
        P.N = LS_HIDDEN;
        P.K = kLayerK[l];
        P.lda = kLayerK[l];
        P.ldb = kLayerK[l];
        P.a_trans = 0;
        P.b_trans = 1;
        P.k_per_split = kLayerK[l];
        P.epi = LM_EPI_BIAS_SILU;
        P.bias = S.b[l];
        P.out_bf16 = act[l + 1];
        P.z_out = z ? z[l] : nullptr;
        P.ldc = LS_HIDDEN;
        LS_CHECK(this, sg_gemm(P, 1, s));
        x = act[l + 1];
    }
    return LINON_OK;
}

linon_status Context::surrogate_pack() {
    Surrogate& S = sg;
    if (!S.d_wpack) LS_CHECK(this, hipMalloc(&S.d_wpack, (size_t)LM_PACK_TOTAL * sizeof(lm::u16)));
    for (int l = 0; l < 4; ++l) LS_CHECK(this, sg_mlp_pack(S.W[l], kLayerK[l], S.d_wpack + lm::pack_offset(l), stream));
    LS_CHECK(this, hipStreamSynchronize(stream));
    return LINON_OK;
}

// Raw network outputs for n states, through the fused inference kernel or the training path's
// kernels (features, four GEMMs, output layer); both give the same bits.
linon_status Context::surrogate_infer(const float* states, int n, float* out, bool fused) {
    Surrogate& S = sg;
    if (!S.trained) return fail(LINON_ERR_STATE, "surrogate_infer: no trained surrogate");
    if (!states || !out || n < 1 || n > (1 << 24)) return fail(LINON_ERR_ARGUMENT, "surrogate_infer: bad arguments");
    LS_CHECK(this, hipSetDevice(device));
    hipStream_t s = stream;
    float *d_in = nullptr, *d_out = nullptr;
    lm::u16 *feat = nullptr, *a0 = nullptr, *a1 = nullptr;
    auto cleanup = [&]() {
        for (void* p : {(void*)d_in, (void*)d_out, (void*)feat, (void*)a0, (void*)a1})
            if (p) (void)hipFree(p);
    };
    linon_status st = LINON_OK;
    if (dalloc(d_in, (size_t)n * 6) != hipSuccess || dalloc(d_out, (size_t)n * LS_OUT) != hipSuccess ||
        (!fused && (dalloc(feat, (size_t)n * LS_FEATS_PAD) != hipSuccess ||
                    dalloc(a0, (size_t)n * LS_HIDDEN) != hipSuccess || dalloc(a1, (size_t)n * LS_HIDDEN) != hipSuccess))) {
        cleanup();
        return fail(LINON_ERR_HIP, "surrogate_infer: out of device memory");
    }
    do {
        if (hipMemcpy(d_in, states, (size_t)n * 24, hipMemcpyHostToDevice) != hipSuccess) {
            st = fail(LINON_ERR_HIP, "surrogate_infer: copy");
            break;
        }
        if (fused) {
            lm::MlpInferParams M;
            std::memset(&M, 0, sizeof(M));
            M.states = d_in;
            M.n = n;
            M.radius = S.cfg.radius;
            M.freqs = S.d_freqs;
            M.wpack = S.d_wpack;
            for (int l = 0; l < 4; ++l) M.bias[l] = S.b[l];
            M.w_out = S.W[4];
            M.b_out = S.b[4];
            M.out = d_out;
            if (sg_mlp_infer(M, s) != hipSuccess) { st = fail(LINON_ERR_HIP, "surrogate_infer: fused"); break; }
        } else {
            ls::SurrogateIO io{d_in, n, S.cfg.radius, S.d_freqs, feat};
            if (sg_features(io, s) != hipSuccess) { st = fail(LINON_ERR_HIP, "surrogate_infer: features"); break; }
            lm::u16* const acts[5] = {feat, a0, a1, a0, a1};
            st = surrogate_forward(n, feat, acts, nullptr, s);
            if (st != LINON_OK) break;
            if (sg_out_layer(a1, S.W[4], S.b[4], n, d_out, s) != hipSuccess) {
                st = fail(LINON_ERR_HIP, "surrogate_infer: output layer");
                break;
            }
        }
        if (hipStreamSynchronize(s) != hipSuccess ||
            hipMemcpy(out, d_out, (size_t)n * LS_OUT * 4, hipMemcpyDeviceToHost) != hipSuccess)
            st = fail(LINON_ERR_HIP, "surrogate_infer: copy back");
    } while (false);
    (void)hipDeviceSynchronize();
    cleanup();
    return st;
}

linon_status Context::surrogate_train(const linon_surrogate_config& cfg, linon_surrogate_stats* stats) {
    if (!sg_field_supported(cfg.field))
        return fail(LINON_ERR_ARGUMENT, "surrogate: field must be swirl, lens, hot-air, warp, identity or "
                                        "mirage (spherical sigmoid)");
    if (cfg.train_rays < 1024 || cfg.batch < 256 || cfg.batch > (1 << 20) || cfg.batch % 64 || cfg.steps < 0 ||
        !(cfg.lr > 0.0f) || !(cfg.h > 0.0f) || cfg.max_steps < 1 || !(cfg.radius > 0.0f) ||
        !(cfg.coarse_h >= 0.0f) || (cfg.coarse_h > 0.0f && !(cfg.coarse_h > cfg.h)))
        return fail(LINON_ERR_ARGUMENT, "surrogate: bad configuration (coarse_h must be 0 or above h)");
    Surrogate& S = sg;
    S.release();
    S.cfg = cfg;
    const int N = cfg.train_rays, B = cfg.batch;
    const int n_hold = std::min(65536, std::max(1024, N / 100));
    const int n_train = N - n_hold;
    hipStream_t s = stream;
    LS_CHECK(this, hipSetDevice(device));
    linon_status st = surrogate_alloc_params();
    if (st != LINON_OK) return st;
    st = surrogate_init_params(cfg.seed * 7919ull + 17ull);
    if (st != LINON_OK) return st;

    // ---- data: uniform entry states, labels from the FP32 RK4 flow map ----
    auto t0 = std::chrono::steady_clock::now();
    LS_CHECK(this, dalloc(S.d_states, (size_t)N * 6));
    LS_CHECK(this, dalloc(S.d_targets, (size_t)N * 6));
    LS_CHECK(this, sg_sample_states(cfg.seed, N, cfg.radius, cfg.h, S.d_states, s));
    ls::FlowmapParams F;
    std::memset(&F, 0, sizeof(F));
    F.step = surrogate_step_params();
    F.states = S.d_states;
    F.n = N;
    F.exits = S.d_targets;
    LS_CHECK(this, wgsl_math_for(cfg.field) ? linon_wgslmath::sg_flowmap(kernel_field(cfg.field), F, s)
                                              : sg_flowmap(kernel_field(cfg.field), F, s));
    // the coarse flow map of the same states, and the network targets: (exit - coarse) * scale,
    // each scale making its residual's RMS 1/4; without a coarse map, exit / R and the direction
    float* base = nullptr;
    if (cfg.coarse_h > 0.0f) {
        LS_CHECK(this, dalloc(base, (size_t)N * 6));
        ls::FlowmapParams Fc = F;
        Fc.step = surrogate_coarse_step_params();
        Fc.exits = base;
        LS_CHECK(this, wgsl_math_for(cfg.field) ? linon_wgslmath::sg_flowmap(kernel_field(cfg.field), Fc, s)
                                                  : sg_flowmap(kernel_field(cfg.field), Fc, s));
        float* acc = nullptr;
        LS_CHECK(this, dalloc(acc, 2));
        LS_CHECK(this, hipMemsetAsync(acc, 0, 8, s));
        LS_CHECK(this, sg_residual_sumsq(S.d_targets, base, N, acc, s));
        float h_acc[2] = {0.0f, 0.0f};
        LS_CHECK(this, hipMemcpyAsync(h_acc, acc, 8, hipMemcpyDeviceToHost, s));
        LS_CHECK(this, hipStreamSynchronize(s));
        (void)hipFree(acc);
        const double rms_p = std::sqrt(h_acc[0] / (3.0 * N)), rms_d = std::sqrt(h_acc[1] / (3.0 * N));
        S.scale_p = (float)(0.25 / std::max(rms_p, 1e-7));
        S.scale_d = (float)(0.25 / std::max(rms_d, 1e-7));
    } else {
        S.scale_p = 1.0f / cfg.radius;
        S.scale_d = 1.0f;
    }
    LS_CHECK(this, dalloc(S.d_tscaled, (size_t)N * 6));
    LS_CHECK(this, sg_make_targets(S.d_targets, base, N, S.scale_p, S.scale_d, S.d_tscaled, s));
    LS_CHECK(this, hipStreamSynchronize(s));
    if (base) (void)hipFree(base);
    auto t1 = std::chrono::steady_clock::now();

    // ---- buffers ----
    LS_CHECK(this, dalloc(S.d_bstates, (size_t)B * 6));
    LS_CHECK(this, dalloc(S.d_btargets, (size_t)B * 6));
    LS_CHECK(this, dalloc(S.d_bfine, (size_t)B * 6));
    LS_CHECK(this, dalloc(S.d_feat, (size_t)B * LS_FEATS_PAD));
    for (int l = 1; l < 5; ++l) LS_CHECK(this, dalloc(S.d_act[l], (size_t)B * LS_HIDDEN));
    for (int l = 0; l < 4; ++l) LS_CHECK(this, dalloc(S.d_z[l], (size_t)B * LS_HIDDEN));
    LS_CHECK(this, dalloc(S.d_out, (size_t)B * LS_OUT));
    LS_CHECK(this, dalloc(S.d_dout, (size_t)B * LS_OUT));
    LS_CHECK(this, dalloc(S.d_err, (size_t)B));
    LS_CHECK(this, dalloc(S.d_dz[0], (size_t)B * LS_HIDDEN));
    LS_CHECK(this, dalloc(S.d_dz[1], (size_t)B * LS_HIDDEN));
    const int ksplit = 1024;  // batch rows per split-K chunk of a weight gradient
    const int splits = (B + ksplit - 1) / ksplit;
    // one [K][LS_HIDDEN] partial per split, for the widest K of the four layers
    const int max_k = *std::max_element(kLayerK, kLayerK + 4);
    const size_t part_elems = (size_t)splits * max_k * LS_HIDDEN;
    LS_CHECK(this, dalloc(S.d_part, part_elems));
    LS_CHECK(this, dalloc(S.d_loss, 1));

    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float conf_w = 0.1f;
    float last_loss = 0.0f;
    for (int step = 0; step < cfg.steps; ++step) {
        LS_CHECK(this, sg_gather(S.d_states, S.d_tscaled, S.d_targets, n_train, cfg.seed ^ 0x5DEECE66Dull, step, B,
                                 S.d_bstates, S.d_btargets, S.d_bfine, s));
        ls::SurrogateIO io{S.d_bstates, B, cfg.radius, S.d_freqs, S.d_feat};
        LS_CHECK(this, sg_features(io, s));
        lm::u16* const acts[5] = {S.d_feat, S.d_act[1], S.d_act[2], S.d_act[3], S.d_act[4]};
        float* const zs[4] = {S.d_z[0], S.d_z[1], S.d_z[2], S.d_z[3]};
        st = surrogate_forward(B, S.d_feat, acts, zs, s);
        if (st != LINON_OK) return st;
        LS_CHECK(this, sg_out_layer(S.d_act[4], S.W[4], S.b[4], B, S.d_out, s));
        const bool want_loss = (step % 50 == 0) || step == cfg.steps - 1;
        if (want_loss) LS_CHECK(this, hipMemsetAsync(S.d_loss, 0, 4, s));
        LS_CHECK(this, sg_loss(S.d_out, S.d_btargets, S.d_bfine, B, S.scale_p, cfg.tolerance, conf_w, S.d_dout,
                               S.d_err, S.d_loss, s));
        // output layer (FP32)
        LS_CHECK(this, sg_out_layer_grad(S.d_act[4], S.d_dout, B, S.gW[4], S.gb[4], s));
        LS_CHECK(this, sg_out_layer_back(S.d_dout, S.W[4], S.d_z[3], B, S.d_dz[0], s));
        // hidden layers, top down; dz[cur] holds dZ of layer l
        int cur = 0;
        for (int l = 3; l >= 0; --l) {
            const int K = kLayerK[l];
            if ((size_t)splits * K * LS_HIDDEN > part_elems)
                return fail(LINON_ERR_STATE, "surrogate: split-K partial buffer too small");
            lm::GemmParams G;  // dW = X^T dZ, split over the batch
            std::memset(&G, 0, sizeof(G));
            G.A = acts[l];
            G.B = S.d_dz[cur];
            G.M = K;
            G.N = LS_HIDDEN;
            G.K = B;
            G.lda = K;
            G.ldb = LS_HIDDEN;
            G.a_trans = 1;
            G.b_trans = 0;
            G.k_per_split = ksplit;
            G.epi = LM_EPI_PARTIAL;
            G.out_f32 = S.d_part;
            G.ldc = LS_HIDDEN;
            LS_CHECK(this, sg_gemm(G, splits, s));
            LS_CHECK(this, sg_reduce(S.d_part, splits, K * LS_HIDDEN, S.gW[l], s));
            LS_CHECK(this, sg_colsum(S.d_dz[cur], B, LS_HIDDEN, S.gb[l], s));
            if (l > 0) {  // dZ of layer l-1 = (dZ W^T) * silu'(z_{l-1})
                lm::GemmParams D;
                std::memset(&D, 0, sizeof(D));
                D.A = S.d_dz[cur];
                D.B = S.Wt[l];
                D.M = B;
                D.N = K;
                D.K = LS_HIDDEN;
                D.lda = LS_HIDDEN;
                D.ldb = K;
                D.a_trans = 0;
                D.b_trans = 0;
                D.k_per_split = LS_HIDDEN;
                D.epi = LM_EPI_SILU_GRAD;
                D.z_in = S.d_z[l - 1];
                D.out_bf16 = S.d_dz[cur ^ 1];
                D.ldc = K;
                LS_CHECK(this, sg_gemm(D, 1, s));
                cur ^= 1;
            }
        }
        // Adam, the learning rate on a cosine from lr down to lr / 100 over the run
        ++S.adam_t;
        const float bc1 = 1.0f - std::pow(b1, (float)S.adam_t), bc2 = 1.0f - std::pow(b2, (float)S.adam_t);
        const float lr = cfg.lr * (0.01f + 0.99f * 0.5f * (1.0f + std::cos(3.14159265f * step / std::max(1, cfg.steps))));
        for (int l = 0; l < 5; ++l) {
            const int K = l < 4 ? kLayerK[l] : LS_HIDDEN;
            const int Nn = l < 4 ? LS_HIDDEN : LS_OUT;
            LS_CHECK(this, sg_adam(S.W[l], S.mW[l], S.vW[l], S.gW[l], K, Nn, lr, b1, b2, eps, bc1, bc2,
                                   l < 4 ? S.Wt[l] : nullptr, K, s));
            LS_CHECK(this, sg_adam(S.b[l], S.mb[l], S.vb[l], S.gb[l], 1, Nn, lr, b1, b2, eps, bc1, bc2, nullptr, 0,
                                   s));
        }
        if (want_loss) {
            LS_CHECK(this, hipMemcpyAsync(&last_loss, S.d_loss, 4, hipMemcpyDeviceToHost, s));
            LS_CHECK(this, hipStreamSynchronize(s));
            if (!std::isfinite(last_loss)) return fail(LINON_ERR_STATE, "surrogate: training diverged (loss NaN)");
        }
    }
    LS_CHECK(this, hipStreamSynchronize(s));
    auto t2 = std::chrono::steady_clock::now();

    // ---- hold-out error (the last n_hold states, never sampled for training) ----
    std::vector<float> err;
    st = surrogate_eval(S.d_states + 6 * (size_t)n_train, S.d_tscaled + 6 * (size_t)n_train,
                        S.d_targets + 6 * (size_t)n_train, n_hold, err);
    if (st != LINON_OK) return st;
    st = surrogate_pack();
    if (st != LINON_OK) return st;
    S.trained = true;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->train_loss = last_loss;
        stats->label_seconds = std::chrono::duration<float>(t1 - t0).count();
        stats->train_seconds = std::chrono::duration<float>(t2 - t1).count();
        std::vector<float> e = err;
        std::sort(e.begin(), e.end());
        stats->holdout_err_q50 = e[e.size() / 2];
        stats->holdout_err_q99 = e[std::min(e.size() - 1, (size_t)(0.99 * e.size()))];
        stats->holdout_frac_within_tol =
            (float)(std::lower_bound(e.begin(), e.end(), cfg.tolerance) - e.begin()) / (float)e.size();
        stats->train_rays = (uint64_t)N;
    }
    S.train_stats = stats ? *stats : linon_surrogate_stats{};
    return LINON_OK;
}

StepParams Context::surrogate_step_params() const {
    StepParams I;
    std::memset(&I, 0, sizeof(I));
    I.h_initial = sg.cfg.h;
    I.t_max = 3.0e38f;
    I.field_weight = 1.0f;
    I.escape_r2 = sg.cfg.radius * sg.cfg.radius;
    I.adaptive = 0;
    I.max_steps = sg.cfg.max_steps;
    I.mode = LINON_MODE_ESCAPE;
    I.need_mapping = 1;
    I.clip_exit = 1;
    return I;
}

// The coarse flow map: the same loop at step coarse_h, with the step cap scaled to the same
// path length.
StepParams Context::surrogate_coarse_step_params() const {
    StepParams I = surrogate_step_params();
    I.h_initial = sg.cfg.coarse_h;
    I.max_steps = std::max(1, (int)std::ceil((double)sg.cfg.max_steps * sg.cfg.h / sg.cfg.coarse_h));
    return I;
}

// Exit-point error of the surrogate on n states (host copy of the errors): `targets` are the
// network targets of the states, `fine` their RK4 exits.
linon_status Context::surrogate_eval(const float* states, const float* targets, const float* fine, int n,
                                     std::vector<float>& err) {
    Surrogate& S = sg;
    hipStream_t s = stream;
    lm::u16* feat = nullptr;
    lm::u16 *a0 = nullptr, *a1 = nullptr;
    float *out = nullptr, *dout = nullptr, *e = nullptr, *loss = nullptr;
    auto cleanup = [&]() {
        for (void* p : {(void*)feat, (void*)a0, (void*)a1, (void*)out, (void*)dout, (void*)e, (void*)loss})
            if (p) (void)hipFree(p);
    };
    if (dalloc(feat, (size_t)n * LS_FEATS_PAD) != hipSuccess || dalloc(a0, (size_t)n * LS_HIDDEN) != hipSuccess ||
        dalloc(a1, (size_t)n * LS_HIDDEN) != hipSuccess || dalloc(out, (size_t)n * LS_OUT) != hipSuccess ||
        dalloc(dout, (size_t)n * LS_OUT) != hipSuccess || dalloc(e, (size_t)n) != hipSuccess ||
        dalloc(loss, 1) != hipSuccess) {
        cleanup();
        return fail(LINON_ERR_HIP, "surrogate_eval: out of device memory");
    }
    linon_status st = LINON_OK;
    do {
        ls::SurrogateIO io{states, n, S.cfg.radius, S.d_freqs, feat};
        if (sg_features(io, s) != hipSuccess) { st = fail(LINON_ERR_HIP, "eval features"); break; }
        lm::u16* const acts[5] = {feat, a0, a1, a0, a1};
        st = surrogate_forward(n, feat, acts, nullptr, s);
        if (st != LINON_OK) break;
        if (sg_out_layer(a1, S.W[4], S.b[4], n, out, s) != hipSuccess) { st = fail(LINON_ERR_HIP, "eval out"); break; }
        if (hipMemsetAsync(loss, 0, 4, s) != hipSuccess ||
            sg_loss(out, targets, fine, n, S.scale_p, S.cfg.tolerance, 0.1f, dout, e, loss, s) != hipSuccess) {
            st = fail(LINON_ERR_HIP, "eval loss");
            break;
        }
        err.resize(n);
        // the kernels ran on the context's (non-blocking) stream: wait for them before copying
        if (hipStreamSynchronize(s) != hipSuccess ||
            hipMemcpy(err.data(), e, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess)
            st = fail(LINON_ERR_HIP, "eval copy");
    } while (false);
    (void)hipDeviceSynchronize();
    cleanup();
    return st;
}

// One surrogate frame of the current view (no mesh: the mapping texture and the overlay):
//   1. RK4 from the camera to the bounding sphere (rays that never enter end there);
//   2. compaction of the rays that entered; the coarse flow map of their entry states (the
//      multi-fidelity form) and the MLP, accepting confident predictions;
//   3. RK4 from the entry state for rays below the confidence threshold;
//   4. the overlay's Lyapunov exponent on that mapping; accepted rays above the threshold go
//      to RK4 as well, and the overlay runs again;
//   5. re-judging: a hash-spread sample of the accepted rays is also run with RK4 and the
//      surrogate's exit-point error on it recorded.
// The compacted lists' lengths stay on the device: every launch is sized for the whole image
// and reads its count there, so the frame never waits for the host and can be captured into a
// HIP graph.  The statistics (counts, re-judged errors) are read when they are asked for; the
// phase timings are recorded when the stream is not being captured.
linon_status Context::surrogate_frame(hipStream_t s) {
    Surrogate& S = sg;
    if (!S.trained) return fail(LINON_ERR_STATE, "render_surrogate: no trained surrogate (linon_surrogate_train)");
    if (!d_color) return fail(LINON_ERR_STATE, "render_surrogate: no image buffers (the last linon_resize failed)");
    const int W = (int)width, H = (int)height, npx = W * H;
    LS_CHECK(this, hipSetDevice(device));
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    LS_CHECK(this, hipStreamIsCapturing(s, &cap));
    const bool capturing = cap != hipStreamCaptureStatusNone;
    if (S.frame_px != npx) {
        if (capturing) return fail(LINON_ERR_STATE, "render_surrogate: render once at this size before capturing");
        LS_CHECK(this, dalloc(S.f_entry, (size_t)npx * 6));
        LS_CHECK(this, dalloc(S.f_entered, (size_t)npx));
        LS_CHECK(this, dalloc(S.f_states, (size_t)npx * 6));
        LS_CHECK(this, dalloc(S.f_texel, (size_t)npx));
        LS_CHECK(this, dalloc(S.f_count, 4));
        LS_CHECK(this, dalloc(S.f_pred, (size_t)npx * 3));
        LS_CHECK(this, dalloc(S.f_accepted, (size_t)npx));
        LS_CHECK(this, dalloc(S.f_list, (size_t)npx));
        LS_CHECK(this, dalloc(S.f_exits, (size_t)npx * 6));
        LS_CHECK(this, dalloc(S.f_errs, (size_t)npx));
        LS_CHECK(this, dalloc(S.f_out, (size_t)npx * LS_OUT));
        LS_CHECK(this, dalloc(S.f_base, (size_t)npx * 6));
        LS_CHECK(this, dalloc(S.f_mask, (size_t)npx));
        S.frame_px = npx;
    }
    if (!S.ev[0])
        for (auto& x : S.ev) LS_CHECK(this, hipEventCreate(&x));
    hipEvent_t* e = S.ev;  // e[6], e[7]: around the MLP kernel, inside the MLP phase
    auto mark = [&](int k) { return capturing ? hipSuccess : hipEventRecord(e[k], s); };
    // device counts: [0] entered, [1] confidence fallback, [2] Lyapunov fallback, [3] re-judged
    int* const n_in = S.f_count;
    int* const n_fb = S.f_count + 1;
    int* const n_ly = S.f_count + 2;
    int* const n_rj = S.f_count + 3;
    const StepParams I = surrogate_step_params();
    const bool q_is_v0 = S.cfg.field == LINON_FIELD_SWIRL || S.cfg.field == LINON_FIELD_WARP ||
                         (S.cfg.field >= 5 && S.cfg.field <= 7);
    const bool wgsl = wgsl_math_for(S.cfg.field);
    const int kf = kernel_field(S.cfg.field);
    auto flowmap = [&](const ls::FlowmapParams& F) {
        return wgsl ? linon_wgslmath::sg_flowmap(kf, F, s) : sg_flowmap(kf, F, s);
    };
    LS_CHECK(this, mark(0));
    LS_CHECK(this, sg_fill_i32(S.f_count, 4, 0, s));
    // 1. approach
    ls::ApproachParams A;
    std::memset(&A, 0, sizeof(A));
    fill_camera(A.cam, camera, width, height);
    A.step = I;
    A.radius = S.cfg.radius;
    A.tiles_x = (W + LINON_TILE - 1) / LINON_TILE;
    A.entry_states = S.f_entry;
    A.entered = S.f_entered;
    A.mapping = d_mapping;
    A.steps = d_steps;
    A.q_is_v0 = q_is_v0 ? 1 : 0;
    LS_CHECK(this, wgsl ? linon_wgslmath::sg_approach(kf, A, s) : sg_approach(kf, A, s));
    LS_CHECK(this, sg_compact_entered(S.f_entered, S.f_entry, W, H, S.f_states, S.f_texel, n_in, s));
    LS_CHECK(this, mark(1));
    // 2. coarse flow map (multi-fidelity form), then the MLP in one fused launch
    const bool coarse = S.cfg.coarse_h > 0.0f;
    if (coarse) {
        ls::FlowmapParams Fc;
        std::memset(&Fc, 0, sizeof(Fc));
        Fc.step = surrogate_coarse_step_params();
        Fc.states = S.f_states;
        Fc.n = npx;
        Fc.n_dev = n_in;
        Fc.exits = S.f_base;
        LS_CHECK(this, flowmap(Fc));
    }
    LS_CHECK(this, mark(6));
    lm::MlpInferParams M;
    std::memset(&M, 0, sizeof(M));
    M.states = S.f_states;
    M.n = npx;
    M.n_dev = n_in;
    M.radius = S.cfg.radius;
    M.freqs = S.d_freqs;
    M.wpack = S.d_wpack;
    for (int l = 0; l < 4; ++l) M.bias[l] = S.b[l];
    M.w_out = S.W[4];
    M.b_out = S.b[4];
    M.out = S.f_out;
    LS_CHECK(this, sg_mlp_infer(M, s));
    LS_CHECK(this, mark(7));
    LS_CHECK(this, sg_accept(S.f_out, coarse ? S.f_base : nullptr, S.f_texel, npx, n_in, S.scale_p,
                             S.cfg.conf_threshold, d_mapping, S.f_pred, S.f_list, n_fb, S.f_accepted, s));
    LS_CHECK(this, mark(2));
    // 3. RK4 for the unconfident rays
    ls::FlowmapParams F;
    std::memset(&F, 0, sizeof(F));
    F.step = I;
    F.states = S.f_states;
    F.exits = S.f_exits;
    F.index = S.f_list;
    F.n = npx;
    F.n_dev = n_fb;
    // the approach's iterations count toward max_steps, as in the plain frame's one RK4 run
    F.iters_at = d_steps;
    F.texel_of = S.f_texel;
    LS_CHECK(this, flowmap(F));
    LS_CHECK(this, sg_apply_rk4(S.f_list, npx, n_fb, S.f_exits, S.f_texel, S.f_pred, d_mapping, nullptr, 1, s));
    LS_CHECK(this, mark(3));
    // 4. Lyapunov gate (the overlay computes lambda from the mapping)
    OverlayParams O;
    std::memset(&O, 0, sizeof(O));
    O.width = W;
    O.height = H;
    O.overlay_mode = 1;
    O.delta = settings.central_difference_delta;
    O.scaling = settings.lyapunov_scaling;
    O.srgb = srgb_output ? 1 : 0;
    O.ray_casting = d_color;
    O.mapping = d_mapping;
    O.final_color = d_final;
    O.exponents = d_exponents;
    O.lyapunov = d_lyap;
    LS_CHECK(this, sg_fill_i32(reinterpret_cast<int*>(d_color), npx, (int)0xff000000u, s));  // black, alpha 1
    LS_CHECK(this, launch_overlay(O, s));
    LS_CHECK(this, sg_lyapunov_gate(S.f_accepted, S.f_texel, npx, n_in, d_lyap, S.cfg.lyapunov_threshold, S.f_list,
                                    n_ly, s));
    F.n_dev = n_ly;
    LS_CHECK(this, flowmap(F));
    LS_CHECK(this, sg_apply_rk4(S.f_list, npx, n_ly, S.f_exits, S.f_texel, S.f_pred, d_mapping, nullptr, 1, s));
    LS_CHECK(this, launch_overlay(O, s));
    LS_CHECK(this, mark(4));
    LS_CHECK(this, sg_accepted_mask(S.f_accepted, S.f_texel, npx, n_in, npx, S.f_mask, s));
    // 5. re-judge a sample of the accepted rays
    const int stride = S.cfg.rejudge_fraction > 0.0f ? std::max(1, (int)std::lround(1.0 / S.cfg.rejudge_fraction)) : 0;
    if (stride > 0) {
        LS_CHECK(this, sg_pick_rejudge(S.f_accepted, npx, n_in, stride, 0xC0FFEEull + S.frames, S.f_list, n_rj, s));
        F.n_dev = n_rj;
        LS_CHECK(this, flowmap(F));
        LS_CHECK(this, sg_apply_rk4(S.f_list, npx, n_rj, S.f_exits, S.f_texel, S.f_pred, d_mapping, S.f_errs, 0, s));
    }
    LS_CHECK(this, mark(5));
    S.frame_timed = !capturing;
    S.has_frame = true;
    ++S.frames;
    return LINON_OK;
}

// Counts and re-judged errors of the last frame that ran on the device (read back on every
// call: a replayed graph rewrites them without the host), and the phase timings when the last
// frame the host enqueued was not captured (zero otherwise).
linon_status Context::surrogate_collect_stats() {
    Surrogate& S = sg;
    if (!S.has_frame) return LINON_OK;
    LS_CHECK(this, hipSetDevice(device));
    LS_CHECK(this, hipDeviceSynchronize());
    int cnt[4] = {0, 0, 0, 0};
    LS_CHECK(this, hipMemcpy(cnt, S.f_count, sizeof(cnt), hipMemcpyDeviceToHost));
    linon_surrogate_stats& F2 = S.frame_stats;
    F2 = S.train_stats;
    F2.frame_rays = (uint64_t)S.frame_px;
    F2.entered = (uint64_t)cnt[0];
    F2.conf_fallback = (uint64_t)cnt[1];
    F2.lyapunov_fallback = (uint64_t)cnt[2];
    F2.accepted = (uint64_t)(cnt[0] - cnt[1] - cnt[2]);
    F2.rejudged = (uint64_t)cnt[3];
    if (S.frame_timed) {
        hipEvent_t* e = S.ev;
        float ms[5];
        for (int k = 0; k < 5; ++k) (void)hipEventElapsedTime(&ms[k], e[k], e[k + 1]);
        F2.approach_ms = ms[0];
        F2.fallback_ms = ms[2];
        F2.overlay_ms = ms[3];
        F2.rejudge_ms = ms[4];
        F2.frame_ms = ms[0] + ms[1] + ms[2] + ms[3];  // ms[1]: coarse flow map + MLP + verdicts
        (void)hipEventElapsedTime(&F2.coarse_ms, e[1], e[6]);
        (void)hipEventElapsedTime(&F2.mlp_ms, e[6], e[7]);  // the fused MLP kernel alone
    }
    if (cnt[3] > 0) {
        std::vector<float> rj(cnt[3]);
        LS_CHECK(this, hipMemcpy(rj.data(), S.f_errs, (size_t)cnt[3] * 4, hipMemcpyDeviceToHost));
        std::sort(rj.begin(), rj.end());
        F2.rejudge_err_q50 = rj[rj.size() / 2];
        F2.rejudge_err_q99 = rj[std::min(rj.size() - 1, (size_t)(0.99 * rj.size()))];
        F2.rejudge_err_max = rj.back();
        F2.rejudge_frac_within_tol =
            (float)(std::upper_bound(rj.begin(), rj.end(), S.cfg.tolerance) - rj.begin()) / (float)rj.size();
    }
    return LINON_OK;
}

namespace {

// Saved-surrogate blob: header, then per layer l = 0..4 the [K][N] FP32 weights and N biases.
struct SurrogateBlobHeader {
    char magic[8];
    uint32_t version, header_bytes;
    uint32_t dims[5][2];  // K, N per layer
    float freqs[LS_FREQS];
    float scale[2];  // output scales of the exit point and direction
    linon_surrogate_config cfg;
    linon_surrogate_stats stats;
};
const char kBlobMagic[8] = {'L', 'I', 'N', 'O', 'N', 'S', 'G', '1'};

size_t blob_bytes() {
    size_t n = sizeof(SurrogateBlobHeader);
    for (int l = 0; l < 5; ++l) {
        const size_t K = l < 4 ? kLayerK[l] : LS_HIDDEN, N = l < 4 ? LS_HIDDEN : LS_OUT;
        n += (K * N + N) * sizeof(float);
    }
    return n;
}

}  // namespace

linon_status Context::surrogate_save(void* out, size_t capacity, size_t* needed) {
    const size_t total = blob_bytes();
    if (needed) *needed = total;
    if (!out) return LINON_OK;
    if (!sg.trained) return fail(LINON_ERR_STATE, "surrogate_save: no trained surrogate");
    if (capacity < total) return fail(LINON_ERR_ARGUMENT, "surrogate_save: buffer smaller than the blob");
    LS_CHECK(this, hipSetDevice(device));
    LS_CHECK(this, hipDeviceSynchronize());
    SurrogateBlobHeader h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kBlobMagic, 8);
    h.version = 1;
    h.header_bytes = sizeof(h);
    for (int l = 0; l < 5; ++l) {
        h.dims[l][0] = l < 4 ? kLayerK[l] : LS_HIDDEN;
        h.dims[l][1] = l < 4 ? LS_HIDDEN : LS_OUT;
    }
    LS_CHECK(this, hipMemcpy(h.freqs, sg.d_freqs, sizeof(h.freqs), hipMemcpyDeviceToHost));
    h.scale[0] = sg.scale_p;
    h.scale[1] = sg.scale_d;
    h.cfg = sg.cfg;
    h.stats = sg.train_stats;
    char* p = static_cast<char*>(out);
    std::memcpy(p, &h, sizeof(h));
    p += sizeof(h);
    for (int l = 0; l < 5; ++l) {
        const size_t K = h.dims[l][0], N = h.dims[l][1];
        LS_CHECK(this, hipMemcpy(p, sg.W[l], K * N * 4, hipMemcpyDeviceToHost));
        p += K * N * 4;
        LS_CHECK(this, hipMemcpy(p, sg.b[l], N * 4, hipMemcpyDeviceToHost));
        p += N * 4;
    }
    return LINON_OK;
}

linon_status Context::surrogate_load(const void* data, size_t bytes) {
    const size_t total = blob_bytes();
    SurrogateBlobHeader h;
    if (!data || bytes != total) return fail(LINON_ERR_ARGUMENT, "surrogate_load: blob size mismatch");
    std::memcpy(&h, data, sizeof(h));
    if (std::memcmp(h.magic, kBlobMagic, 8) != 0 || h.version != 1 || h.header_bytes != sizeof(h))
        return fail(LINON_ERR_ARGUMENT, "surrogate_load: not a surrogate blob of this version");
    for (int l = 0; l < 5; ++l)
        if (h.dims[l][0] != (uint32_t)(l < 4 ? kLayerK[l] : LS_HIDDEN) || h.dims[l][1] != (uint32_t)(l < 4 ? LS_HIDDEN : LS_OUT))
            return fail(LINON_ERR_ARGUMENT, "surrogate_load: layer sizes differ from this build's network");
    if (!sg_field_supported(h.cfg.field) || !(h.cfg.radius > 0.0f) || !(h.cfg.h > 0.0f) || h.cfg.max_steps < 1 ||
        !(h.cfg.coarse_h >= 0.0f) || !(h.scale[0] > 0.0f) || !(h.scale[1] > 0.0f))
        return fail(LINON_ERR_ARGUMENT, "surrogate_load: configuration out of range");
    LS_CHECK(this, hipSetDevice(device));
    LS_CHECK(this, hipDeviceSynchronize());
    Surrogate& S = sg;
    S.release();
    S.cfg = h.cfg;
    linon_status st = surrogate_alloc_params();
    if (st != LINON_OK) return st;
    LS_CHECK(this, hipMemcpy(S.d_freqs, h.freqs, sizeof(h.freqs), hipMemcpyHostToDevice));
    const char* p = static_cast<const char*>(data) + sizeof(h);
    for (int l = 0; l < 5; ++l) {
        const size_t K = h.dims[l][0], N = h.dims[l][1];
        std::vector<float> w(K * N);
        std::memcpy(w.data(), p, K * N * 4);
        p += K * N * 4;
        LS_CHECK(this, hipMemcpy(S.W[l], w.data(), K * N * 4, hipMemcpyHostToDevice));
        LS_CHECK(this, hipMemcpy(S.b[l], p, N * 4, hipMemcpyHostToDevice));
        p += N * 4;
        if (l < 4) {  // the BF16 [N][K] copy the GEMMs read
            std::vector<uint16_t> wt(K * N);
            for (size_t k = 0; k < K; ++k)
                for (size_t n = 0; n < N; ++n) {
                    const float r = host_bf16_round(w[k * N + n]);
                    uint32_t u;
                    std::memcpy(&u, &r, 4);
                    wt[n * K + k] = (uint16_t)(u >> 16);
                }
            LS_CHECK(this, hipMemcpy(S.Wt[l], wt.data(), wt.size() * 2, hipMemcpyHostToDevice));
        }
    }
    S.scale_p = h.scale[0];
    S.scale_d = h.scale[1];
    S.train_stats = h.stats;
    st = surrogate_pack();
    if (st != LINON_OK) return st;
    S.trained = true;
    return LINON_OK;
}

}  // namespace linon

using linon::Context;

extern "C" {

linon_status linon_surrogate_train(linon_context* ctx, const linon_surrogate_config* cfg,
                                   linon_surrogate_stats* stats) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return LINON_ERR_ARGUMENT;
    c->err.clear();
    if (!cfg) return c->fail(LINON_ERR_ARGUMENT, "surrogate_train: null config");
    return c->surrogate_train(*cfg, stats);
}

linon_status linon_render_surrogate(linon_context* ctx, void* stream) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return LINON_ERR_ARGUMENT;
    c->err.clear();
    return c->surrogate_frame(stream ? reinterpret_cast<hipStream_t>(stream) : c->stream);
}

linon_status linon_surrogate_save(linon_context* ctx, void* out, size_t capacity, size_t* needed) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return LINON_ERR_ARGUMENT;
    c->err.clear();
    return c->surrogate_save(out, capacity, needed);
}

linon_status linon_surrogate_load(linon_context* ctx, const void* data, size_t bytes) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return LINON_ERR_ARGUMENT;
    c->err.clear();
    return c->surrogate_load(data, bytes);
}

linon_status linon_read_surrogate_accepted(linon_context* ctx, uint8_t* mask, size_t bytes) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return LINON_ERR_ARGUMENT;
    c->err.clear();
    if (!mask) return c->fail(LINON_ERR_ARGUMENT, "read_surrogate_accepted: null buffer");
    if (!c->sg.frames || !c->sg.f_mask || c->sg.frame_px != (int)(c->width * c->height))
        return c->fail(LINON_ERR_STATE, "read_surrogate_accepted: no surrogate frame at the current size");
    return c->read(c->sg.f_mask, mask, (size_t)c->sg.frame_px, bytes, "read_surrogate_accepted");
}

linon_status linon_surrogate_infer(linon_context* ctx, const float* states, int n, float* out, int fused) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return LINON_ERR_ARGUMENT;
    c->err.clear();
    if (fused != 0 && fused != 1) return c->fail(LINON_ERR_ARGUMENT, "surrogate_infer: fused must be 0 or 1");
    return c->surrogate_infer(states, n, out, fused == 1);
}

linon_status linon_get_surrogate_stats(linon_context* ctx, linon_surrogate_stats* out) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c || !out) return LINON_ERR_ARGUMENT;
    c->err.clear();
    if (c->sg.frames) {
        const linon_status st = c->surrogate_collect_stats();
        if (st != LINON_OK) return st;
    }
    *out = c->sg.frames ? c->sg.frame_stats : c->sg.train_stats;
    return LINON_OK;
}

}  // extern "C"

extern "C" linon_status linon_flowmap(int device, int field, int math, const linon_integrator* integrator,
                                      float field_weight, const float* states, int n, float* exits, int32_t* steps) {
    if (!integrator || !states || !exits || n < 1 || n > (1 << 26) || integrator->mode != LINON_MODE_ESCAPE ||
        math < LINON_MATH_AUTO || math > LINON_MATH_WGSL || !(integrator->h_initial > 0.0f) ||
        integrator->max_steps < 1 || (integrator->clip_exit != 0 && integrator->clip_exit != 1))
        return LINON_ERR_ARGUMENT;
    if (!linon::sg_field_supported(field)) return LINON_ERR_UNSUPPORTED;
    if (hipSetDevice(device) != hipSuccess) return LINON_ERR_HIP;
    float *d_in = nullptr, *d_out = nullptr;
    int* d_steps = nullptr;
    linon_status st = LINON_OK;
    auto ok = [&](hipError_t e) {
        if (e != hipSuccess) st = LINON_ERR_HIP;
        return e == hipSuccess;
    };
    const size_t bytes = (size_t)n * 6 * sizeof(float);
    if (ok(hipMalloc(&d_in, bytes)) && ok(hipMalloc(&d_out, bytes)) && ok(hipMalloc(&d_steps, (size_t)n * 4)) &&
        ok(hipMemcpy(d_in, states, bytes, hipMemcpyHostToDevice)) && ok(hipMemset(d_steps, 0, (size_t)n * 4))) {
        ls::FlowmapParams F;
        std::memset(&F, 0, sizeof(F));
        linon::fill_step(F.step, *integrator, field_weight);
        F.states = d_in;
        F.n = n;
        F.exits = d_out;
        F.steps = d_steps;
        const int kf = linon::kernel_field_id(math, field);
        if (ok(linon::wgsl_math_default(math, field) ? linon_wgslmath::sg_flowmap(kf, F, nullptr)
                                                      : linon::sg_flowmap(kf, F, nullptr)) &&
            ok(hipDeviceSynchronize()) && ok(hipMemcpy(exits, d_out, bytes, hipMemcpyDeviceToHost)) && steps)
            ok(hipMemcpy(steps, d_steps, (size_t)n * 4, hipMemcpyDeviceToHost));
    }
    for (void* p : {(void*)d_in, (void*)d_out, (void*)d_steps})
        if (p) (void)hipFree(p);
    return st;
}

extern "C" linon_status linon_selftest_gemm(int device, const uint16_t* A, const uint16_t* B, int M, int N, int K,
                                            int a_trans, int b_trans, int splits, const float* bias, int silu,
                                            float* C, uint16_t* C_bf16) {
    if (!A || !B || !C || M < 1 || N < 1 || K < 1 || splits < 1 || (K % splits) != 0 || ((K / splits) % LM_BK) != 0)
        return LINON_ERR_ARGUMENT;
    if (hipSetDevice(device) != hipSuccess) return LINON_ERR_HIP;
    lm::u16 *dA = nullptr, *dB = nullptr, *dC16 = nullptr;
    float *dC = nullptr, *dP = nullptr, *dBias = nullptr, *dZ = nullptr;
    linon_status st = LINON_OK;
    auto ok = [&](hipError_t e) {
        if (e != hipSuccess) st = LINON_ERR_HIP;
        return e == hipSuccess;
    };
    if (ok(hipMalloc(&dA, (size_t)M * K * 2)) && ok(hipMalloc(&dB, (size_t)K * N * 2)) &&
        ok(hipMalloc(&dC, (size_t)M * N * 4)) && ok(hipMalloc(&dP, (size_t)splits * M * N * 4)) &&
        ok(hipMalloc(&dBias, (size_t)N * 4)) && ok(hipMalloc(&dC16, (size_t)M * N * 2)) &&
        ok(hipMalloc(&dZ, (size_t)M * N * 4)) && ok(hipMemcpy(dA, A, (size_t)M * K * 2, hipMemcpyHostToDevice)) &&
        ok(hipMemcpy(dB, B, (size_t)K * N * 2, hipMemcpyHostToDevice)) &&
        ok(hipMemset(dBias, 0, (size_t)N * 4)) &&
        (!bias || ok(hipMemcpy(dBias, bias, (size_t)N * 4, hipMemcpyHostToDevice)))) {
        lm::GemmParams P;
        std::memset(&P, 0, sizeof(P));
        P.A = dA;
        P.B = dB;
        P.M = M;
        P.N = N;
        P.K = K;
        P.lda = a_trans ? M : K;
        P.ldb = b_trans ? K : N;
        P.a_trans = a_trans;
        P.b_trans = b_trans;
        P.k_per_split = K / splits;
        P.ldc = N;
        P.bias = dBias;
        if (silu) {
            P.epi = LM_EPI_BIAS_SILU;
            P.out_bf16 = dC16;
            P.z_out = dC;
        } else if (splits > 1) {
            P.epi = LM_EPI_PARTIAL;
            P.out_f32 = dP;
        } else {
            P.epi = LM_EPI_STORE_F32;
            P.out_f32 = dC;
        }
        if (ok(linon::sg_gemm(P, splits, nullptr)) && splits > 1 && !silu)
            ok(linon::sg_reduce(dP, splits, M * N, dC, nullptr));
        if (st == LINON_OK && ok(hipDeviceSynchronize())) {
            ok(hipMemcpy(C, dC, (size_t)M * N * 4, hipMemcpyDeviceToHost));
            if (silu && C_bf16) ok(hipMemcpy(C_bf16, dC16, (size_t)M * N * 2, hipMemcpyDeviceToHost));
        }
    }
    for (void* p : {(void*)dA, (void*)dB, (void*)dC, (void*)dP, (void*)dBias, (void*)dC16, (void*)dZ})
        if (p) (void)hipFree(p);
    return st;
}
