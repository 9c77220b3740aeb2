// Kernels of the learned flow-map surrogate: training (Fourier features, BF16-MFMA GEMMs with
// fused epilogues, an FP32 output layer, loss, split-K weight-gradient reduction, Adam) and a
// frame (RK4 approach to the bounding sphere, MLP inference, RK4 fallback, re-judging).
#include <hip/hip_runtime.h>

#include "device/fields.hpp"
#include "device/gemm_nt.hpp"
#include "device/mlp.hpp"
#include "device/mlp_fused.hpp"
#include "device/surrogate.hpp"
#include "surrogate_kernels.hpp"

namespace LINON_KERNEL_NS {

using lm::u16;

__global__ void __launch_bounds__(256) gemm_kernel(lm::GemmParams P) { lm::gemm_body(P); }
__global__ void __launch_bounds__(256) gemm_nt_kernel(lm::GemmParams P) { lm::gemm_nt_body(P); }

__global__ void __launch_bounds__(256) features_kernel(ls::SurrogateIO S) { ls::features_body(S); }
__global__ void __launch_bounds__(256) mlp_pack_kernel(const float* W, int K, u16* dst) { lm::mlp_pack_body(W, K, dst); }
// two 128-ray workgroups per CU (75 KB of LDS each)
__global__ void __launch_bounds__(256, 2) mlp_infer_kernel(lm::MlpInferParams P) { lm::mlp_infer_body(P); }

template <class F>
__global__ void __launch_bounds__(256) approach_kernel(ls::ApproachParams A) {
    ls::approach_body<F>(A);
}

template <class F>
__global__ void __launch_bounds__(256) flowmap_kernel(ls::FlowmapParams P) {
    ls::flowmap_body<F>(P);
}

// Uniform entry states: p uniform on the sphere of radius R, pulled inside by up to one step
// (u * inset), q uniform over directions pointing into the sphere (q . p < 0); counter-based
// SplitMix64 stream, so the data set is a function of the seed.
__device__ __forceinline__ float u01(unsigned long long seed, unsigned long long i) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (float)(z >> 40) * (1.0f / 16777216.0f);
}
__global__ void __launch_bounds__(256) sample_states_kernel(unsigned long long seed, int n, float radius, float inset,
                                                            float* states) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long b = 6ull * (unsigned long long)i;
    float z1 = 2.0f * u01(seed, b) - 1.0f, ph1 = 6.283185307179586f * u01(seed, b + 1);
    float z2 = 2.0f * u01(seed, b + 2) - 1.0f, ph2 = 6.283185307179586f * u01(seed, b + 3);
    float r = radius - inset * u01(seed, b + 4);
    float s1 = sqrtf(fmaxf(0.0f, 1.0f - z1 * z1)), s2 = sqrtf(fmaxf(0.0f, 1.0f - z2 * z2));
    float px = s1 * cosf(ph1), py = s1 * sinf(ph1), pz = z1;
    float qx = s2 * cosf(ph2), qy = s2 * sinf(ph2), qz = z2;
    if (px * qx + py * qy + pz * qz > 0.0f) {
        qx = -qx;
        qy = -qy;
        qz = -qz;
    }
    float* o = states + 6 * (size_t)i;
    o[0] = r * px; o[1] = r * py; o[2] = r * pz;
    o[3] = qx; o[4] = qy; o[5] = qz;
}

// Picks roughly one accepted ray in `stride` (hash-spread) for re-judging with RK4.
__global__ void __launch_bounds__(256) pick_rejudge_kernel(const int* accepted, int n, const int* n_dev, int stride,
                                                           unsigned long long seed, int* list, int* count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (n_dev ? *n_dev : n) || !accepted[i]) return;
    if ((int)(u01(seed, (unsigned long long)i) * (float)stride) != 0) return;
    int slot = atomicAdd(count, 1);
    list[slot] = i;
}

// Network targets of the training set: (exit - base) * scale, per row (point scale for the
// three exit-point values, direction scale for the exit direction); base = the coarse flow map,
// or none (0) when the network predicts the exit itself.
__global__ void __launch_bounds__(256) make_targets_kernel(const float* fine, const float* base, int n,
                                                           float scale_p, float scale_d, float* tgt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 6 * n) return;
    const int k = t % 6;
    const float b = base ? base[t] : 0.0f;
    tgt[t] = (fine[t] - b) * (k < 3 ? scale_p : scale_d);
}

// Sums of squares of (fine - base) over all rows: acc[0] over the exit-point values, acc[1]
// over the directions (block reduction, one atomic per block and accumulator).
__global__ void __launch_bounds__(256) residual_sumsq_kernel(const float* fine, const float* base, int n, float* acc) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    float sp = 0.0f, sd = 0.0f;
    if (t < n) {
        for (int k = 0; k < 6; ++k) {
            const float d = fine[6 * (size_t)t + k] - (base ? base[6 * (size_t)t + k] : 0.0f);
            (k < 3 ? sp : sd) += d * d;
        }
    }
    __shared__ float red[2][256];
    red[0][threadIdx.x] = sp;
    red[1][threadIdx.x] = sd;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(acc, red[0][0]);
        atomicAdd(acc + 1, red[1][0]);
    }
}

// Gathers a mini-batch: rows idx[i] of the training states, network targets and RK4 exits (idx
// from a counter-based hash of (seed, step, i)).
__global__ void __launch_bounds__(256) gather_batch_kernel(const float* states, const float* targets,
                                                           const float* fine, int n_data, unsigned long long seed,
                                                           int step, int batch, float* bstates, float* btargets,
                                                           float* bfine) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= batch) return;
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * ((unsigned long long)step * (unsigned long long)batch + i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const int r = (int)(z % (unsigned long long)n_data);
    for (int k = 0; k < 6; ++k) {
        bstates[6 * (size_t)i + k] = states[6 * (size_t)r + k];
        btargets[6 * (size_t)i + k] = targets[6 * (size_t)r + k];
        bfine[6 * (size_t)i + k] = fine[6 * (size_t)r + k];
    }
}

// Exit-point error of a prediction against the RK4 exit, as the parity tests measure it:
// max over x, y, z of |pred - exit| / max(1, |exit|).
__device__ __forceinline__ float exit_error(float px, float py, float pz, const float* e) {
    const float ex = fabsf(px - e[0]) / fmaxf(1.0f, fabsf(e[0]));
    const float ey = fabsf(py - e[1]) / fmaxf(1.0f, fabsf(e[1]));
    const float ez = fabsf(pz - e[2]) / fmaxf(1.0f, fabsf(e[2]));
    return fmaxf(ex, fmaxf(ey, ez));
}

// FP32 output layer: out[i][o] = sum_k act[i][k] * W[k][o] + b[o] (bf16 activations, fp32
// weights), one lane per (row, output).
__global__ void __launch_bounds__(256) out_layer_kernel(const u16* act, const float* W, const float* b, int n,
                                                        float* out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * LS_OUT) return;
    const int i = t / LS_OUT, o = t % LS_OUT;
    const u16* a = act + (size_t)i * LS_HIDDEN;
    float s = b[o];
    for (int k = 0; k < LS_HIDDEN; ++k) s = __builtin_fmaf(lm::bf2f(a[k]), W[k * LS_OUT + o], s);
    out[(size_t)i * LS_OUT + o] = s;
}

// Loss and output gradient.  Regression: mean squared error of the 6 scaled targets.
// Confidence: logistic loss against the label [exit-point error <= tol], the error being the
// prediction's (base + output / scale_p) against the RK4 exit `fine`.  Writes dOut [n][LS_OUT]
// (already / n) and the per-row error.
__global__ void __launch_bounds__(256) loss_kernel(const float* out, const float* target, const float* fine, int n,
                                                   float scale_p, float tol, float conf_weight, float* dout,
                                                   float* err, float* loss_sum) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < n) {
        const float* o = out + (size_t)i * LS_OUT;
        const float* y = target + 6 * (size_t)i;
        const float* f = fine + 6 * (size_t)i;
        float pe[3];
        for (int k = 0; k < 6; ++k) {
            float d = o[k] - y[k];
            l += d * d / 6.0f;
            dout[(size_t)i * LS_OUT + k] = 2.0f * d / (6.0f * n);
            if (k < 3) pe[k] = f[k] + d / scale_p;  // the prediction: exit + (output - target) / scale
        }
        const float e = exit_error(pe[0], pe[1], pe[2], f);
        err[i] = e;
        const float label = e <= tol ? 1.0f : 0.0f;
        const float c = o[6];
        const float s = 1.0f / (1.0f + __expf(-c));
        l += conf_weight * (fmaxf(c, 0.0f) - c * label + log1pf(__expf(-fabsf(c))));
        dout[(size_t)i * LS_OUT + 6] = conf_weight * (s - label) / n;
    }
    // block reduction of the loss
    __shared__ float red[256];
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss_sum, red[0] / n);
}

// Output-layer backward: dA[i][k] = sum_o dOut[i][o] W[k][o], gated by silu'(z) -> bf16 dZ.
__global__ void __launch_bounds__(256) out_layer_back_kernel(const float* dout, const float* W, const float* z, int n,
                                                             u16* dz) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * LS_HIDDEN) return;
    const int i = t / LS_HIDDEN, k = t % LS_HIDDEN;
    float s = 0.0f;
    for (int o = 0; o < LS_OUT; ++o) s = __builtin_fmaf(dout[(size_t)i * LS_OUT + o], W[k * LS_OUT + o], s);
    dz[t] = lm::f2bf(s * lm::silu_grad(z[t]));
}

// Output-layer weight/bias gradients: one block per k (0..LS_HIDDEN, where k == LS_HIDDEN is
// the bias), reducing over the batch.
__global__ void __launch_bounds__(256) out_layer_grad_kernel(const u16* act, const float* dout, int n, float* dW,
                                                             float* db) {
    const int k = blockIdx.x;
    float s[LS_OUT];
    for (int o = 0; o < LS_OUT; ++o) s[o] = 0.0f;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const float a = (k < LS_HIDDEN) ? lm::bf2f(act[(size_t)i * LS_HIDDEN + k]) : 1.0f;
        for (int o = 0; o < LS_OUT; ++o) s[o] = __builtin_fmaf(a, dout[(size_t)i * LS_OUT + o], s[o]);
    }
    __shared__ float red[LS_OUT][256];
    for (int o = 0; o < LS_OUT; ++o) red[o][threadIdx.x] = s[o];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int o = 0; o < LS_OUT; ++o) red[o][threadIdx.x] += red[o][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < LS_OUT) {
        if (k < LS_HIDDEN)
            dW[k * LS_OUT + threadIdx.x] = red[threadIdx.x][0];
        else
            db[threadIdx.x] = red[threadIdx.x][0];
    }
}

// Sums split-K partials: out[j] = sum_s part[s][j].
__global__ void __launch_bounds__(256) reduce_partials_kernel(const float* part, int splits, int n, float* out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float s = 0.0f;
    for (int k = 0; k < splits; ++k) s += part[(size_t)k * n + j];
    out[j] = s;
}

// Bias gradient of a hidden layer: column sums of bf16 dZ [n][cols].
__global__ void __launch_bounds__(256) colsum_kernel(const u16* dz, int n, int cols, float* out) {
    const int c = blockIdx.x;
    float s = 0.0f;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += lm::bf2f(dz[(size_t)i * cols + c]);
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
}

// Adam on FP32 master weights [K][N]; optionally refreshes the bf16 transposed copy [N][ldt].
__global__ void __launch_bounds__(256) adam_kernel(float* w, float* m, float* v, const float* g, int K, int N,
                                                   float lr, float b1, float b2, float eps, float bc1, float bc2,
                                                   u16* wt, int ldt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= K * N) return;
    const float gr = g[t];
    const float mm = b1 * m[t] + (1.0f - b1) * gr;
    const float vv = b2 * v[t] + (1.0f - b2) * gr * gr;
    m[t] = mm;
    v[t] = vv;
    const float nw = w[t] - lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    w[t] = nw;
    if (wt) {
        const int k = t / N, n = t % N;
        wt[(size_t)n * ldt + k] = lm::f2bf(nw);
    }
}

// Surrogate verdict per entered ray: accept (write the exit point to the mapping) when the
// confidence reaches the threshold; otherwise queue the ray for RK4.  The exit point is the
// base (coarse flow map, or 0) plus the output over the point scale.
__global__ void __launch_bounds__(256) accept_kernel(const float* out, const float* base, const int* texel_of, int n,
                                                     const int* n_dev, float scale_p, float conf_threshold,
                                                     float* mapping, float* pred_exit, int* fallback,
                                                     int* n_fallback, int* accepted) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (n_dev ? *n_dev : n)) return;
    const float* o = out + (size_t)i * LS_OUT;
    const float conf = 1.0f / (1.0f + __expf(-o[6]));
    const int texel = texel_of[i];
    const float inv = 1.0f / scale_p;
    const float* b = base ? base + 6 * (size_t)i : nullptr;
    const float px = (b ? b[0] : 0.0f) + o[0] * inv, py = (b ? b[1] : 0.0f) + o[1] * inv,
                pz = (b ? b[2] : 0.0f) + o[2] * inv;
    pred_exit[3 * (size_t)i] = px;
    pred_exit[3 * (size_t)i + 1] = py;
    pred_exit[3 * (size_t)i + 2] = pz;
    if (conf >= conf_threshold) {
        reinterpret_cast<float4*>(mapping)[texel] = make_float4(px, py, pz, 1.0f);
        accepted[i] = 1;
    } else {
        accepted[i] = 0;
        int slot = atomicAdd(n_fallback, 1);
        fallback[slot] = i;
    }
}

// Compacts the rays that entered the sphere into a prefix of entry-state rows; texel_of[i]
// remembers where ray i's result goes.  Threads walk the image in the approach kernel's 8x8
// tile order, one wave per tile, so each wave appends one tile's entering rays together: the
// RK4 legs that run over the compacted rows (coarse flow map, fallback, re-judging) then see
// neighbouring pixels side by side in a wave, as the plain kernel's tiles do, rather than 64
// pixels of one image row.
__global__ void __launch_bounds__(256) compact_entered_kernel(const int* entered, const float* entry_states, int W,
                                                              int H, float* states, int* texel_of, int* count) {
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int tiles_x = (W + LINON_TILE - 1) / LINON_TILE;
    const int gx = (tile % tiles_x) * LINON_TILE + (lane & (LINON_TILE - 1));
    const int gy = (tile / tiles_x) * LINON_TILE + (lane / LINON_TILE);
    const int t = (H - 1 - gy) * W + gx;  // the approach kernel's texel of (gx, gy)
    const bool e = gx < W && gy < H && entered[t];
    // wave-aggregated append: one atomic per wave, lanes placed by their rank among the
    // entering lanes of the wave
    const unsigned long long mask = __ballot(e);
    int base = 0;
    if (lane == 0 && mask) base = atomicAdd(count, __popcll(mask));
    base = __shfl(base, 0, 64);
    if (e) {
        const int slot = base + __popcll(mask & ((1ull << lane) - 1ull));
        for (int k = 0; k < 6; ++k) states[6 * (size_t)slot + k] = entry_states[6 * (size_t)t + k];
        texel_of[slot] = t;
    }
}

// Rays re-judged by RK4 or sent to RK4: write the RK4 exit into the mapping; for re-judged
// rays also record the surrogate's error against the RK4 exit (exit_error).
__global__ void __launch_bounds__(256) apply_rk4_kernel(const int* list, int n, const int* n_dev, const float* exits,
                                                        const int* texel_of, const float* pred_exit, float* mapping,
                                                        float* err_out, int write_mapping) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (n_dev ? *n_dev : n)) return;
    const int r = list[i];
    const float* e = exits + 6 * (size_t)r;
    if (err_out)
        err_out[i] = exit_error(pred_exit[3 * (size_t)r], pred_exit[3 * (size_t)r + 1], pred_exit[3 * (size_t)r + 2], e);
    if (write_mapping) reinterpret_cast<float4*>(mapping)[texel_of[r]] = make_float4(e[0], e[1], e[2], 1.0f);
}

// Accepted rays whose Lyapunov exponent (from the surrogate's own mapping) is above the
// threshold are queued for RK4 as well, and no longer count as accepted.
__global__ void __launch_bounds__(256) lyapunov_gate_kernel(int* accepted, const int* texel_of, int n,
                                                            const int* n_dev, const float* lyap, float threshold,
                                                            int* list, int* count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (n_dev ? *n_dev : n) || !accepted[i]) return;
    const float l = lyap[texel_of[i]];
    if (l > threshold) {
        accepted[i] = 0;
        int slot = atomicAdd(count, 1);
        list[slot] = i;
    }
}

// Per texel: 1 where the frame kept the surrogate's exit (mask cleared beforehand).
__global__ void __launch_bounds__(256) accepted_mask_kernel(const int* accepted, const int* texel_of, int n,
                                                            const int* n_dev, unsigned char* mask) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (n_dev ? *n_dev : n) && accepted[i]) mask[texel_of[i]] = 1;
}

__global__ void __launch_bounds__(256) fill_u8_kernel(unsigned char* p, int n, unsigned char value) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

__global__ void __launch_bounds__(256) fill_i32_kernel(int* p, int n, int value) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

// ---- launchers ------------------------------------------------------------------------------
hipError_t sg_gemm(const lm::GemmParams& P, int splits, hipStream_t s) {
    void* args[] = {const_cast<lm::GemmParams*>(&P)};
    if (splits == 1 && lm::gemm_nt_applies(P)) {  // the forward layout: 128 x 128 tiles, 16-B loads
        dim3 grid((P.M + LM_NT_BM - 1) / LM_NT_BM, P.N / LM_NT_BN, 1);
        return hipLaunchKernel((const void*)&gemm_nt_kernel, grid, dim3(256), args, 0, s);
    }
    dim3 grid((P.M + LM_BM - 1) / LM_BM, (P.N + LM_BN - 1) / LM_BN, splits);
    return hipLaunchKernel((const void*)&gemm_kernel, grid, dim3(256), args, 0, s);
}

#define LS_GRID(n) dim3(((n) + 255) / 256), dim3(256)

hipError_t sg_mlp_pack(const float* W, int K, u16* dst, hipStream_t s) {
    hipLaunchKernelGGL(mlp_pack_kernel, LS_GRID((K / 16) * 8 * 64), 0, s, W, K, dst);
    return hipGetLastError();
}
hipError_t sg_mlp_infer(const lm::MlpInferParams& P, hipStream_t s) {
    if (P.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_infer_kernel, dim3((P.n + LM_FUSED_ROWS - 1) / LM_FUSED_ROWS), dim3(256), 0, s, P);
    return hipGetLastError();
}
hipError_t sg_features(const ls::SurrogateIO& S, hipStream_t s) {
    hipLaunchKernelGGL(features_kernel, LS_GRID(S.n), 0, s, S);
    return hipGetLastError();
}
hipError_t sg_make_targets(const float* fine, const float* base, int n, float scale_p, float scale_d, float* tgt,
                           hipStream_t s) {
    hipLaunchKernelGGL(make_targets_kernel, LS_GRID(6 * n), 0, s, fine, base, n, scale_p, scale_d, tgt);
    return hipGetLastError();
}
hipError_t sg_residual_sumsq(const float* fine, const float* base, int n, float* acc, hipStream_t s) {
    hipLaunchKernelGGL(residual_sumsq_kernel, LS_GRID(n), 0, s, fine, base, n, acc);
    return hipGetLastError();
}
hipError_t sg_gather(const float* states, const float* targets, const float* fine, int n_data, unsigned long long seed,
                     int step, int batch, float* bstates, float* btargets, float* bfine, hipStream_t s) {
    hipLaunchKernelGGL(gather_batch_kernel, LS_GRID(batch), 0, s, states, targets, fine, n_data, seed, step, batch,
                       bstates, btargets, bfine);
    return hipGetLastError();
}
hipError_t sg_out_layer(const u16* act, const float* W, const float* b, int n, float* out, hipStream_t s) {
    hipLaunchKernelGGL(out_layer_kernel, LS_GRID(n * LS_OUT), 0, s, act, W, b, n, out);
    return hipGetLastError();
}
hipError_t sg_loss(const float* out, const float* target, const float* fine, int n, float scale_p, float tol,
                   float conf_weight, float* dout, float* err, float* loss_sum, hipStream_t s) {
    hipLaunchKernelGGL(loss_kernel, LS_GRID(n), 0, s, out, target, fine, n, scale_p, tol, conf_weight, dout, err,
                       loss_sum);
    return hipGetLastError();
}
hipError_t sg_out_layer_back(const float* dout, const float* W, const float* z, int n, u16* dz, hipStream_t s) {
    hipLaunchKernelGGL(out_layer_back_kernel, LS_GRID(n * LS_HIDDEN), 0, s, dout, W, z, n, dz);
    return hipGetLastError();
}
hipError_t sg_out_layer_grad(const u16* act, const float* dout, int n, float* dW, float* db, hipStream_t s) {
    hipLaunchKernelGGL(out_layer_grad_kernel, dim3(LS_HIDDEN + 1), dim3(256), 0, s, act, dout, n, dW, db);
    return hipGetLastError();
}
hipError_t sg_reduce(const float* part, int splits, int n, float* out, hipStream_t s) {
    hipLaunchKernelGGL(reduce_partials_kernel, LS_GRID(n), 0, s, part, splits, n, out);
    return hipGetLastError();
}
hipError_t sg_colsum(const u16* dz, int n, int cols, float* out, hipStream_t s) {
    hipLaunchKernelGGL(colsum_kernel, dim3(cols), dim3(256), 0, s, dz, n, cols, out);
    return hipGetLastError();
}
hipError_t sg_adam(float* w, float* m, float* v, const float* g, int K, int N, float lr, float b1, float b2, float eps,
                   float bc1, float bc2, u16* wt, int ldt, hipStream_t s) {
    hipLaunchKernelGGL(adam_kernel, LS_GRID(K * N), 0, s, w, m, v, g, K, N, lr, b1, b2, eps, bc1, bc2, wt, ldt);
    return hipGetLastError();
}
hipError_t sg_accept(const float* out, const float* base, const int* texel_of, int n, const int* n_dev, float scale_p,
                     float thr, float* mapping, float* pred_exit, int* fallback, int* n_fallback, int* accepted,
                     hipStream_t s) {
    hipLaunchKernelGGL(accept_kernel, LS_GRID(n), 0, s, out, base, texel_of, n, n_dev, scale_p, thr, mapping,
                       pred_exit, fallback, n_fallback, accepted);
    return hipGetLastError();
}
hipError_t sg_compact_entered(const int* entered, const float* entry_states, int W, int H, float* states,
                              int* texel_of, int* count, hipStream_t s) {
    const unsigned tiles = (unsigned)(((W + LINON_TILE - 1) / LINON_TILE) * ((H + LINON_TILE - 1) / LINON_TILE));
    hipLaunchKernelGGL(compact_entered_kernel, dim3((tiles + 3) / 4), dim3(256), 0, s, entered, entry_states, W, H,
                       states, texel_of, count);
    return hipGetLastError();
}
hipError_t sg_apply_rk4(const int* list, int n, const int* n_dev, const float* exits, const int* texel_of,
                        const float* pred_exit, float* mapping, float* err_out, int write_mapping, hipStream_t s) {
    hipLaunchKernelGGL(apply_rk4_kernel, LS_GRID(n), 0, s, list, n, n_dev, exits, texel_of, pred_exit, mapping, err_out,
                       write_mapping);
    return hipGetLastError();
}
hipError_t sg_accepted_mask(const int* accepted, const int* texel_of, int n, const int* n_dev, int npx,
                            unsigned char* mask, hipStream_t s) {
    hipLaunchKernelGGL(fill_u8_kernel, LS_GRID(npx), 0, s, mask, npx, (unsigned char)0);
    if (n > 0) hipLaunchKernelGGL(accepted_mask_kernel, LS_GRID(n), 0, s, accepted, texel_of, n, n_dev, mask);
    return hipGetLastError();
}
hipError_t sg_lyapunov_gate(int* accepted, const int* texel_of, int n, const int* n_dev, const float* lyap, float thr,
                            int* list, int* count, hipStream_t s) {
    hipLaunchKernelGGL(lyapunov_gate_kernel, LS_GRID(n), 0, s, accepted, texel_of, n, n_dev, lyap, thr, list, count);
    return hipGetLastError();
}
hipError_t sg_sample_states(unsigned long long seed, int n, float radius, float inset, float* states, hipStream_t s) {
    hipLaunchKernelGGL(sample_states_kernel, LS_GRID(n), 0, s, seed, n, radius, inset, states);
    return hipGetLastError();
}
hipError_t sg_pick_rejudge(const int* accepted, int n, const int* n_dev, int stride, unsigned long long seed, int* list,
                           int* count, hipStream_t s) {
    hipLaunchKernelGGL(pick_rejudge_kernel, LS_GRID(n), 0, s, accepted, n, n_dev, stride, seed, list, count);
    return hipGetLastError();
}
hipError_t sg_fill_i32(int* p, int n, int value, hipStream_t s) {
    hipLaunchKernelGGL(fill_i32_kernel, LS_GRID(n), 0, s, p, n, value);
    return hipGetLastError();
}

#define LS_FIELD_SWITCH(KERNEL)                                              \
    switch (field) {                                                         \
        case 17: return (const void*)&KERNEL<lw::FieldSwirl>;                \
        case 18: return (const void*)&KERNEL<lw::FieldLens>;                 \
        case 19: return (const void*)&KERNEL<lw::FieldHotAir>;               \
        case 20: return (const void*)&KERNEL<lw::FieldWarp>;                 \
        case LINON_KERNEL_HOT_AIR_FAITHFUL_ROOT:                             \
            return (const void*)&KERNEL<lw::FieldHotAirFaithfulRoot>;        \
        case LINON_KERNEL_LENS_RECIP:                                        \
            return (const void*)&KERNEL<lw::FieldLensRecip>;                 \
        case 3: return (const void*)&KERNEL<lw::FieldMirageSphericalSigmoid>; \
        case 16: return (const void*)&KERNEL<lw::FieldIdentity>;             \
        default: return nullptr;                                             \
    }

static const void* approach_for(int field) { LS_FIELD_SWITCH(approach_kernel) }
static const void* flowmap_for(int field) { LS_FIELD_SWITCH(flowmap_kernel) }

bool sg_field_supported(int field) { return approach_for(field) != nullptr; }

hipError_t sg_approach(int field, const ls::ApproachParams& A, hipStream_t s) {
    const void* fn = approach_for(field);
    if (!fn) return hipErrorInvalidDeviceFunction;
    const int tiles = A.tiles_x * ((A.cam.height + LINON_TILE - 1) / LINON_TILE);
    void* args[] = {const_cast<ls::ApproachParams*>(&A)};
    return hipLaunchKernel(fn, dim3((tiles + 3) / 4), dim3(256), args, 0, s);
}
hipError_t sg_flowmap(int field, const ls::FlowmapParams& P, hipStream_t s) {
    const void* fn = flowmap_for(field);
    if (!fn) return hipErrorInvalidDeviceFunction;
    if (P.n <= 0) return hipSuccess;
    void* args[] = {const_cast<ls::FlowmapParams*>(&P)};
    return hipLaunchKernel(fn, dim3((P.n + 255) / 256), dim3(256), args, 0, s);
}

}  // namespace LINON_KERNEL_NS
