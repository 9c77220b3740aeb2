// Fused inference of the surrogate MLP: Fourier features, the four hidden layers and the FP32
// output layer for a tile of 128 rays in one workgroup, activations kept in LDS between layers.
//
//   * 256 threads = 4 waves; wave w computes all 128 rays for columns 64 w .. +63 of each hidden
//     layer as 2 x 4 tiles of v_mfma_f32_32x32x16_bf16 (128 accumulators per lane), so every
//     weight fragment is loaded by one wave of the workgroup.
//   * Layer input: bf16 [128][264] in LDS (264 = 256 + 8 bf16 of padding, so consecutive rows
//     start 16 B apart in the bank space); each lane reads its 8 consecutive k of one ray with
//     one 16-byte LDS read.
//   * Weights: prepacked once per weight update in MFMA fragment order ([k/16][column
//     tile][lane][8 bf16], see mlp_pack_body), so a wave's fragment is one contiguous, coalesced
//     1 KB global load (L2 resident: 496 KB for all four layers), two k steps ahead of the
//     MFMAs that use it.
//   * Epilogue: bias + SiLU, rounded to bf16 into the same LDS tile after a barrier, 4 columns
//     of a ray per 8-byte LDS write.  The last hidden layer is followed by the FP32 output
//     layer (256 -> 7) from LDS.
//
// The arithmetic is the unfused path's (features_body, gemm_nt_body with the bias-SiLU
// epilogue, out_layer_kernel): the same roundings in the same order (k in ascending steps of 16
// into one accumulator, sequential FMA over k in the output layer), so both give the same bits.
#pragma once
#include "mlp.hpp"
#include "surrogate.hpp"

namespace lm {

#define LM_FUSED_ROWS 128
#define LM_FUSED_LDA (LS_HIDDEN + 8)
#define LM_FUSED_KSTEPS0 (LS_FEATS_PAD / 16)     // 14 k steps of the first layer
#define LM_FUSED_KSTEPS (LS_HIDDEN / 16)         // 16 k steps of the others
// bf16 elements of the packed weights of one layer
#define LM_PACK_LAYER0 (LS_FEATS_PAD * LS_HIDDEN)
#define LM_PACK_LAYER (LS_HIDDEN * LS_HIDDEN)
#define LM_PACK_TOTAL (LM_PACK_LAYER0 + 3 * LM_PACK_LAYER)

__host__ __device__ inline size_t pack_offset(int layer) {
    return layer == 0 ? 0 : (size_t)LM_PACK_LAYER0 + (size_t)(layer - 1) * LM_PACK_LAYER;
}

// Packs one hidden layer's FP32 master weights W [K][256] into MFMA B fragments: element e of
// fragment (ks, ct, lane) is bf16(W[16 ks + 8 (lane / 32) + e][32 ct + lane % 32]), i.e. what
// lane `lane` supplies as B (k 8 (lane / 32) .. +7 of column lane % 32) for the k step ks of
// column tile ct.  One thread per fragment.
__device__ __forceinline__ void mlp_pack_body(const float* W, int K, u16* dst) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (K / 16) * 8 * 64) return;
    const int lane = t & 63, ct = (t >> 6) & 7, ks = t >> 9;
    const int k0 = 16 * ks + 8 * (lane >> 5), n = 32 * ct + (lane & 31);
    for (int e = 0; e < 8; ++e) dst[(size_t)t * 8 + e] = f2bf(W[(size_t)(k0 + e) * LS_HIDDEN + n]);
}

struct MlpInferParams {
    const float* states;  // [n][6] entry states (p, q)
    int n;
    float radius;
    const float* freqs;   // [LS_FREQS]
    const u16* wpack;     // LM_PACK_TOTAL packed hidden weights
    const float* bias[4]; // hidden-layer biases [256]
    const float* w_out;   // [256][LS_OUT] FP32
    const float* b_out;   // [LS_OUT]
    float* out;           // [n][LS_OUT]
    const int* n_dev;     // optional: n on the device (then n is the launch's capacity)
};

typedef unsigned int u32x4v __attribute__((ext_vector_type(4)));

// The k loop of one hidden layer: this wave's weight fragments (column tiles 2 wave and
// 2 wave + 1 of each k step, 1 KB each) are loaded two k steps ahead of the MFMAs that use them
// (L2 latency against 8 MFMAs per step), into two register sets of two k steps each that the
// loop body, four k steps long, uses in turn (fixed registers: no copies between sets, and no
// wait before a fragment is used).  A k step count of 4 m + 2 (the first layer's 14) takes its
// first two steps before the loop.
template <int KSTEPS>
__device__ __forceinline__ void mlp_layer_k(const u16* wl, const u16 (*act)[LM_FUSED_LDA], int wave, int lane,
                                            f32x16 (&acc)[2][4]) {
    static_assert(KSTEPS % 2 == 0, "k steps come in pairs");
    auto load = [&](int ks, u32x4v (&w)[2][2]) {
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                w[h][j] = *reinterpret_cast<const u32x4v*>(wl + ((size_t)((ks + h) * 8 + 2 * wave + j) * 64 + lane) * 8);
        // keep the loads here: the scheduler would otherwise sink them next to their first use
        __builtin_amdgcn_sched_barrier(0);
    };
    auto step = [&](int ks, const u32x4v (&w)[2]) {
        bf16x8 x[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            x[i] = *reinterpret_cast<const bf16x8*>(&act[32 * i + (lane & 31)][16 * ks + 8 * (lane >> 5)]);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            bf16x8 wv;
            __builtin_memcpy(&wv, &w[j], 16);
#pragma unroll
            for (int i = 0; i < 4; ++i) acc[j][i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wv, x[i], acc[j][i], 0, 0, 0);
        }
    };
    constexpr int k0 = KSTEPS % 4;  // 0 or 2
    u32x4v wa[2][2], wb[2][2];
    if (k0) {
        u32x4v wc[2][2];
        load(0, wc);
        load(2, wa);
        step(0, wc[0]);
        step(1, wc[1]);
    } else {
        load(0, wa);
    }
#pragma nounroll
    for (int ks = k0; ks < KSTEPS; ks += 4) {
        load(ks + 2, wb);
        step(ks, wa[0]);
        step(ks + 1, wa[1]);
        // the last pass reloads its own fragments (never used): an unconditional load keeps the
        // count of loads in flight the same on every path into the next two steps, so their
        // waits are exact
        load(ks + 4 < KSTEPS ? ks + 4 : ks, wa);
        step(ks + 2, wb[0]);
        step(ks + 3, wb[1]);
    }
}

__device__ __forceinline__ void mlp_infer_body(const MlpInferParams& P) {
    __shared__ __attribute__((aligned(16))) u16 act[LM_FUSED_ROWS][LM_FUSED_LDA];
    __shared__ __attribute__((aligned(16))) float w5[LS_HIDDEN * 8];  // [k][8]: the 7 outputs and a zero
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int row0 = blockIdx.x * LM_FUSED_ROWS;
    const int n = P.n_dev ? *P.n_dev : P.n;
    if (row0 >= n) return;  // whole workgroup, before any barrier
    for (int i = tid; i < LS_HIDDEN * 8; i += blockDim.x) w5[i] = (i & 7) < LS_OUT ? P.w_out[(i >> 3) * LS_OUT + (i & 7)] : 0.0f;

    // ---- features: two threads per ray, eight frequencies each (features_body's arithmetic) ----
    {
        const int r = tid >> 1, half = tid & 1;
        const int g = row0 + r;
        u16* o = act[r];
        if (g < n) {
            float x[LS_IN];
            const float inv_r = 1.0f / P.radius;
            for (int k = 0; k < 3; ++k) x[k] = P.states[6 * (size_t)g + k] * inv_r;
            for (int k = 3; k < 6; ++k) x[k] = P.states[6 * (size_t)g + k];
            if (half == 0)
                for (int k = 0; k < LS_IN; ++k) o[k] = f2bf(x[k]);
            for (int f = 8 * half; f < 8 * half + 8; ++f) {
                const float w = P.freqs[f];
                for (int k = 0; k < LS_IN; ++k) {
                    float s, c;
                    __sincosf(w * x[k], &s, &c);
                    o[LS_IN + 2 * (f * LS_IN + k)] = f2bf(s);
                    o[LS_IN + 2 * (f * LS_IN + k) + 1] = f2bf(c);
                }
            }
            if (half == 1)
                for (int k = LS_FEATS; k < LS_FEATS_PAD; ++k) o[k] = 0;
        } else {
            for (int k = half; k < LS_FEATS_PAD; k += 2) o[k] = 0;
        }
    }
    __syncthreads();

    for (int l = 0; l < 4; ++l) {
        const u16* wl = P.wpack + pack_offset(l);
        // The weights are the MFMA's A operand and the layer input its B operand, so a lane's D
        // registers hold 4 consecutive output columns of one ray (stored as one 8-byte LDS
        // write): acc[jn][im] is the tile of columns 64 wave + 32 jn .. +31 and rays
        // 32 im .. +31.  Every wave reads its own weight columns (no two waves of a workgroup
        // load the same fragment) and all 128 rays from LDS.  The products and their k order
        // are those of the unswapped product (each D element sums the same 16 products per
        // instruction).
        f32x16 acc[2][4];
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[j][i][e] = 0.0f;
        if (l == 0)
            mlp_layer_k<LM_FUSED_KSTEPS0>(wl, act, wave, lane, acc);
        else
            mlp_layer_k<LM_FUSED_KSTEPS>(wl, act, wave, lane, acc);
        __syncthreads();  // every wave has read this layer's input
        // D of register e: column n = 8 (e / 4) + 4 (lane / 32) + e % 4 of the tile, ray
        // lane % 32: registers 4g .. 4g + 3 are columns n0 .. n0 + 3 of one ray
        const float* bias = P.bias[l];
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int n0 = 64 * wave + 32 * j + 8 * g + 4 * (lane >> 5);
                const float4 bv = *reinterpret_cast<const float4*>(bias + n0);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int m = 32 * i + (lane & 31);
                    const unsigned lo = f2bf2(silu(acc[j][i][4 * g] + bv.x), silu(acc[j][i][4 * g + 1] + bv.y));
                    const unsigned hi = f2bf2(silu(acc[j][i][4 * g + 2] + bv.z), silu(acc[j][i][4 * g + 3] + bv.w));
                    *reinterpret_cast<uint2*>(&act[m][n0]) = make_uint2(lo, hi);
                }
            }
        __syncthreads();
    }

    // ---- output layer (FP32, out_layer_kernel's order): two threads per ray ----
    {
        const int r = tid >> 1, half = tid & 1;
        const int g = row0 + r;
        if (g < n) {
            // outputs 0..3 (half 0) or 4..7 (half 1; the eighth, a zero column, is dropped)
            const int o0 = 4 * half;
            float s[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) s[q] = P.b_out[min(o0 + q, LS_OUT - 1)];
            for (int k8 = 0; k8 < LS_HIDDEN; k8 += 8) {
                const uint4 a8 = *reinterpret_cast<const uint4*>(&act[r][k8]);
                const unsigned aw[4] = {a8.x, a8.y, a8.z, a8.w};
#pragma unroll
                for (int kk = 0; kk < 8; ++kk) {
                    const float a = __uint_as_float(kk & 1 ? (aw[kk >> 1] & 0xffff0000u) : (aw[kk >> 1] << 16));
                    const float4 w = *reinterpret_cast<const float4*>(&w5[(k8 + kk) * 8 + o0]);
                    s[0] = __builtin_fmaf(a, w.x, s[0]);
                    s[1] = __builtin_fmaf(a, w.y, s[1]);
                    s[2] = __builtin_fmaf(a, w.z, s[2]);
                    s[3] = __builtin_fmaf(a, w.w, s[3]);
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (o0 + q < LS_OUT) P.out[(size_t)g * LS_OUT + o0 + q] = s[q];
        }
    }
}

}  // namespace lm
