// Fused inference of the surrogate MLP: Fourier features, the four hidden layers and the FP32
// output layer for a tile of 128 rays in one workgroup, activations kept in LDS between layers.
//
//   * 256 threads = 4 waves in a 2 x 2 arrangement: wave (wr, wc) computes rows 64 wr .. +63 and
//     columns 128 wc .. +127 of each hidden layer as 2 x 4 tiles of v_mfma_f32_32x32x16_bf16
//     (128 accumulators per lane).
//   * A operand: the layer input, bf16 [128][264] in LDS (264 = 256 + 8 bf16 of padding, so
//     consecutive rows start 16 B apart in the bank space); each lane reads its 8 consecutive k
//     of one row with one 16-byte LDS read.
//   * B operand: the weights, prepacked once per weight update in MFMA fragment order
//     ([k/16][column tile][lane][8 bf16], see mlp_pack_body), so a wave's fragment is one
//     contiguous, coalesced 1 KB global load (L2 resident: 496 KB for all four layers).  The
//     next k step's fragments are loaded while the current step's MFMAs run.
//   * Epilogue: bias + SiLU, rounded to bf16 into the same LDS tile after a barrier.  The last
//     hidden layer is followed by the FP32 output layer (256 -> 7) from LDS.
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
};

__device__ __forceinline__ void mlp_infer_body(const MlpInferParams& P) {
    __shared__ __attribute__((aligned(16))) u16 act[LM_FUSED_ROWS][LM_FUSED_LDA];
    __shared__ float w5[LS_HIDDEN * LS_OUT + 1];  // + 1: the zero the fourth sum of half 1 reads last
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wr = wave >> 1, wc = wave & 1;
    const int row0 = blockIdx.x * LM_FUSED_ROWS;
    for (int i = tid; i < LS_HIDDEN * LS_OUT; i += blockDim.x) w5[i] = P.w_out[i];
    if (tid == 0) w5[LS_HIDDEN * LS_OUT] = 0.0f;

    // ---- features: two threads per ray, eight frequencies each (features_body's arithmetic) ----
    {
        const int r = tid >> 1, half = tid & 1;
        const int g = row0 + r;
        u16* o = act[r];
        if (g < P.n) {
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

    typedef unsigned int u32x4v __attribute__((ext_vector_type(4)));
    for (int l = 0; l < 4; ++l) {
        const int ksteps = l == 0 ? LM_FUSED_KSTEPS0 : LM_FUSED_KSTEPS;
        const u16* wl = P.wpack + pack_offset(l);
        f32x16 acc[2][4];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;
        // this wave's fragments of k step ks: tiles 4 wc .. 4 wc + 3, 1 KB each
        auto frag_ptr = [&](int ks, int j) {
            return reinterpret_cast<const u32x4v*>(wl + ((size_t)(ks * 8 + 4 * wc + j) * 64 + lane) * 8);
        };
        u32x4v bn[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) bn[j] = *frag_ptr(0, j);
        for (int ks = 0; ks < ksteps; ++ks) {
            u32x4v bc[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) bc[j] = bn[j];
            if (ks + 1 < ksteps)
#pragma unroll
                for (int j = 0; j < 4; ++j) bn[j] = *frag_ptr(ks + 1, j);
            bf16x8 a[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                a[i] = *reinterpret_cast<const bf16x8*>(&act[64 * wr + 32 * i + (lane & 31)][16 * ks + 8 * (lane >> 5)]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    bf16x8 b;
                    __builtin_memcpy(&b, &bc[j], 16);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b, acc[i][j], 0, 0, 0);
                }
        }
        __syncthreads();  // every wave has read this layer's input
        // D of register e: row 8 (e / 4) + 4 (lane / 32) + e % 4, column lane % 32
        const float* bias = P.bias[l];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = 128 * wc + 32 * j + (lane & 31);
            const float bv = bias[col];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const int row = 64 * wr + 32 * i + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
                    act[row][col] = f2bf(silu(acc[i][j][e] + bv));
                }
        }
        __syncthreads();
    }

    // ---- output layer (FP32, out_layer_kernel's order): two threads per ray ----
    {
        const int r = tid >> 1, half = tid & 1;
        const int g = row0 + r;
        if (g < P.n) {
            // outputs 0..3 (half 0) or 4..6 (half 1; its fourth sum is computed and dropped)
            const int o0 = half ? 4 : 0;
            // (the fourth sum of half 1 reads row k + 1's first weight, and the pad after the last row)
            float s[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) s[q] = P.b_out[min(o0 + q, LS_OUT - 1)];
            for (int k = 0; k < LS_HIDDEN; ++k) {
                const float a = bf2f(act[r][k]);
                const float* wk = w5 + k * LS_OUT + o0;
#pragma unroll
                for (int q = 0; q < 4; ++q) s[q] = __builtin_fmaf(a, wk[q], s[q]);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (o0 + q < LS_OUT) P.out[(size_t)g * LS_OUT + o0 + q] = s[q];
        }
    }
}

}  // namespace lm
