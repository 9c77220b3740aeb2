// BF16 MFMA building blocks of the flow-map surrogate (a Fourier-feature MLP):
//   * gemm_body: C[M x N] = A[M x K] * B[K x N] on v_mfma_f32_32x32x16_bf16 with FP32
//     accumulation.  A 64x64 block tile, 4 waves each owning one 32x32 tile, K in steps of 32
//     staged through LDS.  LDS holds A as [m][k] and B as [n][k] (k contiguous), so each lane
//     reads its 8 consecutive k-values of one row / column with a single 16-byte load; the
//     global->LDS copy absorbs either storage order of A and B (row-major or transposed).
//   * fused epilogues: bias + SiLU (+ the pre-activation kept for the backward pass), plain
//     bias (output layer), SiLU' gating of the backward pass, and split-K partial sums for
//     weight gradients.
// Operand layout of v_mfma_f32_32x32x16_bf16 for lane l: A row (l % 32), k 8*(l/32)..+7; B
// column (l % 32), same k; D[i][j] in VGPR r: j = l % 32, i = 8*(r/4) + 4*(l/32) + r%4.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

namespace lm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned short u16;

#define LM_BM 64
#define LM_BN 64
#define LM_BK 32

// epilogue kinds
#define LM_EPI_STORE_F32 0      // C (+bias) -> fp32 out
#define LM_EPI_BIAS_SILU 1      // Z = C + bias -> z_out (fp32), A = silu(Z) -> bf16 out
#define LM_EPI_SILU_GRAD 2      // dZ = C * silu'(Z_in) -> bf16 out (and fp32 copy when out_f32)
#define LM_EPI_PARTIAL 3        // split-K partial: fp32 out + blockIdx.z * M * N

struct GemmParams {
    const u16* A;   // bf16 bits
    const u16* B;
    int M, N, K;
    int lda, ldb;   // leading dimensions of the stored matrices
    int a_trans;    // 0: A stored [M][K] (row-major); 1: stored [K][M]
    int b_trans;    // 0: B stored [K][N]; 1: stored [N][K]
    int k_per_split;  // split-K chunk (multiple of LM_BK); = K when no split
    int epi;
    const float* bias;   // [N]
    const float* z_in;   // [M][N] fp32 pre-activations (LM_EPI_SILU_GRAD)
    float* out_f32;      // [M][N] (or partials)
    u16* out_bf16;       // [M][N]
    float* z_out;        // [M][N] fp32 pre-activations (LM_EPI_BIAS_SILU)
    int ldc;
};

// float -> bf16 bits, rounded to nearest even (one v_cvt_pk_bf16_f32 on gfx950; NaN stays NaN)
__device__ __forceinline__ u16 f2bf(float x) {
    const __bf16 b = (__bf16)x;
    u16 u;
    __builtin_memcpy(&u, &b, 2);
    return u;
}
// two floats -> packed bf16 pair (a in the low half), one conversion instruction
__device__ __forceinline__ unsigned f2bf2(float a, float b) {
    typedef float f32x2v __attribute__((ext_vector_type(2)));
    typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
    const bf16x2v v = __builtin_convertvector((f32x2v){a, b}, bf16x2v);
    unsigned u;
    __builtin_memcpy(&u, &v, 4);
    return u;
}
__device__ __forceinline__ float bf2f(u16 b) { return __uint_as_float(((unsigned)b) << 16); }
// SiLU and its derivative through the hardware reciprocal (v_rcp_f32, 1 ulp): a network
// needs no correctly rounded division, and every kernel (training GEMM epilogues, the fused
// inference) evaluates these same expressions.
__device__ __forceinline__ float sigmoid_fast(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigmoid_fast(z); }
__device__ __forceinline__ float silu_grad(float z) {
    float s = sigmoid_fast(z);
    return s * (1.0f + z * (1.0f - s));
}

// One 64x64 output tile per 256-thread block; blockIdx.x -> M tiles, blockIdx.y -> N tiles,
// blockIdx.z -> split-K chunk.  Rows/cols beyond M/N and k beyond K read as zero.
__device__ __forceinline__ void gemm_body(const GemmParams& P) {
    __shared__ __attribute__((aligned(16))) u16 sA[LM_BM][LM_BK + 8];  // +8: stagger LDS banks
    __shared__ __attribute__((aligned(16))) u16 sB[LM_BN][LM_BK + 8];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.x * LM_BM, n0 = blockIdx.y * LM_BN;
    const int kbeg = blockIdx.z * P.k_per_split;
    const int kend = min(P.K, kbeg + P.k_per_split);
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;  // this wave's 32x32 sub-tile
    f32x16 acc;
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;

    for (int k0 = kbeg; k0 < kend; k0 += LM_BK) {
        // ---- global -> LDS: 64 x 32 elements of A and of B, 8 per thread ----
        if (!P.a_trans) {
            const int r = tid >> 2, kc = (tid & 3) * 8;
            const int gm = m0 + r;
            for (int e = 0; e < 8; ++e) {
                const int gk = k0 + kc + e;
                sA[r][kc + e] = (gm < P.M && gk < kend) ? P.A[(size_t)gm * P.lda + gk] : (u16)0;
            }
        } else {  // stored [K][M]: read along m, scatter into [m][k]
            const int kk = tid >> 3, mc = (tid & 7) * 8;
            const int gk = k0 + kk;
            for (int e = 0; e < 8; ++e) {
                const int gm = m0 + mc + e;
                sA[mc + e][kk] = (gm < P.M && gk < kend) ? P.A[(size_t)gk * P.lda + gm] : (u16)0;
            }
        }
        if (P.b_trans) {  // stored [N][K]
            const int r = tid >> 2, kc = (tid & 3) * 8;
            const int gn = n0 + r;
            for (int e = 0; e < 8; ++e) {
                const int gk = k0 + kc + e;
                sB[r][kc + e] = (gn < P.N && gk < kend) ? P.B[(size_t)gn * P.ldb + gk] : (u16)0;
            }
        } else {  // stored [K][N]: read along n, scatter into [n][k]
            const int kk = tid >> 3, nc = (tid & 7) * 8;
            const int gk = k0 + kk;
            for (int e = 0; e < 8; ++e) {
                const int gn = n0 + nc + e;
                sB[nc + e][kk] = (gn < P.N && gk < kend) ? P.B[(size_t)gk * P.ldb + gn] : (u16)0;
            }
        }
        __syncthreads();
        // ---- two MFMA k-steps of 16 ----
        for (int ks = 0; ks < LM_BK; ks += 16) {
            const int kq = ks + 8 * (lane >> 5);
            bf16x8 a = *reinterpret_cast<const bf16x8*>(&sA[wm + (lane & 31)][kq]);
            bf16x8 b = *reinterpret_cast<const bf16x8*>(&sB[wn + (lane & 31)][kq]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }

    // ---- epilogue ----
    const int j = n0 + wn + (lane & 31);
    for (int r = 0; r < 16; ++r) {
        const int i = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (i >= P.M || j >= P.N) continue;
        float c = acc[r];
        const size_t o = (size_t)i * P.ldc + j;
        switch (P.epi) {
            case LM_EPI_STORE_F32:
                P.out_f32[o] = c + (P.bias ? P.bias[j] : 0.0f);
                break;
            case LM_EPI_BIAS_SILU: {
                float z = c + P.bias[j];
                if (P.z_out) P.z_out[o] = z;
                P.out_bf16[o] = f2bf(silu(z));
                break;
            }
            case LM_EPI_SILU_GRAD: {
                float g = c * silu_grad(P.z_in[o]);
                P.out_bf16[o] = f2bf(g);
                if (P.out_f32) P.out_f32[o] = g;
                break;
            }
            default:  // LM_EPI_PARTIAL
                P.out_f32[(size_t)blockIdx.z * P.M * P.N + (size_t)i * P.N + j] = c;
                break;
        }
    }
}

}  // namespace lm
