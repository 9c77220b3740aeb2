// Fast path of the surrogate GEMM for the layout of every forward pass: A [M][K] row-major
// (activations) times B stored [N][K] (the BF16 weight copies), both k-contiguous.
//   * block tile 128 x 128, 4 waves, each wave 64 x 64 = 2 x 2 v_mfma_f32_32x32x16_bf16 tiles;
//   * k in steps of 32: every lane moves 2 x 16 B of A and 2 x 16 B of B per step (one 128-bit
//     global load each), staged in LDS as [row][k] with 8 bf16 of padding (80-B rows);
//   * LDS double buffer: the next step's tile is loaded into registers while the current one
//     feeds the MFMAs, then written to the other buffer (one barrier per step).
// Requires K % 32 == 0 and N % 128 == 0 (the MLP: K in {224, 256}, N = 256); rows >= M read as
// zero and are not stored.
#pragma once
#include "mlp.hpp"

namespace lm {

#define LM_NT_BM 128
#define LM_NT_BN 128
#define LM_NT_BK 32
#define LM_NT_LD (LM_NT_BK + 8)

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline bool gemm_nt_applies(const GemmParams& P) {
    return !P.a_trans && P.b_trans && (P.K % LM_NT_BK) == 0 && (P.N % LM_NT_BN) == 0 && P.k_per_split == P.K &&
           (P.epi == LM_EPI_STORE_F32 || P.epi == LM_EPI_BIAS_SILU) && (P.lda % 8) == 0 && (P.ldb % 8) == 0;
}

__device__ __forceinline__ void gemm_nt_body(const GemmParams& P) {
    __shared__ __attribute__((aligned(16))) u16 sA[2][LM_NT_BM][LM_NT_LD];
    __shared__ __attribute__((aligned(16))) u16 sB[2][LM_NT_BN][LM_NT_LD];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.x * LM_NT_BM, n0 = blockIdx.y * LM_NT_BN;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    // this lane's two 16-B chunks of each tile: rows r and r + 64, k offset kc
    const int r = tid >> 2, kc = (tid & 3) * 8;
    const u32x4 zero = {0u, 0u, 0u, 0u};
    auto load_a = [&](int k0, u32x4& x0, u32x4& x1) {
        const int g0 = m0 + r, g1 = m0 + r + 64;
        x0 = g0 < P.M ? *reinterpret_cast<const u32x4*>(P.A + (size_t)g0 * P.lda + k0 + kc) : zero;
        x1 = g1 < P.M ? *reinterpret_cast<const u32x4*>(P.A + (size_t)g1 * P.lda + k0 + kc) : zero;
    };
    auto load_b = [&](int k0, u32x4& y0, u32x4& y1) {
        y0 = *reinterpret_cast<const u32x4*>(P.B + (size_t)(n0 + r) * P.ldb + k0 + kc);
        y1 = *reinterpret_cast<const u32x4*>(P.B + (size_t)(n0 + r + 64) * P.ldb + k0 + kc);
    };
    f32x16 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;

    u32x4 a0, a1, b0, b1;
    load_a(0, a0, a1);
    load_b(0, b0, b1);
    int buf = 0;
    *reinterpret_cast<u32x4*>(&sA[0][r][kc]) = a0;
    *reinterpret_cast<u32x4*>(&sA[0][r + 64][kc]) = a1;
    *reinterpret_cast<u32x4*>(&sB[0][r][kc]) = b0;
    *reinterpret_cast<u32x4*>(&sB[0][r + 64][kc]) = b1;
    __syncthreads();
    for (int k0 = 0; k0 < P.K; k0 += LM_NT_BK) {
        const bool more = k0 + LM_NT_BK < P.K;
        if (more) {  // prefetch the next step into registers
            load_a(k0 + LM_NT_BK, a0, a1);
            load_b(k0 + LM_NT_BK, b0, b1);
        }
        for (int ks = 0; ks < LM_NT_BK; ks += 16) {
            const int kq = ks + 8 * (lane >> 5);
            bf16x8 fa[2], fb[2];
            for (int i = 0; i < 2; ++i)
                fa[i] = *reinterpret_cast<const bf16x8*>(&sA[buf][wm + 32 * i + (lane & 31)][kq]);
            for (int j = 0; j < 2; ++j)
                fb[j] = *reinterpret_cast<const bf16x8*>(&sB[buf][wn + 32 * j + (lane & 31)][kq]);
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
        if (more) {
            const int nb = buf ^ 1;
            *reinterpret_cast<u32x4*>(&sA[nb][r][kc]) = a0;
            *reinterpret_cast<u32x4*>(&sA[nb][r + 64][kc]) = a1;
            *reinterpret_cast<u32x4*>(&sB[nb][r][kc]) = b0;
            *reinterpret_cast<u32x4*>(&sB[nb][r + 64][kc]) = b1;
            __syncthreads();
            buf = nb;
        }
    }
    // epilogue: D[i][j] of register e at row 8*(e/4) + 4*(lane/32) + e%4, column lane%32
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            const int col = n0 + wn + 32 * j + (lane & 31);
            const float bias = P.bias ? P.bias[col] : 0.0f;
            for (int e = 0; e < 16; ++e) {
                const int row = m0 + wm + 32 * i + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
                if (row >= P.M) continue;
                const size_t o = (size_t)row * P.ldc + col;
                const float z = acc[i][j][e] + bias;
                if (P.epi == LM_EPI_STORE_F32) {
                    P.out_f32[o] = z;
                } else {
                    if (P.z_out) P.z_out[o] = z;
                    P.out_bf16[o] = f2bf(silu(z));
                }
            }
        }
}

}  // namespace lm
