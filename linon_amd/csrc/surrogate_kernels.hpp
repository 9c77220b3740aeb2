// Host-side launchers of the surrogate kernels (surrogate.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "device/mlp.hpp"
#include "device/mlp_fused.hpp"
#include "device/surrogate.hpp"

// Compiled twice like kernels.hip: LINON_KERNEL_NS is linon (IEEE) or linon_wgslmath (WGSL
// accuracy); the RK4 legs of a surrogate frame use the build the plain kernel of that field uses.
#ifndef LINON_KERNEL_NS
#define LINON_KERNEL_NS linon
#endif

#define LINON_DECLARE_SURROGATE_LAUNCHERS \
    bool sg_field_supported(int field); \
    hipError_t sg_gemm(const lm::GemmParams& P, int splits, hipStream_t s); \
    hipError_t sg_features(const ls::SurrogateIO& S, hipStream_t s); \
    hipError_t sg_mlp_pack(const float* W, int K, lm::u16* dst, hipStream_t s); \
    hipError_t sg_mlp_infer(const lm::MlpInferParams& P, hipStream_t s); \
    hipError_t sg_make_targets(const float* fine, const float* base, int n, float scale_p, float scale_d, float* tgt, \
                               hipStream_t s); \
    hipError_t sg_residual_sumsq(const float* fine, const float* base, int n, float* acc, hipStream_t s); \
    hipError_t sg_gather(const float* states, const float* targets, const float* fine, int n_data, \
                         unsigned long long seed, int step, int batch, float* bstates, float* btargets, float* bfine, \
                         hipStream_t s); \
    hipError_t sg_out_layer(const lm::u16* act, const float* W, const float* b, int n, float* out, hipStream_t s); \
    hipError_t sg_loss(const float* out, const float* target, const float* fine, int n, float scale_p, float tol, \
                       float conf_weight, float* dout, float* err, float* loss_sum, hipStream_t s); \
    hipError_t sg_out_layer_back(const float* dout, const float* W, const float* z, int n, lm::u16* dz, hipStream_t s); \
    hipError_t sg_out_layer_grad(const lm::u16* act, const float* dout, int n, float* dW, float* db, hipStream_t s); \
    hipError_t sg_reduce(const float* part, int splits, int n, float* out, hipStream_t s); \
    hipError_t sg_colsum(const lm::u16* dz, int n, int cols, float* out, hipStream_t s); \
    hipError_t sg_adam(float* w, float* m, float* v, const float* g, int K, int N, float lr, float b1, float b2, float eps, \
                       float bc1, float bc2, lm::u16* wt, int ldt, hipStream_t s); \
    hipError_t sg_accept(const float* out, const float* base, const int* texel_of, int n, const int* n_dev, \
                         float scale_p, float thr, float* mapping, float* pred_exit, int* fallback, int* n_fallback, \
                         int* accepted, hipStream_t s); \
    hipError_t sg_compact_entered(const int* entered, const float* entry_states, int W, int H, float* states, \
                                  int* texel_of, int* count, hipStream_t s); \
    hipError_t sg_apply_rk4(const int* list, int n, const int* n_dev, const float* exits, const int* texel_of, \
                            const float* pred_exit, float* mapping, float* err_out, int write_mapping, hipStream_t s); \
    hipError_t sg_accepted_mask(const int* accepted, const int* texel_of, int n, const int* n_dev, int npx, \
                                unsigned char* mask, hipStream_t s); \
    hipError_t sg_lyapunov_gate(int* accepted, const int* texel_of, int n, const int* n_dev, const float* lyap, \
                                float thr, int* list, int* count, hipStream_t s); \
    hipError_t sg_fill_i32(int* p, int n, int value, hipStream_t s); \
    hipError_t sg_sample_states(unsigned long long seed, int n, float radius, float inset, float* states, hipStream_t s); \
    hipError_t sg_pick_rejudge(const int* accepted, int n, const int* n_dev, int stride, unsigned long long seed, \
                               int* list, int* count, hipStream_t s); \
    hipError_t sg_approach(int field, const ls::ApproachParams& A, hipStream_t s); \
    hipError_t sg_flowmap(int field, const ls::FlowmapParams& P, hipStream_t s);

namespace linon {
LINON_DECLARE_SURROGATE_LAUNCHERS
}  // namespace linon
namespace linon_wgslmath {
LINON_DECLARE_SURROGATE_LAUNCHERS
}  // namespace linon_wgslmath
