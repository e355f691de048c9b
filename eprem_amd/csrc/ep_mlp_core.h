// ep_mlp_core.h -- the fused network evaluation shared by the inference
// kernel (k_mlp_fwd, ep_mlp.hip) and the surrogate EP update (ep_sur.hip).
//
// A wave evaluates kFNT tiles of 32 samples transposed, C^T = W^T X^T, on
// v_mfma_f32_32x32x16_bf16: lane (col, h) = (l % 32, l / 32) carries
// sample `col` of a tile and, as the B operand of K-step ks, its input
// features 16 ks + 8 h + q (q = 0..7); the accumulators it receives hold
// output features 8 (r / 4) + 4 h + r % 4 (r = 0..15) of each 32-feature
// tile, which after bias, ReLU and bf16 rounding are the next layer's B
// operand (the weight packing, k_mlp_pack, folds in the permutation).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace ep {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_mlp_t __attribute__((ext_vector_type(16)));

constexpr int kFH = 128;   // hidden width of the fused kernel
constexpr int kFNT = 2;    // 32-sample tiles per wave (one A fragment feeds two MFMAs)
constexpr int kFWaves = 4; // waves per block (one per SIMD; all share the LDS weights)

// One layer on NT sample tiles: acc[t][ot] = sum_ks A(ks, ot) B[t][ks].
template <int KS, int OT, int NT = kFNT>
__device__ __forceinline__ void mlp_layer(const bf16x8_t *__restrict__ frag, const bf16x8_t (&b)[NT][KS],
                                          f32x16_mlp_t (&acc)[NT][OT], int lane) {
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int ot = 0; ot < OT; ++ot)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[t][ot][r] = 0.0f;
#pragma unroll
  for (int ks = 0; ks < KS; ++ks)
#pragma unroll
    for (int ot = 0; ot < OT; ++ot) {
      const bf16x8_t a = frag[(ks * OT + ot) * 64 + lane];
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[t][ot] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b[t][ks], acc[t][ot], 0, 0, 0);
    }
}

// The same with the fragments read from LDS two products ahead and a
// scheduling barrier after each product: left alone the compiler loads
// every fragment of the layer up front (fast, but it then holds them all
// in registers and leaves no room for anything else in the kernel).
template <int KS, int OT, int NT>
__device__ __forceinline__ void mlp_layer_lean(const bf16x8_t *__restrict__ frag, const bf16x8_t (&b)[NT][KS],
                                               f32x16_mlp_t (&acc)[NT][OT], int lane) {
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int ot = 0; ot < OT; ++ot)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[t][ot][r] = 0.0f;
  constexpr int NF = KS * OT;
  bf16x8_t a0 = frag[lane];
  bf16x8_t a1 = frag[(NF > 1 ? 1 : 0) * 64 + lane];
#pragma unroll
  for (int i = 0; i < NF; ++i) {
    const int ks = i / OT, ot = i % OT;
    const bf16x8_t a = a0;
    a0 = a1;
    if (i + 2 < NF) a1 = frag[(i + 2) * 64 + lane];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t][ot] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b[t][ks], acc[t][ot], 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Bias, ReLU and bf16 rounding of a hidden layer's accumulators into the
// next layer's B operands: K-step 2 ot + u takes registers 8u..8u+7 of
// tile ot (features 32 ot + 16 u + 8 (q/4) + 4 h + q%4).
template <int NT = kFNT>
__device__ __forceinline__ void mlp_relu_to_b(const f32x16_mlp_t (&acc)[NT][kFH / 32], const float *bias, int h,
                                              bf16x8_t (&b)[NT][kFH / 16]) {
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int ot = 0; ot < kFH / 32; ++ot)
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        bf16x8_t v;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int r = 8 * u + q;
          float z = acc[t][ot][r] + bias[32 * ot + 8 * (r >> 2) + 4 * h + (r & 3)];
          z = z > 0.0f ? z : 0.0f;
          v[q] = (__bf16)z;
        }
        b[t][2 * ot + u] = v;
      }
}

// The fused network's packed weights (bf16 fragments, then biases) for
// the current parameters, packing them first if they changed; sizes = the
// layer widths.  Fails when no network is initialised or its shape is not
// one the fused kernel takes (input <= 128, hidden layers of 128, output
// <= 32).
hipError_t mlp_fused_weights(hipStream_t st, const uint4 **packed, int *packed16, int *bias16,
                             std::vector<int> *sizes);

}  // namespace ep
