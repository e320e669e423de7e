// cnn_mfma.hpp -- BF16 MFMA building blocks of the learned preconditioner (gfx950).
//
// A 3x3 periodic convolution of a channels-last activation u[site][C] is an implicit
// GEMM: out[site][co] = sum_k A[site][k] W[k][co] with k = tap * C + ci and
// A[site][k] = u[site + off(tap)][ci], tap = (dt+1)*3 + (dx+1).  One wave computes a
// 16-site x 16-channel output tile with v_mfma_f32_16x16x32_bf16 (FP32 accumulate),
// K padded to a multiple of 32 with zeros.
//
// Operand layouts of v_mfma_f32_16x16x32_bf16 (lane l = 0..63):
//   A (16 x 32): lane l holds A[l % 16][8 (l / 16) + i], i = 0..7
//   B (32 x 16): lane l holds B[8 (l / 16) + i][l % 16], i = 0..7
//   D (16 x 16): lane l holds D[4 (l / 16) + i][l % 16], i = 0..3
// Weights are pre-arranged into per-lane B fragments frag[ks][lane][8] so a lane
// loads its 16 bytes of B with one access.
#pragma once
#include <hip/hip_runtime.h>

namespace ycn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kCnnNW1 = 9 * 8 * 16, kCnnNW2 = 9 * 16 * 16, kCnnNW3 = 9 * 16 * 8;
constexpr int kCnnNWG = 9 * 16;                        // one gate layer: G[tap][16], input q = g phi
constexpr int kCnnOffG1 = kCnnNW1 + kCnnNW2 + kCnnNW3, kCnnOffG2 = kCnnOffG1 + kCnnNWG;
constexpr int kCnnNW = kCnnOffG2 + kCnnNWG;            // 4896 weights
constexpr int ksteps(int cin) { return (9 * cin + 31) / 32; }
// B-fragment sizes (bf16x8 per lane per K-step)
constexpr int kFrag1 = ksteps(8) * 64, kFrag2 = ksteps(16) * 64, kFrag3 = ksteps(16) * 64;

__device__ __forceinline__ f32x4 mfma_bf16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 bf16x8_zero() {
    bf16x8 v;
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = (__bf16)0.0f;
    return v;
}

// gate value 1 + tanh(t) = 2 / (1 + e^{-2t}) in [0, 2] (hardware exp and reciprocal,
// about 1 ulp each)
__device__ __forceinline__ float gate_fn(float t) { return 2.f * __builtin_amdgcn_rcpf(1.f + __expf(-2.f * t)); }

}  // namespace ycn
