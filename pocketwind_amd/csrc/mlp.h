// mlp.h -- the learned wall closure's network on gfx950 matrix cores.
//
//   h1 = silu(W1 x + b1), h2 = silu(W2 h1 + b2), y = W3 h2 + b3     (12 -> 64 -> 64 -> 3)
//
// Products run on v_mfma_f32_16x16x32_bf16 (bf16 operands, f32 accumulation): a wavefront
// multiplies a 16 x 32 tile of samples (rows) by a 32 x 16 tile of weights. The inputs, weights,
// hidden activations and backward deltas are rounded to bf16 (round to nearest even) where
// they enter a product; pre-activations, biases and all sums stay f32. The closure has no
// counterpart in the reference, whose wall treatment is the wall law of
// src/lbm/shaders/stream_collide.wgsl:1-80 (kept as this network's fallback).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pw {
namespace mlp {

constexpr int kIn = 12, kHidden = 64, kOut = 3;
constexpr int kInPad = 32, kOutPad = 16;
// Flat f32 parameter vector: W1 (64 x 12), b1, W2 (64 x 64), b2, W3 (3 x 64), b3.
constexpr int oW1 = 0, oB1 = oW1 + kHidden * kIn, oW2 = oB1 + kHidden, oB2 = oW2 + kHidden * kHidden,
              oW3 = oB2 + kHidden, oB3 = oW3 + kOut * kHidden, kParams = oB3 + kOut;

// Device copies the matrix cores read: bf16 bits, zero padded, plus transposes for the
// backward products.
struct Packed {
  uint16_t w1[kHidden * kInPad];    // [out][in, padded to 32]
  uint16_t w2[kHidden * kHidden];   // [out][in]
  uint16_t w2t[kHidden * kHidden];  // [in][out]
  uint16_t w3[kOutPad * kHidden];   // [out, padded to 16][in]
  uint16_t w3t[kHidden * kInPad];   // [in][out, padded to 32]
  float b1[kHidden], b2[kHidden], b3[kOutPad];
};

__host__ __device__ __forceinline__ uint16_t bf16_bits(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef uint16_t u16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// acc += A (16 x 32) B (32 x 16). A is row-major at `a` with row stride lda; B is given by its
// transpose `bt` (row n of bt is column n of B) with row stride ldb. Lane l supplies row l % 16
// of A and column l % 16 of B at k = 8 (l / 16) .. 8 (l / 16) + 7, and receives rows
// 4 (l / 16) + j, j = 0..3, of column l % 16 of the product.
__device__ __forceinline__ f32x4_t mma(const uint16_t* a, int lda, const uint16_t* bt, int ldb, f32x4_t acc) {
  const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
  const u16x8_t av = *reinterpret_cast<const u16x8_t*>(a + r * lda + 8 * g);
  const u16x8_t bv = *reinterpret_cast<const u16x8_t*>(bt + r * ldb + 8 * g);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, av), __builtin_bit_cast(bf16x8_t, bv),
                                                 acc, 0, 0, 0);
}

// Orders this wavefront's LDS writes before its later reads of other lanes' data.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float sigmoid(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float silu_prime(float z) {
  const float s = sigmoid(z);
  return s * (1.0f + z * (1.0f - s));
}

// Forward pass of 16 samples. xs: 16 x 32 bf16 in LDS (features 12..31 zero); h1, h2: 16 x 64
// bf16 LDS scratch of this wavefront. Lane l receives outputs of rows 4 (l / 16) + j, output
// l % 16 (meaningful for l % 16 < 3).
__device__ __forceinline__ f32x4_t forward16(const Packed* __restrict__ P, const uint16_t* xs, uint16_t* h1,
                                             uint16_t* h2) {
  const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
  const f32x4_t zero = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    const f32x4_t acc = mma(xs, kInPad, P->w1 + nb * 16 * kInPad, kInPad, zero);
#pragma unroll
    for (int j = 0; j < 4; ++j) h1[(4 * g + j) * kHidden + 16 * nb + r] = bf16_bits(silu(acc[j] + P->b1[16 * nb + r]));
  }
  wave_sync();
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    f32x4_t acc = zero;
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) acc = mma(h1 + 32 * kb, kHidden, P->w2 + nb * 16 * kHidden + 32 * kb, kHidden, acc);
#pragma unroll
    for (int j = 0; j < 4; ++j) h2[(4 * g + j) * kHidden + 16 * nb + r] = bf16_bits(silu(acc[j] + P->b2[16 * nb + r]));
  }
  wave_sync();
  f32x4_t y = zero;
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) y = mma(h2 + 32 * kb, kHidden, P->w3 + 32 * kb, kHidden, y);
#pragma unroll
  for (int j = 0; j < 4; ++j) y[j] += P->b3[r];
  return y;
}

}  // namespace mlp
}  // namespace pw
