// The learned step-size proposer's network, shared by the batched controller
// (inference) and the training kernels: MLP 8 -> 64 -> 64 -> 1 (tanh, tanh,
// linear), FP16 parameters, FP32 arithmetic.  No reference counterpart (an
// extension of the batched path).
//
// Parameter packing (4801 values): W1[64][8], b1[64], W2[64][64], b2[64],
// W3[64], b3.
#pragma once

#include <hip/hip_runtime.h>

#include "kmt_kernels.hpp"

namespace kmt {

constexpr int kIn = 8, kHid = 64;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kHid * kIn, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid,
              kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kHid, kNParams = kOffB3 + 1;
static_assert(kHid == kWave, "one lane per hidden unit");
static_assert(kOffW2 % 8 == 0 && kIn == 8, "16-byte rows of W1 and W2");

__device__ __forceinline__ float hf(uint16_t u) { return (float)__builtin_bit_cast(_Float16, u); }
__device__ __forceinline__ float hf_lo(unsigned u) { return hf((uint16_t)(u & 0xffffu)); }
__device__ __forceinline__ float hf_hi(unsigned u) { return hf((uint16_t)(u >> 16)); }

// Forward pass by one wave: lane i computes hidden unit i of both layers
// (its rows of W1 and W2 as 16-byte loads; the first layer's activations
// reach the other lanes by shuffles), the output is a wave sum.  Every lane
// returns the same output; h1 and h2 are lane i's activations.  The wave's
// 64 lanes must all be active.  w must be 16-byte aligned.
//   h1_i = tanh(b1_i + sum_j W1_ij x_j)        (j ascending, fma)
//   h2_i = tanh(b2_i + sum_j W2_ij h1_j)       (j ascending, fma)
//   out  = (wave sum of W3_i h2_i) + b3
__device__ __forceinline__ float mlp_wave_forward(const uint16_t *w, const float (&x)[kIn], float &h1, float &h2) {
  const int i = threadIdx.x % kWave;
  const u32x4 r1 = *reinterpret_cast<const u32x4 *>(w + kOffW1 + i * kIn);
  float a = hf(w[kOffB1 + i]);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    a = fmaf(hf_lo(r1[q]), x[2 * q], a);
    a = fmaf(hf_hi(r1[q]), x[2 * q + 1], a);
  }
  h1 = tanhf(a);
  a = hf(w[kOffB2 + i]);
  const u32x4 *row2 = reinterpret_cast<const u32x4 *>(w + kOffW2 + i * kHid);
#pragma unroll
  for (int c = 0; c < kHid / 8; ++c) {
    const u32x4 r = row2[c];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a = fmaf(hf_lo(r[q]), __shfl(h1, 8 * c + 2 * q, kWave), a);
      a = fmaf(hf_hi(r[q]), __shfl(h1, 8 * c + 2 * q + 1, kWave), a);
    }
  }
  h2 = tanhf(a);
  return wave_sum(hf(w[kOffW3 + i]) * h2) + hf(w[kOffB3]);
}

}  // namespace kmt
