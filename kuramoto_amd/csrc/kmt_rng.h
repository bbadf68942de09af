// Counter-based Philox4x32-10 (Salmon et al., SC'11) and the library's
// factor streams.  Host-side only: the factor X is drawn on the host in
// double precision (so its bits do not depend on device libm), uploaded, and
// the dense adjacency is assembled from it on the device with exact fma
// chains.  This replaces the reference's cuRAND Philox_4x32_10 draws in
// makeu1 / m1 / m2 (kuramoto.cu:62-101, :128-135), whose seeding/offset
// mapping is cuRAND-internal and is not reproduced.
#pragma once
#include <cmath>
#include <cstdint>

namespace kmt {

struct Philox4x32 {
  static constexpr uint32_t kMulA = 0xD2511F53u;
  static constexpr uint32_t kMulB = 0xCD9E8D57u;
  static constexpr uint32_t kBumpA = 0x9E3779B9u;  // golden ratio
  static constexpr uint32_t kBumpB = 0xBB67AE85u;  // sqrt(3) - 1

  // Ten rounds of the Philox S-P network on a 128-bit counter block.
  static void block(uint32_t x[4], uint32_t k0, uint32_t k1) {
    for (int round = 0; round < 10; ++round) {
      const uint64_t pa = (uint64_t)kMulA * x[0];
      const uint64_t pb = (uint64_t)kMulB * x[2];
      const uint32_t y0 = (uint32_t)(pb >> 32) ^ x[1] ^ k0;
      const uint32_t y1 = (uint32_t)pb;
      const uint32_t y2 = (uint32_t)(pa >> 32) ^ x[3] ^ k1;
      const uint32_t y3 = (uint32_t)pa;
      x[0] = y0; x[1] = y1; x[2] = y2; x[3] = y3;
      k0 += kBumpA;
      k1 += kBumpB;
    }
  }
};

// Stream identifiers and key tag of the library's factor draws.
constexpr uint32_t kKeyTag = 0x4B4D5441u;         // "KMTA"
constexpr uint32_t kStreamGauss = 0x58470000u;    // Gaussian factor
constexpr uint32_t kStreamSign = 0x52410000u;     // Rademacher factor

// Gaussian factor, X[n*K+m]: element e comes from counter e; two 53-bit
// uniforms u, v in (0,1) and the Box–Muller cosine branch
// sqrt(-2 ln u) cos(2 pi v).  Only log, sqrt and cos are used, so the bits do
// not depend on whether a compiler fuses sin and cos into sincos.
inline void gaussian_factor(long N, int K, uint32_t seed, double *X) {
  const uint64_t total = (uint64_t)N * (uint64_t)K;
  for (uint64_t e = 0; e < total; ++e) {
    uint32_t x[4] = {(uint32_t)e, (uint32_t)(e >> 32), kStreamGauss, 0u};
    Philox4x32::block(x, seed, kKeyTag);
    const uint64_t w0 = ((uint64_t)x[1] << 32) | x[0];
    const uint64_t w1 = ((uint64_t)x[3] << 32) | x[2];
    const double u = ((double)(w0 >> 11) + 0.5) * 0x1.0p-53;
    const double v = ((double)(w1 >> 11) + 0.5) * 0x1.0p-53;
    X[e] = std::sqrt(-2.0 * std::log(u)) * std::cos(2.0 * M_PI * v);
  }
}

// Rademacher factor: 128 signs per counter block, bit b -> element 128*blk+b.
inline void sign_factor(long N, int K, uint32_t seed, double *X) {
  const uint64_t total = (uint64_t)N * (uint64_t)K;
  for (uint64_t blk = 0; 128 * blk < total; ++blk) {
    uint32_t x[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), kStreamSign, 0u};
    Philox4x32::block(x, seed, kKeyTag);
    const uint64_t base = 128 * blk;
    for (int b = 0; b < 128 && base + b < total; ++b) X[base + b] = ((x[b / 32] >> (b % 32)) & 1u) ? 1.0 : -1.0;
  }
}

}  // namespace kmt
