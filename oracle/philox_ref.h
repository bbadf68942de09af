/* ORACLE (test infrastructure only) — never linked into the product.
 *
 * Philox4x32-10 counter-based generator, restated from the published
 * description in Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11 (section 4.2, constants in Table 2).  The
 * reference draws its adjacency bits from cuRAND's Philox4_32_10
 * (kuramoto.cu:57-60, :128-135); cuRAND's seeding/offset mapping is not
 * restated — the build uses its own counter layout (see oracle.hpp,
 * gaussian_factor / rademacher_factor).
 *
 * Pinned by the Random123 known-answer vectors (tests/test_oracle_generators.py).
 */
#ifndef KMT_ORACLE_PHILOX_REF_H
#define KMT_ORACLE_PHILOX_REF_H
#include <stdint.h>

static inline void orc_mulhilo32(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

/* One Philox4x32 bijection with R rounds: ctr (4 words), key (2 words). */
static inline void orc_philox4x32(const uint32_t ctr_in[4], const uint32_t key_in[2], int rounds, uint32_t out[4]) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;   /* multipliers */
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;   /* Weyl key increments */
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < rounds; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    orc_mulhilo32(M0, c0, &hi0, &lo0);
    orc_mulhilo32(M1, c2, &hi1, &lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += W0; k1 += W1;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
#endif
