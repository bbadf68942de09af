// The volcano factor of the reference's -A mode with any rank: X_nm = +-1,
// M = X diag(w) X^T with w_m = (-1)^m / N (makeu1 / makeu2 / m1 / m2,
// kuramoto.cu:62-143).  The reference regenerates the signs from Philox
// counters inside m1 / m2 and never stores them; here they are drawn once on
// the device from the library's own counter-based Philox stream (the same
// bits as the host generator sign_factor in kmt_rng.h: element e = n K + m is
// bit e mod 128 of block e / 128) and kept packed at one bit per entry, in
// two layouts:
//   bits  [N][ldw]      row n, word q: m = 32 q + b in bit b (zero past K)
//   bitsT [32 Kw][ldwT] row m, word q: n = 32 q + b in bit b (zero past N)
// (1.25 GB each at N = K = 10^5, against 40 GB for X in FP32).
//
// Coupling (one RHS, four launches):
//   sign_sum_kernel  pass 1 over bits:  partial sums of sum_n X_nm v_n for
//                    every m (v = cos, sin), per range of n
//   sign_gram_kernel G_m = w_m * (sum of the partials), zero past K
//   sign_sum_kernel  pass 2 over bitsT: partial sums of sum_m X_nm G_m
//   sign_stage_kernel C_n, S_n = sum of the partials, then the stage epilogue
// Every sum is in double and runs in a fixed order (deterministic), and the
// sums are those of the oracle's factored path (X_nm v_n and X_nm G_m are
// +-v_n and +-G_m exactly), so only the order of the double additions
// differs from it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kmt_kernels.hpp"

namespace kmt {

constexpr uint32_t kSignKeyTag = 0x4B4D5441u;     // the host generator's key tag
constexpr uint32_t kSignStream = 0x52410000u;     // its Rademacher stream

// Philox4x32-10 (Salmon et al., SC'11), the host generator's round function.
__device__ __forceinline__ void philox10(uint32_t x[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    const uint32_t hb = __umulhi(0xCD9E8D57u, x[2]), lb = 0xCD9E8D57u * x[2];
    const uint32_t ha = __umulhi(0xD2511F53u, x[0]), la = 0xD2511F53u * x[0];
    x[0] = hb ^ x[1] ^ k0;
    x[1] = lb;
    x[2] = ha ^ x[3] ^ k1;
    x[3] = la;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Word W of the sign stream: elements 32 W .. 32 W + 31 (bit b -> element 32 W + b).
__device__ __forceinline__ uint32_t sign_stream_word(uint64_t W, uint32_t seed) {
  uint32_t x[4] = {(uint32_t)(W >> 2), (uint32_t)(W >> 34), kSignStream, 0u};
  philox10(x, seed, kSignKeyTag);
  const uint32_t sel = (uint32_t)(W & 3);
  return sel == 0 ? x[0] : sel == 1 ? x[1] : sel == 2 ? x[2] : x[3];
}

// bits[n][q] = signs of m = 32 q .. 32 q + 31 (m < K) of row n: stream
// elements n K + 32 q + b, i.e. 32 consecutive stream bits from position
// e0 = n K + 32 q (at most two stream words).
[[maybe_unused]] static __global__ __launch_bounds__(kBlock) void gen_sign_bits_kernel(uint32_t *bits, long N, long K, long ldw,
                                                                uint32_t seed) {
  const long idx = (long)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= N * ldw) return;
  const long n = idx / ldw, q = idx % ldw;
  uint32_t word = 0;
  if (32 * q < K) {
    const long nb = K - 32 * q < 32 ? K - 32 * q : 32;
    const uint64_t e0 = (uint64_t)n * (uint64_t)K + (uint64_t)(32 * q);
    const uint32_t sh = (uint32_t)(e0 & 31);
    word = sign_stream_word(e0 >> 5, seed) >> sh;
    if (sh != 0 && sh + nb > 32) word |= sign_stream_word((e0 >> 5) + 1, seed) << (32 - sh);
    if (nb < 32) word &= (1u << nb) - 1u;
  }
  bits[idx] = word;
}

// bitsT from bits, 32 x 64 bit blocks: lane l holds row n0 + l's word q; the
// ballot of bit b over the wave is column m = 32 q + b for those 64 rows,
// i.e. two words of bitsT row m.  Lane b (< 32) writes row m = 32 q + b.
[[maybe_unused]] static __global__ __launch_bounds__(kBlock) void transpose_sign_bits_kernel(const uint32_t *bits, long N, long ldw, long Kw,
                                                                      uint32_t *bitsT, long ldwT) {
  const long wave = ((long)blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const long nblk = (N + 63) / 64;
  if (wave >= nblk * Kw) return;  // whole waves: the ballot needs every lane
  const long q = wave / nblk, n0 = (wave % nblk) * 64;
  const uint32_t w = (n0 + lane < N) ? bits[(n0 + lane) * ldw + q] : 0u;
  uint32_t lo = 0, hi = 0;
#pragma unroll
  for (int b = 0; b < 32; ++b) {
    const unsigned long long m = __ballot((w >> b) & 1u);
    if (lane == b) {
      lo = (uint32_t)m;
      hi = (uint32_t)(m >> 32);
    }
  }
  if (lane < 32) {
    uint32_t *dst = bitsT + (32 * q + lane) * ldwT + n0 / 32;
    dst[0] = lo;
    if (n0 / 32 + 1 < ldwT) dst[1] = hi;
  }
}

// Partial signed sums over a range of outer rows o (the rows of `bits`):
//   part[split][32 q + b][0/1] = sum_o (bit b of bits[o][q] ? +1 : -1) * (vc_o, vs_o)
// One thread per word q (32 inner indices, 64 double accumulators); the
// outer vector is the same in every lane.  Formed as 2 P_b - T with
// P_b = sum_o bit_b(o) v_o (an fma by the bit as 0.0 / 1.0: one bit-field
// extract, one conversion and two fmas per entry) and T = sum_o v_o, shared
// by the thread's 32 inner indices.  All in double; against the oracle's
// sequential sum of +-v_o only the rounding of the double additions
// differs (relative 1e-16 times the cancellation, ~sqrt of the row count).
template <class R, class V>
__global__ __launch_bounds__(kBlock) void sign_sum_kernel(const uint32_t *__restrict__ bits, long ldw, long nwords,
                                                           long n_outer, long per_split, const V *__restrict__ vc,
                                                           const V *__restrict__ vs, long vstride,
                                                           double *__restrict__ part, long ldp, const State<R> *st,
                                                           int check_state) {
  if (check_state && (!st->active || st->commit_pending)) return;
  const long q = (long)blockIdx.x * kBlock + threadIdx.x;
  if (q >= nwords) return;
  const long o0 = (long)blockIdx.y * per_split;
  const long o1 = o0 + per_split < n_outer ? o0 + per_split : n_outer;
  double pc[32], ps[32], tc = 0.0, ts = 0.0;
#pragma unroll
  for (int b = 0; b < 32; ++b) { pc[b] = 0.0; ps[b] = 0.0; }
  auto row = [&](uint32_t wd, double c, double s) {
    tc += c;
    ts += s;
#pragma unroll
    for (int b = 0; b < 32; ++b) {
      const double x = (double)((wd >> b) & 1u);
      pc[b] = __builtin_fma(x, c, pc[b]);
      ps[b] = __builtin_fma(x, s, ps[b]);
    }
  };
  constexpr int U = 4;  // outer rows per batch of loads
  long o = o0;
  for (; o + U <= o1; o += U) {
    uint32_t wd[U];
    double c[U], s[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      wd[u] = __builtin_nontemporal_load(bits + (o + u) * ldw + q);
      c[u] = (double)vc[(o + u) * vstride];
      s[u] = (double)vs[(o + u) * vstride];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) row(wd[u], c[u], s[u]);
  }
  for (; o < o1; ++o) row(bits[o * ldw + q], (double)vc[o * vstride], (double)vs[o * vstride]);
  double *dst = part + ((size_t)blockIdx.y * ldp + 32 * q) * 2;
#pragma unroll
  for (int b = 0; b < 32; ++b) {
    dst[2 * b] = 2.0 * pc[b] - tc;
    dst[2 * b + 1] = 2.0 * ps[b] - ts;
  }
}

// G_m = w_m * sum over the splits of the pass-1 partials (in split order),
// interleaved (G_c, G_s); zero for m >= K.
template <class R>
__global__ __launch_bounds__(kBlock) void sign_gram_kernel(const double *part, int nsplit, long ldp, const R *w,
                                                            long K, double *G, const State<R> *st, int check_state) {
  if (check_state && (!st->active || st->commit_pending)) return;
  const long m = (long)blockIdx.x * kBlock + threadIdx.x;
  if (m >= ldp) return;
  double c = 0.0, s = 0.0;
  if (m < K) {
    for (int sp = 0; sp < nsplit; ++sp) {
      c += part[((size_t)sp * ldp + m) * 2];
      s += part[((size_t)sp * ldp + m) * 2 + 1];
    }
    c *= (double)w[m];
    s *= (double)w[m];
  }
  G[2 * m] = c;
  G[2 * m + 1] = s;
}

// C_n, S_n = sum over the splits of the pass-2 partials, then the stage
// epilogue of row n; stage 7 writes one error partial per workgroup.
template <class R, int KIND>
__global__ __launch_bounds__(kBlock) void sign_stage_kernel(const double *part, int nsplit, long ldp, Vecs<R> v) {
  if constexpr (KIND >= 2) {
    if (v.check_state && (!v.st->active || v.st->commit_pending)) return;
  }
  __shared__ double errsq[kBlock];
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  double sq = 0.0;
  if (i < v.N) {
    double C = 0.0, S = 0.0;
    for (int sp = 0; sp < nsplit; ++sp) {
      C += part[((size_t)sp * ldp + i) * 2];
      S += part[((size_t)sp * ldp + i) * 2 + 1];
    }
    const R h = (KIND >= 2) ? v.st->h : (R)0;
    sq = row_epilogue<R, KIND>(v, i, (R)C, (R)S, h);
  }
  if constexpr (KIND == 7) {
    const double tot = block_sum(sq, errsq);
    if (threadIdx.x == 0) v.err_part[blockIdx.x] = tot;
  }
}

// Dense rows from the packed signs with the dense generator's chain
// (gen_adj_kernel): M_ij = sum_m fma(w_m, X_im X_jm, acc) in the working
// precision, X_im X_jm = +1 where the two bits agree, -1 where they differ.
// A workgroup owns ROWS rows and 256 columns; each thread reads its column's
// bit row one word (32 m) at a time, the rows' words are the same in every
// lane.
template <class R, class S, int ROWS>
__global__ __launch_bounds__(kBlock) void gen_adj_sign_kernel(const uint32_t *bits, long ldw, const R *w, long K,
                                                               long N, long row0, long nrows, S *A, long lda,
                                                               R inv_scale) {
  const long i0 = row0 + (long)blockIdx.y * ROWS;
  const long j = (long)blockIdx.x * kBlock + threadIdx.x;
  const long jj = j < N ? j : N - 1;
  R acc[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) acc[r] = (R)0;
  const uint32_t *bj = bits + jj * ldw;
  const long kw = (K + 31) / 32;
  for (long q = 0; q < kw; ++q) {
    const uint32_t wj = bj[q];
    uint32_t wi[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const long i = i0 + r < N ? i0 + r : N - 1;
      wi[r] = bits[i * ldw + q];
    }
    const int nb = K - 32 * q < 32 ? (int)(K - 32 * q) : 32;
    for (int b = 0; b < nb; ++b) {
      const R wm = w[32 * q + b];
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        const R x = (((wi[r] ^ wj) >> b) & 1u) ? (R)-1 : (R)1;
        acc[r] = rfma<R>(wm, x, acc[r]);
      }
    }
  }
  if (j < N) {
#pragma unroll
    for (int r = 0; r < ROWS; ++r)
      if (i0 + r < row0 + nrows && i0 + r < N)
        A[(size_t)(i0 + r - row0) * lda + j] = store_value<R, S>(acc[r], inv_scale);
  }
}

}  // namespace kmt
