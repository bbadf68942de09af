// The volcano factor of the reference's -A mode with any rank: X_nm = +-1,
// M = X diag(w) X^T with w_m = (-1)^m / N (makeu1 / makeu2 / m1 / m2,
// kuramoto.cu:62-143).  The reference regenerates the signs from Philox
// counters inside m1 / m2 and never stores them; here they are drawn once on
// the device from the library's own counter-based Philox stream (the same
// bits as the host generator sign_factor in kmt_rng.h: element e = n K + m is
// bit e mod 128 of block e / 128) and kept packed at one bit per entry, once
// per int8 pass, in that pass's tile order (sign_tile_word):
//   pass 2 tiles: row n, word q: m = 32 q + b in bit b (zero past K)
//   pass 1 tiles: row m, word q: n = 32 q + b in bit b (zero past N)
// (1.25 GB each at N = K = 10^5, against 40 GB for X in FP32), each with one
// more row of ones (k < the pass's k range) and zero rows up to its padded
// row count.
//
// Coupling (one RHS, five launches), on the int8 matrix cores with exact
// integer sums:
//   sign_limbs_kernel  (cos, sin)_n as 56-bit fixed point (2^-54 units) in
//                      7 signed base-256 digits each: the B operand of pass 1
//   sign_mfma_kernel   pass 1 over the pass-1 tiles: P_m = sum_n bit(m, n) * digit(n),
//                      per digit and per range of n, plus T = sum_n digit(n)
//   sign_gram_kernel   sum_n X_nm v_n = 2 P_m - T (X = 2 bit - 1) combined to
//                      double, G_m = w_m * that, and G as digits: the B
//                      operand of pass 2
//   sign_mfma_kernel   pass 2 over the pass-2 tiles: the same with rows n, k = m
//   sign_stage_kernel  C_n, S_n = 2 P_n - T combined, then the stage epilogue
// Each digit's sum over k is exact in int32 (|sum| <= 128 k < 2^31 for
// k < 2^24), so every coupling sum is the exact sum of the fixed-point
// inputs rounded once to double: of the oracle's factored path (kuramoto.cu:
// 62-101 restated) it differs only by the inputs' rounding to 2^-54 (cos,
// sin and |G_m| <= 1, as w_m = +-1/N) and by the oracle's own rounding of its
// sequential double additions.  Deterministic.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

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

// Tile order of a pass's bits: rows in groups of 64 (one wave's four 16-row
// MFMA tiles), words in pairs of 32 (1024 k, two super-blocks); the group's
// 64 rows x 32 words of pair P are 8 KB of consecutive memory, in the order
// one wave loads them: load (t, s) of lane (g, rl) is the 16 bytes of row
// 16 t + rl, words 32 P + 8 g + 4 s .. + 3, at 16-byte slot
// (2 t + s) 64 + 16 g + rl.  A wave thus streams one contiguous run as P
// advances (row-major bits gave each lane's row 128 bytes every pair, rows
// 12.5 KB apart at N = 10^5, and HBM served that at about half its rate).
__host__ __device__ __forceinline__ size_t sign_tile_word(long r, long q, long npairs) {
  const long rg = r >> 6, t = (r >> 4) & 3, rl = r & 15;
  const long P = q >> 5, g = (q >> 3) & 3, s = (q >> 2) & 1, e = q & 3;
  return ((((size_t)rg * npairs + P) * 8 + (2 * t + s)) * 64 + (16 * g + rl)) * 4 + e;
}

// Row n of the pass-2 tiles, word q: signs of m = 32 q .. 32 q + 31 (m < K):
// stream elements n K + 32 q + b, i.e. 32 consecutive stream bits from
// position e0 = n K + 32 q (at most two stream words).  One thread per word
// of the rows n < N, in tile order (coalesced stores).
[[maybe_unused]] static __global__ __launch_bounds__(kBlock) void gen_sign_bits_kernel(uint32_t *tiles, long N, long K,
                                                                                      long npairs, uint32_t seed) {
  const long o = (long)blockIdx.x * kBlock + threadIdx.x;
  if (o >= (N + 63) / 64 * npairs * 2048) return;
  const long piece = o >> 2, lane = piece & 63, ts = (piece >> 6) & 7, tile = piece >> 9;
  const long n = 64 * (tile / npairs) + 16 * (ts >> 1) + (lane & 15);
  const long q = 32 * (tile % npairs) + 8 * (lane >> 4) + 4 * (ts & 1) + (o & 3);
  if (n >= N) return;  // the ones row and zero rows are written separately
  uint32_t word = 0;
  if (32 * q < K) {
    const long nb = K - 32 * q < 32 ? K - 32 * q : 32;
    const uint64_t e0 = (uint64_t)n * (uint64_t)K + (uint64_t)(32 * q);
    const uint32_t sh = (uint32_t)(e0 & 31);
    word = sign_stream_word(e0 >> 5, seed) >> sh;
    if (sh != 0 && sh + nb > 32) word |= sign_stream_word((e0 >> 5) + 1, seed) << (32 - sh);
    if (nb < 32) word &= (1u << nb) - 1u;
  }
  tiles[o] = word;
}

// The ones row (row `rows`) of a pass's tiles: word q = ones[q].
[[maybe_unused]] static __global__ __launch_bounds__(kBlock) void sign_tile_ones_kernel(uint32_t *tiles, long rows,
                                                                                       const uint32_t *ones,
                                                                                       long npairs) {
  const long q = (long)blockIdx.x * kBlock + threadIdx.x;
  if (q < 32 * npairs) tiles[sign_tile_word(rows, q, npairs)] = ones[q];
}

// Pass-1 tiles from pass-2 tiles, 32 x 64 bit blocks: lane l holds row
// n0 + l's word q; the ballot of bit b over the wave is column m = 32 q + b
// for those 64 rows, i.e. words n0 / 32 and n0 / 32 + 1 of pass-1 row m.
// Lane b (< 32) writes row m = 32 q + b (rows m >= K get zero words; the
// ones row is written afterwards).
[[maybe_unused]] static __global__ __launch_bounds__(kBlock) void transpose_sign_bits_kernel(const uint32_t *tiles2, long N,
                                                                                            long npairs2, long Kw,
                                                                                            uint32_t *tiles1,
                                                                                            long npairs1) {
  const long wave = ((long)blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const long nblk = (N + 63) / 64;
  if (wave >= nblk * Kw) return;  // whole waves: the ballot needs every lane
  const long q = wave / nblk, n0 = (wave % nblk) * 64;
  const uint32_t w = (n0 + lane < N) ? tiles2[sign_tile_word(n0 + lane, q, npairs2)] : 0u;
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
    const long m = 32 * q + lane, c = n0 / 32;
    tiles1[sign_tile_word(m, c, npairs1)] = lo;
    if (c + 1 < 32 * npairs1) tiles1[sign_tile_word(m, c + 1, npairs1)] = hi;
  }
}

// ---------------------------------------------------------------------
// The int8 passes.  An MFMA v_mfma_i32_16x16x64_i8 multiplies a 16 x 64 int8
// A tile by a 64 x 16 int8 B tile into 16 x 16 int32: lane l holds A row l%16
// and B column l%16 with 16 bytes each for its quarter g = l/16 of the 64 k,
// and D rows 4g..4g+3 of column l%16.  A = the bits as bytes 0/1; B = 16
// columns: the 7 digits of the cos (or G_c) value in 0..6, of sin (G_s) in
// 7..13, zero in 14 and 15.
//
// k is walked in pairs of 512-wide super-blocks (32 bit words = 128 bytes
// per row).  Lane (row, g) loads words 32 P + 8 g .. + 7 of its row as two
// 16-byte vectors from the pass's tiles (sign_tile_word: the wave's 8 KB of
// the pair are consecutive); super-block S = 2 P + s takes words
// 32 P + 8 g + 4 s + u, u = 0..3.  For word u and
// half h (0, 1) the A operand is the four dwords (word >> (4 h + j)) &
// 0x01010101, j = 0..3, so byte i of dword j is bit 8 i + 4 h + j:
// k = 32 (32 P + 8 g + 4 s + u) + 8 i + 4 h + j.  The B operand of (S, u, h)
// is stored in exactly that order (sign_bfrag_offset), 1 KB per MFMA, 8 KB
// per super-block, so the expansion of the bits costs two VALU operations
// per 4 entries and no lane exchange.
// ---------------------------------------------------------------------
typedef int i32x4_t __attribute__((ext_vector_type(4)));
// Register sets of the pass's ring: one pair in use, kSignPF - 1 in flight.
// 2 sets: 124 VGPRs (the occupancy query reports 3 workgroups per CU); 3
// sets (fewer waves per SIMD, about the same bytes in flight per SIMD) ran
// 4% slower at N = K = 10^5.
constexpr int kSignPF = 2;

template <int I, int U, class F> __device__ __forceinline__ void sign_unroll(F &&f) {
  if constexpr (I < U) {
    f(std::integral_constant<int, I>{});
    sign_unroll<I + 1, U>(f);
  }
}
constexpr int kSignSB = 512;                          // k per super-block
constexpr long kSignBias = 0x80808080808080LL;        // 128 * sum_l 256^l, l < 7
constexpr double kSignScale = 0x1p54;                 // fixed point: 2^-54 units
// Largest N and K: a digit sum over k (2 P - T) is at most 3 * 128 k in
// magnitude, 384 * 2^20 < 2^28.6 here, which keeps sign_combine exact.
constexpr long kSignMaxDim = 1L << 20;

// byte of column c (0..15) for index k in the B fragments
__device__ __forceinline__ size_t sign_bfrag_offset(long k, int c) {
  const long q = k >> 5;
  const long S = 2 * (q >> 5) + ((q >> 2) & 1);
  const int g = (int)((q >> 3) & 3), u = (int)(q & 3);
  const int b = (int)(k & 31), i = b >> 3, h = (b >> 2) & 1, j = b & 3;
  return ((((size_t)S * 4 + u) * 2 + h) * 64 + (c + 16 * g)) * 16 + j * 4 + i;
}

// The 7 signed base-256 digits of round(v 2^54), |v| <= 1: e = d + bias is
// non-negative and below 2^56, and digit l = byte l of e minus 128 (as a
// byte: byte ^ 0x80), so that d = sum_l digit_l 256^l.
__device__ __forceinline__ void sign_digits(double v, uint8_t *out) {
  const long long e = __double2ll_rn(v * kSignScale) + kSignBias;
#pragma unroll
  for (int l = 0; l < 7; ++l) out[l] = (uint8_t)(((e >> (8 * l)) & 0xFF) ^ 0x80);
}

// sum_l s_l 256^l 2^-54 rounded once to double: |s_l| < 2^28.6, so the low
// four digits (< 2^52.6) and the high three (< 2^44.6) are exact int64 and
// exact doubles, and one fma rounds their sum.
__device__ __forceinline__ double sign_combine(const long long (&s)[7]) {
  const long long lo = s[0] + s[1] * 256LL + s[2] * 65536LL + s[3] * 16777216LL;
  const long long hi = s[4] + s[5] * 256LL + s[6] * 65536LL;
  return __builtin_fma((double)hi, 0x1p32, (double)lo) * 0x1p-54;
}

// B fragments of pass 1 from the stage vector pair: one thread per k = n
// (zero columns 14, 15 and zero rows for k >= n_used up to the padded end).
template <class R>
__global__ __launch_bounds__(kBlock) void sign_limbs_kernel(const R *__restrict__ vc, const R *__restrict__ vs,
                                                             long n_used, long n_pad, uint8_t *__restrict__ bfrag,
                                                             const State<R> *st, int check_state) {
  if (check_state && (!st->active || st->commit_pending)) return;
  const long k = (long)blockIdx.x * kBlock + threadIdx.x;
  if (k >= n_pad) return;
  uint8_t d[16] = {};
  if (k < n_used) {
    sign_digits((double)vc[k], d);
    sign_digits((double)vs[k], d + 7);
  }
#pragma unroll
  for (int c = 0; c < 16; ++c) bfrag[sign_bfrag_offset(k, c)] = d[c];
}

__device__ __forceinline__ i32x4_t sign_expand(uint32_t w, int h) {
  const uint32_t x = w >> (4 * h), m = 0x01010101u;
  return i32x4_t{(int)(x & m), (int)((x >> 1) & m), (int)((x >> 2) & m), (int)((x >> 3) & m)};
}

// One pass: part[sp][r][c] = sum over the k of split sp of bit(r, k) * B[k][c]
// for the rows r < ldp of the pass's tiles (its bit rows, then the ones row:
// T_c = sum_k B[k][c], then zero rows).  A workgroup = 4 waves on the same
// split of super-block pairs; wave w owns row group 4 (group of blocks) + w,
// NT 16-row tiles.  The B fragments of a pair (16 KB) are loaded once per
// workgroup, 64 bytes per thread, and shared through LDS (double-buffered,
// one barrier per pair); every thread's loads for pair P + kSignPF (its
// share of the B fragments before the MFMAs of pair P, its rows' bit words
// after them) are issued in the order they are used, so kSignPF - 1 pairs
// stay in flight behind the one in use.
template <class R, int NT>
__global__ __launch_bounds__(kBlock) void sign_mfma_kernel(const u32x4 *__restrict__ tiles,
                                                            const u32x4 *__restrict__ bfrag, long npairs,
                                                            long pairs_per, int nsplit, int32_t *__restrict__ part,
                                                            long ldp, const State<R> *st, int check_state) {
  constexpr int NW = kBlock / kWave;
  constexpr int BPT = 2 * 8 * kWave / kBlock;  // 16-byte B pieces per thread and pair
  __shared__ u32x4 bsh[2][2 * 8 * kWave];
  if (check_state && (!st->active || st->commit_pending)) return;
  const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
  const long rgb = blockIdx.x / nsplit;  // the workgroup's 4 row groups
  const int sp = (int)(blockIdx.x % nsplit);
  const long rg = rgb * NW + wave;
  const long p0 = (long)sp * pairs_per, p1 = p0 + pairs_per < npairs ? p0 + pairs_per : npairs;
  const int g = lane >> 4, rl = lane & 15;
  static_assert(NT == 4, "tile order: 64 rows per wave");
  // the wave's 8 KB of pair P start at tw + 512 P (16-byte slots); every
  // lane loads (no divergent branch around the loads, which would make the
  // compiler wait for every load in flight)
  const u32x4 *tw = tiles + (size_t)rg * npairs * 512 + lane;
  i32x4_t acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = i32x4_t{0, 0, 0, 0};
  // pair P: the rows' bit words and this thread's pieces of the pair's B
  // fragments (piece f = tid + kBlock j of the pair's 1024, in the pair's
  // order: super-block s, (u, h), lane).  Nothing here consumes a value.
  auto load_w = [&](long P, u32x4 (&W)[NT][2]) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      W[t][0] = tw[(size_t)P * 512 + (2 * t) * 64];
      W[t][1] = tw[(size_t)P * 512 + (2 * t + 1) * 64];
    }
  };
  auto load_b = [&](long P, u32x4 (&Bst)[BPT]) {
#pragma unroll
    for (int j = 0; j < BPT; ++j) Bst[j] = bfrag[(size_t)P * 2 * 8 * kWave + tid + kBlock * j];
  };
  // the MFMAs of super-block 2 P + s, B fragments from LDS buffer buf
  auto multiply = [&](const u32x4 (&W)[NT][2], int s, int buf) {
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const i32x4_t b = __builtin_bit_cast(i32x4_t, bsh[buf][(s * 8 + u * 2 + h) * kWave + lane]);
#pragma unroll
        for (int t = 0; t < NT; ++t)
          acc[t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(sign_expand(W[t][s][u], h), b, acc[t], 0, 0, 0);
      }
  };
  // Ring of kSignPF register sets, slot d holding pair P + d.  Loads past
  // the split's end re-read its last pair (harmless, and no branch).
  const long plast = p1 - 1;
  auto cl = [](long x, long hi) { return x < hi ? x : hi; };
  u32x4 W[kSignPF][NT][2], Bst[kSignPF][BPT];
  if (p0 < p1)
    sign_unroll<0, kSignPF>([&](auto i) {
      constexpr int d = decltype(i)::value;
      load_b(cl(p0 + d, plast), Bst[d]);
      load_w(cl(p0 + d, plast), W[d]);
    });
  // pair P + d from slot d: its B fragments to LDS and a barrier, the slot's
  // B pieces refilled with pair P + d + kSignPF, the pair's MFMAs, then the
  // slot's bit words refilled (after the MFMAs that read them)
  auto pair = [&](auto i, long P, bool refill) {
    constexpr int d = decltype(i)::value;
    const int buf = (int)((P + d - p0) & 1);
#pragma unroll
    for (int j = 0; j < BPT; ++j) bsh[buf][tid + kBlock * j] = Bst[d][j];
    __syncthreads();
    if (refill) load_b(cl(P + d + kSignPF, plast), Bst[d]);
    __builtin_amdgcn_sched_barrier(0);
    multiply(W[d], 0, buf);
    multiply(W[d], 1, buf);
    __builtin_amdgcn_sched_barrier(0);
    if (refill) load_w(cl(P + d + kSignPF, plast), W[d]);
    __builtin_amdgcn_sched_barrier(0);
  };
  long P = p0;
  for (; P + kSignPF <= p1; P += kSignPF) sign_unroll<0, kSignPF>([&](auto i) { pair(i, P, true); });
  sign_unroll<0, kSignPF - 1>([&](auto i) {
    if (P + decltype(i)::value < p1) pair(i, P, false);
  });
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const long r = rg * 16 * NT + 16 * t + 4 * g + v;
      part[((size_t)sp * ldp + r) * 16 + rl] = acc[t][v];
    }
}

// The 14 digit sums of row r (2 P - T over the splits), exact in int64.
__device__ __forceinline__ void sign_row_sums(const int32_t *__restrict__ part, int nsplit, long ldp, long r,
                                              long trow, long long (&sc)[7], long long (&ss)[7]) {
#pragma unroll
  for (int l = 0; l < 7; ++l) { sc[l] = 0; ss[l] = 0; }
  for (int sp = 0; sp < nsplit; ++sp) {
    const int32_t *pr = part + ((size_t)sp * ldp + r) * 16, *pt = part + ((size_t)sp * ldp + trow) * 16;
#pragma unroll
    for (int l = 0; l < 7; ++l) {
      sc[l] += 2LL * pr[l] - pt[l];
      ss[l] += 2LL * pr[7 + l] - pt[7 + l];
    }
  }
}

// G_m = w_m sum_n X_nm (cos, sin)_n from the pass-1 sums, written as the B
// fragments of pass 2 (one thread per k = m, zero past K up to the padded end).
template <class R>
__global__ __launch_bounds__(kBlock) void sign_gram_kernel(const int32_t *__restrict__ part, int nsplit, long ldp,
                                                            const R *__restrict__ w, long K, long k_pad,
                                                            uint8_t *__restrict__ bfrag, const State<R> *st,
                                                            int check_state) {
  if (check_state && (!st->active || st->commit_pending)) return;
  const long m = (long)blockIdx.x * kBlock + threadIdx.x;
  if (m >= k_pad) return;
  uint8_t d[16] = {};
  if (m < K) {
    long long sc[7], ss[7];
    sign_row_sums(part, nsplit, ldp, m, K, sc, ss);
    const double wm = (double)w[m];
    sign_digits(sign_combine(sc) * wm, d);
    sign_digits(sign_combine(ss) * wm, d + 7);
  }
#pragma unroll
  for (int c = 0; c < 16; ++c) bfrag[sign_bfrag_offset(m, c)] = d[c];
}

// C_n, S_n from the pass-2 sums, then the stage epilogue of row n; stage 7
// writes one error partial per workgroup.
template <class R, int KIND>
__global__ __launch_bounds__(kBlock) void sign_stage_kernel(const int32_t *__restrict__ part, int nsplit, long ldp,
                                                             Vecs<R> v) {
  if constexpr (KIND >= 2) {
    if (v.check_state && (!v.st->active || v.st->commit_pending)) return;
  }
  __shared__ double errsq[kBlock];
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  double sq = 0.0;
  if (i < v.N) {
    double C = 0.0, S = 0.0;
    if (nsplit > 0) {
      long long sc[7], ss[7];
      sign_row_sums(part, nsplit, ldp, i, v.N, sc, ss);
      C = sign_combine(sc);
      S = sign_combine(ss);
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
// bit row (from the pass-2 tiles) one word (32 m) at a time, the rows'
// words are the same in every lane.
template <class R, class S, int ROWS>
__global__ __launch_bounds__(kBlock) void gen_adj_sign_kernel(const uint32_t *tiles, long npairs, const R *w, long K,
                                                               long N, long row0, long nrows, S *A, long lda,
                                                               R inv_scale) {
  const long i0 = row0 + (long)blockIdx.y * ROWS;
  const long j = (long)blockIdx.x * kBlock + threadIdx.x;
  const long jj = j < N ? j : N - 1;
  R acc[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) acc[r] = (R)0;
  const long kw = (K + 31) / 32;
  for (long q = 0; q < kw; ++q) {
    const uint32_t wj = tiles[sign_tile_word(jj, q, npairs)];
    uint32_t wi[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const long i = i0 + r < N ? i0 + r : N - 1;
      wi[r] = tiles[sign_tile_word(i, q, npairs)];
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
