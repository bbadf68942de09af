// Symmetric adjacency stored as its upper block triangle (KMT_ADJ_SYM, in
// the working precision or, with KMT_ADJ_BF16 / KMT_ADJ_F16, in 16 bits
// with a power-of-two scale and FP32 accumulation as in the dense path).
//
// Every adjacency generated from a factor is exactly symmetric (M_ij and
// M_ji are the same products, kuramoto.cu:145-161), so half of the dense
// matrix is redundant.  Here the matrix is cut into kSymT x kSymT blocks and
// only the blocks (I, J) with J >= I are stored (row-major inside a block,
// zero past N), in the order of sym_index.  One RHS streams that triangle
// once, i.e. half the bytes of the dense gemv-pair:
//   sym_tile_kernel    one workgroup per stored block: row sums of the block
//                      against (cos, sin) of its columns -> partials of rows
//                      in block row I, and (off the diagonal) column sums
//                      against (cos, sin) of its rows -> partials of rows in
//                      block row J (M_ji = M_ij);
//   sym_reduce_kernel  one thread per row i (block row K): the nb partials
//                      of row i summed in the order of the other block index
//                      (column-sum partials of blocks (L, K), L < K, then
//                      row-sum partials of blocks (K, L), L >= K), then the
//                      stage epilogue of the dense path (row_epilogue).
// The partials cost 4 kSymT values per block (1.6% of the block's bytes at
// kSymT = 256), written once and read once.  Sums run in a fixed order:
// results are deterministic, and differ from the dense path only by the
// order of the additions.
#pragma once
#include <type_traits>

#include "kmt_kernels.hpp"

namespace kmt {

constexpr int kSymT = 256;  // block edge: rows per block = columns per block = kBlock

__host__ __device__ __forceinline__ long sym_index(long I, long J, long nb) {
  return I * nb - I * (I - 1) / 2 + (J - I);
}

template <class T> struct alignas(2 * sizeof(T)) SymPair {
  T c, s;
};

// R: working precision; S: stored element; Acc: accumulation (double for
// native storage, float for 16-bit storage, as in the dense path).
template <class R, class S> struct SymArgs {
  using Acc = typename Store<S>::Acc;
  const S *A;          // stored blocks
  const int2 *ij;      // (I, J) of every stored block
  long nb;             // blocks per row / column of the padded matrix
  SymPair<Acc> *prow;  // [block][kSymT]: row sums (C, S) of the block
  SymPair<Acc> *pcol;  // [block][kSymT]: column sums of off-diagonal blocks
  R mscale;            // M = mscale * stored
  Vecs<R> v;
};

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// One load of a lane: 16 bytes of 4- and 8-byte elements, 8 bytes (four
// values) of 16-bit ones.
template <class S> struct SymLoad {
  static constexpr int BYTES = sizeof(S) >= 4 ? 16 : 8;
  static constexpr int VEC = BYTES / (int)sizeof(S);
  using Raw = typename std::conditional<BYTES == 16, u32x4, u32x2>::type;
};
template <class S> __device__ __forceinline__ void sym_unpack(const u32x4 v, typename Store<S>::Val *o) { unpack<S>(v, o); }
template <class S> __device__ __forceinline__ void sym_unpack(const u32x2 v, float *o);
template <> __device__ __forceinline__ void sym_unpack<bf16_t>(const u32x2 v, float *o) {
  o[0] = bf16_lo(v.x); o[1] = bf16_hi(v.x); o[2] = bf16_lo(v.y); o[3] = bf16_hi(v.y);
}
template <> __device__ __forceinline__ void sym_unpack<f16_t>(const u32x2 v, float *o) {
  o[0] = f16_lo(v.x); o[1] = f16_hi(v.x); o[2] = f16_lo(v.y); o[3] = f16_hi(v.y);
}

// The 16 values v[0..15] of every lane, each summed over the 64 lanes of
// the wave; lane l returns the sum of slot 8 b5 + 4 b4 + 2 b3 + b2 (b = bits
// of l).  Reduce-scatter: 8 + 4 + 2 + 1 exchanges, then 2 to finish.
template <class R> __device__ __forceinline__ R wave_reduce16(R (&v)[16], int lane) {
  const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8, b2 = lane & 4;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const R keep = b5 ? v[k + 8] : v[k], send = b5 ? v[k] : v[k + 8];
    v[k] = keep + __shfl_xor(send, 32);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const R keep = b4 ? v[k + 4] : v[k], send = b4 ? v[k] : v[k + 4];
    v[k] = keep + __shfl_xor(send, 16);
  }
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const R keep = b3 ? v[k + 2] : v[k], send = b3 ? v[k] : v[k + 2];
    v[k] = keep + __shfl_xor(send, 8);
  }
  R s;
  {
    const R keep = b2 ? v[1] : v[0], send = b2 ? v[0] : v[1];
    s = keep + __shfl_xor(send, 4);
  }
  s += __shfl_xor(s, 2);
  s += __shfl_xor(s, 1);
  return s;
}

// The same for 8 values: lane l returns the sum of slot 4 b5 + 2 b4 + b3.
template <class R> __device__ __forceinline__ R wave_reduce8(R (&v)[8], int lane) {
  const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const R keep = b5 ? v[k + 4] : v[k], send = b5 ? v[k] : v[k + 4];
    v[k] = keep + __shfl_xor(send, 32);
  }
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const R keep = b4 ? v[k + 2] : v[k], send = b4 ? v[k] : v[k + 2];
    v[k] = keep + __shfl_xor(send, 16);
  }
  R s;
  {
    const R keep = b3 ? v[1] : v[0], send = b3 ? v[0] : v[1];
    s = keep + __shfl_xor(send, 8);
  }
  s += __shfl_xor(s, 4);
  s += __shfl_xor(s, 2);
  s += __shfl_xor(s, 1);
  return s;
}

// Wave w of the workgroup owns rows w, w + 4, w + 8, ... of the block, four
// at a time with the next four in flight, so that the workgroup's loads at
// any moment cover one contiguous stretch of the block (6% faster than four
// separate quarter-blocks per workgroup); lane l owns the columns
// q 64 VEC + l VEC + e (q < NLD, e < VEC): four per lane, loaded 16 bytes
// at a time, coalesced per row.
template <class R, class S>
__global__ __launch_bounds__(kBlock) void sym_tile_kernel(const SymArgs<R, S> a) {
  if (a.v.check_state && (!a.v.st->active || a.v.st->commit_pending)) return;
  using Acc = typename Store<S>::Acc;
  using Val = typename Store<S>::Val;
  using Raw = typename SymLoad<S>::Raw;
  constexpr int VEC = SymLoad<S>::VEC;
  constexpr int NLD = kSymT / (kWave * VEC);
  constexpr int NC = NLD * VEC;
  constexpr int RB = 4;
  constexpr int NW = kBlock / kWave;
  constexpr int RW = kSymT / NW;
  static_assert(NC == 4 && RW % (2 * RB) == 0, "block shape");
  __shared__ Acc vi[2][kSymT];
  __shared__ Acc colbuf[NW][2][kSymT];
  const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
  const long blk = blockIdx.x;
  // the block's rows depend on the block index only: the first batch is
  // issued before the block's coordinates and vectors are known
  const S *base = a.A + (size_t)blk * kSymT * kSymT + lane * VEC;
  auto row_of = [&](int k) { return wave + NW * k; };  // the wave's k-th row
  auto load = [&](Raw (&raw)[RB][NLD], int kb) {
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int q = 0; q < NLD; ++q)
        raw[r][q] = __builtin_nontemporal_load(
            reinterpret_cast<const Raw *>(base + (size_t)row_of(kb + r) * kSymT + q * kWave * VEC));
  };
  Raw rawA[RB][NLD], rawB[RB][NLD];
  load(rawA, 0);
  const int2 ij = a.ij[blk];
  const long I = ij.x, J = ij.y, N = a.v.N;
  const R *cosv = a.v.cs_in, *sinv = a.v.cs_in + a.v.ldv;
  {
    const long i = I * kSymT + tid;
    vi[0][tid] = i < N ? (Acc)cosv[i] : (Acc)0;
    vi[1][tid] = i < N ? (Acc)sinv[i] : (Acc)0;
  }
  Acc cj[NC], sj[NC];
#pragma unroll
  for (int q = 0; q < NLD; ++q)
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      const long j = J * kSymT + q * kWave * VEC + lane * VEC + e;
      cj[q * VEC + e] = j < N ? (Acc)cosv[j] : (Acc)0;
      sj[q * VEC + e] = j < N ? (Acc)sinv[j] : (Acc)0;
    }
  __syncthreads();
  Acc colC[NC], colS[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) { colC[k] = 0; colS[k] = 0; }
  // the wave's rows kb .. kb + RB - 1: row sums reduced over the wave,
  // column sums kept per lane
  auto process = [&](const Raw (&raw)[RB][NLD], int kb) {
    Acc part[2 * RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      Val mv[NC];
#pragma unroll
      for (int q = 0; q < NLD; ++q) sym_unpack<S>(raw[r][q], mv + q * VEC);
      Acc m[NC];
#pragma unroll
      for (int k = 0; k < NC; ++k) m[k] = (Acc)mv[k];
      const int row = row_of(kb + r);
      const Acc ci = vi[0][row], si = vi[1][row];
      Acc pc = m[0] * cj[0], ps = m[0] * sj[0];
#pragma unroll
      for (int k = 1; k < NC; ++k) {
        pc = rfma<Acc>(m[k], cj[k], pc);
        ps = rfma<Acc>(m[k], sj[k], ps);
      }
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        colC[k] = rfma<Acc>(m[k], ci, colC[k]);
        colS[k] = rfma<Acc>(m[k], si, colS[k]);
      }
      part[r] = pc;
      part[RB + r] = ps;
    }
    const Acc tot = wave_reduce8<Acc>(part, lane);
    if ((lane & 7) == 0) {
      const int slot = ((lane >> 5) & 1) * 4 + ((lane >> 4) & 1) * 2 + ((lane >> 3) & 1);
      Acc *dst = reinterpret_cast<Acc *>(a.prow + (size_t)blk * kSymT + row_of(kb + (slot % RB)));
      dst[slot / RB] = tot;
    }
  };
  // two row batches in flight: the next one loads while this one is reduced
#pragma unroll 1
  for (int kb = 0; kb < RW; kb += 2 * RB) {
    load(rawB, kb + RB);
    process(rawA, kb);
    if (kb + 2 * RB < RW) load(rawA, kb + 2 * RB);
    process(rawB, kb + RB);
  }
  if (I == J) return;  // the diagonal block is stored whole: its row sums are complete
#pragma unroll
  for (int q = 0; q < NLD; ++q)
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      const int c = q * kWave * VEC + lane * VEC + e;
      colbuf[wave][0][c] = colC[q * VEC + e];
      colbuf[wave][1][c] = colS[q * VEC + e];
    }
  __syncthreads();
  SymPair<Acc> out;
  out.c = colbuf[0][0][tid];
  out.s = colbuf[0][1][tid];
#pragma unroll
  for (int w = 1; w < NW; ++w) {
    out.c += colbuf[w][0][tid];
    out.s += colbuf[w][1][tid];
  }
  a.pcol[(size_t)blk * kSymT + tid] = out;
}

// One thread per row; a workgroup covers exactly one block row (kBlock =
// kSymT).  Stage 7 writes one error partial per workgroup.
template <class R, class S, int KIND>
__global__ __launch_bounds__(kBlock) void sym_reduce_kernel(const SymArgs<R, S> a) {
  using Acc = typename Store<S>::Acc;
  if constexpr (KIND >= 2) {
    if (a.v.check_state && (!a.v.st->active || a.v.st->commit_pending)) return;
  }
  static_assert(kSymT == kBlock, "one block row per workgroup");
  __shared__ double sh[kBlock];
  const long K = blockIdx.x, r = threadIdx.x;
  const long i = K * kSymT + r;
  double sq = 0.0;
  if (i < a.v.N) {
    Acc C = 0, Sn = 0;
    const SymPair<Acc> *pc = a.pcol + r, *pr = a.prow + r;
#pragma unroll 4
    for (long L = 0; L < K; ++L) {
      const SymPair<Acc> p = pc[(size_t)sym_index(L, K, a.nb) * kSymT];
      C += p.c;
      Sn += p.s;
    }
    const SymPair<Acc> *row = pr + (size_t)sym_index(K, K, a.nb) * kSymT;
#pragma unroll 4
    for (long L = K; L < a.nb; ++L) {
      const SymPair<Acc> p = row[(size_t)(L - K) * kSymT];
      C += p.c;
      Sn += p.s;
    }
    const R h = (KIND >= 2) ? a.v.st->h : (R)0;
    sq = row_epilogue<R, KIND>(a.v, i, (R)C * a.mscale, (R)Sn * a.mscale, h);
  }
  if constexpr (KIND == 7) {
    const double s = block_sum(sq, sh);
    if (threadIdx.x == 0) a.v.err_part[blockIdx.x] = s;
  }
}

// Stored blocks from the factor with gen_adj_kernel's formula (the same
// value bits as the dense matrix); one workgroup per block, one column per
// thread.
template <class R, class S>
__global__ __launch_bounds__(kBlock) void gen_adj_sym_kernel(const R *X, const R *w, int K, long N, const int2 *ij,
                                                             S *A, R inv_scale) {
  const long blk = blockIdx.x;
  const int2 b = ij[blk];
  const long j = (long)b.y * kSymT + threadIdx.x;
  S *dst = A + (size_t)blk * kSymT * kSymT + threadIdx.x;
  const R *Xj = X + (j < N ? j : 0) * K;
  for (int r = 0; r < kSymT; ++r) {
    const long i = (long)b.x * kSymT + r;
    R acc = 0;
    if (i < N && j < N) {
      const R *Xi = X + i * K;
      for (int m = 0; m < K; ++m) acc = rfma<R>(w[m], Xi[m] * Xj[m], acc);
    }
    dst[(size_t)r * kSymT] = store_value<R, S>(acc, inv_scale);
  }
}

// Rows [row0, row0 + nrows) of the full matrix from the stored triangle.
template <class R, class S>
__global__ __launch_bounds__(kBlock) void read_sym_rows_kernel(const S *A, long nb, long N, long row0, long nrows,
                                                               R scale, R *dst) {
  const long r = blockIdx.y;
  if (r >= nrows) return;
  const long i = row0 + r, I = i / kSymT, ri = i % kSymT;
  for (long j = (long)blockIdx.x * kBlock + threadIdx.x; j < N; j += (long)gridDim.x * kBlock) {
    const long J = j / kSymT, cj = j % kSymT;
    const S v = I <= J ? A[((size_t)sym_index(I, J, nb) * kSymT + ri) * kSymT + cj]
                       : A[((size_t)sym_index(J, I, nb) * kSymT + cj) * kSymT + ri];
    dst[(size_t)r * N + j] = load_value<R, S>(v, scale);
  }
}

}  // namespace kmt
