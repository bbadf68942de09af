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

// Shape of the block pass: wave w of the workgroup owns 64 rows of the
// block, RB at a time with the next RB in flight; lane l owns the columns
// q 64 VEC + l VEC + e (q < NLD, e < VEC): four per lane, loaded 16 (8)
// bytes at a time, coalesced per row.  With 4- and 8-byte elements the
// wave's rows are w, w + 4, w + 8, ..., so that the workgroup's loads at any
// moment cover one contiguous stretch of the block; with 16-bit elements
// (512-byte rows) they are the block's w-th quarter, four streams per
// workgroup (each form measured faster for its element size: FP64 1.72 ms
// against 1.76, FP32 0.88 against 0.90, BF16 0.47 against 0.50 per RHS at
// N = 50000).
template <class R, class S> struct SymPass {
  using Acc = typename Store<S>::Acc;
  using Val = typename Store<S>::Val;
  using Raw = typename SymLoad<S>::Raw;
  static constexpr int VEC = SymLoad<S>::VEC;
  static constexpr int NLD = kSymT / (kWave * VEC);
  static constexpr int NC = NLD * VEC;
  static constexpr int RB = 8;
  static constexpr int NW = kBlock / kWave;
  static constexpr int RW = kSymT / NW;
  static constexpr bool QUARTER = sizeof(S) == 2;
  // Row sums of 8- and 2-byte elements are staged in LDS and leave the
  // workgroup in one coalesced store per block (FP64 N = 50000: 1.72 ->
  // 1.67 ms per RHS, 16-bit storage unchanged); for 4-byte elements the
  // per-batch store from the wave measured 0.7% faster.
  static constexpr bool STAGE_ROWS = sizeof(S) != 4;
  static_assert(NC == 4 && RW % (2 * RB) == 0 && 2 * RB == 16, "block shape");

  // the wave's k-th row, and the row distance of consecutive k
  static constexpr int KSTEP = QUARTER ? 1 : NW;
  __device__ static int row_of(int wave, int k) { return (QUARTER ? wave * RW : wave) + KSTEP * k; }

  // The block as a buffer resource: its base in scalar registers, so a
  // lane's address is one 32-bit offset (voff, the lane's row-0 columns in
  // the wave's first row) and each load's row is a scalar constant; the
  // hardware returns 0 past the block's bytes.
  __device__ static __amdgpu_buffer_rsrc_t block_rsrc(const S *blk_base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<S *>(blk_base), (short)0, (int)(kSymT * kSymT * sizeof(S)),
                                            0x00020000);
  }
  __device__ static int lane_offset(int wave, int lane) { return (row_of(wave, 0) * kSymT + lane * VEC) * (int)sizeof(S); }

  // rows kb .. kb + RB - 1 of the wave (lane's columns), non-temporal
  __device__ static void load(Raw (&raw)[RB][NLD], __amdgpu_buffer_rsrc_t rs, int voff, int kb) {
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int q = 0; q < NLD; ++q) {
        const int soff = (KSTEP * (kb + r) * kSymT + q * kWave * VEC) * (int)sizeof(S);
        if constexpr (sizeof(Raw) == 16) raw[r][q] = __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 2);
        else raw[r][q] = __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 2);
      }
  }

  // The wave's rows kb .. kb + RB - 1: row sums against the columns'
  // (cos, sin) reduced over the wave and stored to rows[row] (the block's
  // row partials, or their copy in LDS); column sums against the rows'
  // (cos, sin) (vi) kept per lane.
  __device__ static void process(const Raw (&raw)[RB][NLD], int kb, int wave, int lane, const Acc (&vi)[2][kSymT],
                                 const Acc (&cj)[NC], const Acc (&sj)[NC], Acc (&colC)[NC], Acc (&colS)[NC],
                                 SymPair<Acc> *rows) {
    Acc part[2 * RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      Val mv[NC];
#pragma unroll
      for (int q = 0; q < NLD; ++q) sym_unpack<S>(raw[r][q], mv + q * VEC);
      Acc m[NC];
#pragma unroll
      for (int k = 0; k < NC; ++k) m[k] = (Acc)mv[k];
      const int row = row_of(wave, kb + r);
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
    // every lane stores its slot's sum (the 4 lanes of a slot hold the same
    // value): no divergent branch around the store
    const Acc tot = wave_reduce16<Acc>(part, lane);
    const int slot = ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 + ((lane >> 2) & 1);
    Acc *dst = reinterpret_cast<Acc *>(rows + row_of(wave, kb + (slot % RB)));
    dst[slot / RB] = tot;
  }

  // (cos, sin) at index i, zero past N: the load is clamped into [0, N)
  // rather than branched around, so no divergent branch separates loads
  // (a branch around a load makes the compiler wait for every load in
  // flight where the paths join)
  __device__ static void vec_at(const R *cosv, const R *sinv, long i, long N, Acc &c, Acc &s) {
    const long k = i < N ? i : N - 1;
    const Acc cv = (Acc)cosv[k], sv = (Acc)sinv[k];
    c = i < N ? cv : (Acc)0;
    s = i < N ? sv : (Acc)0;
  }

  // (cos, sin) of the lane's columns of block column J (zero past N)
  __device__ static void col_vectors(const R *cosv, const R *sinv, long J, long N, int lane, Acc (&cj)[NC],
                                     Acc (&sj)[NC]) {
#pragma unroll
    for (int q = 0; q < NLD; ++q)
#pragma unroll
      for (int e = 0; e < VEC; ++e)
        vec_at(cosv, sinv, J * kSymT + q * kWave * VEC + lane * VEC + e, N, cj[q * VEC + e], sj[q * VEC + e]);
  }

  // column sums of the block over the NW waves (fixed order) -> pcol_blk
  __device__ static void finish_columns(Acc (&colbuf)[NW][2][kSymT], const Acc (&colC)[NC], const Acc (&colS)[NC],
                                        int wave, int lane, int tid, SymPair<Acc> *pcol_blk) {
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
    pcol_blk[tid] = out;
  }
};

// One workgroup per stored block.
template <class R, class S>
__global__ __launch_bounds__(kBlock) void sym_tile_kernel(const SymArgs<R, S> a) {
  if (a.v.check_state && (!a.v.st->active || a.v.st->commit_pending)) return;
  using P = SymPass<R, S>;
  using Acc = typename P::Acc;
  using Raw = typename P::Raw;
  constexpr int RB = P::RB, NLD = P::NLD, NC = P::NC, NW = P::NW, RW = P::RW;
  __shared__ Acc vi[2][kSymT];
  __shared__ Acc colbuf[NW][2][kSymT];
  __shared__ SymPair<Acc> rowbuf[P::STAGE_ROWS ? kSymT : 1];
  const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
  const long blk = blockIdx.x;
  // the block's rows depend on the block index only: the first two batches
  // are issued before the block's coordinates and vectors are known
  const __amdgpu_buffer_rsrc_t rs = P::block_rsrc(a.A + (size_t)blk * kSymT * kSymT);
  const int voff = P::lane_offset(wave, lane);
  Raw raw0[RB][NLD], raw1[RB][NLD];
  P::load(raw0, rs, voff, 0);
  P::load(raw1, rs, voff, RB);
  const int2 ij = a.ij[blk];
  const long I = ij.x, J = ij.y, N = a.v.N;
  const R *cosv = a.v.cs_in, *sinv = a.v.cs_in + a.v.ldv;
  {
    Acc c, s;
    P::vec_at(cosv, sinv, I * kSymT + tid, N, c, s);
    vi[0][tid] = c;
    vi[1][tid] = s;
  }
  Acc cj[NC], sj[NC];
  P::col_vectors(cosv, sinv, J, N, lane, cj, sj);
  __syncthreads();
  Acc colC[NC], colS[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) { colC[k] = 0; colS[k] = 0; }
  SymPair<Acc> *prow = a.prow + (size_t)blk * kSymT;
  SymPair<Acc> *rows = P::STAGE_ROWS ? rowbuf : prow;
  // a ring of two row batches: batch k + 2 loads into batch k's registers
  // as soon as batch k is reduced, in the order the batches are used.  The
  // loads in the loop are unconditional (the last pair is peeled off), so
  // each batch stays in its own registers across the back edge: a
  // conditional load there becomes a register copy that waits for the
  // batch just issued, and the ring drains every iteration.
#pragma unroll 1
  for (int kb = 0; kb < RW - 2 * RB; kb += 2 * RB) {
    P::process(raw0, kb, wave, lane, vi, cj, sj, colC, colS, rows);
    P::load(raw0, rs, voff, kb + 2 * RB);
    P::process(raw1, kb + RB, wave, lane, vi, cj, sj, colC, colS, rows);
    P::load(raw1, rs, voff, kb + 3 * RB);
  }
  P::process(raw0, RW - 2 * RB, wave, lane, vi, cj, sj, colC, colS, rows);
  P::process(raw1, RW - RB, wave, lane, vi, cj, sj, colC, colS, rows);
  if constexpr (P::STAGE_ROWS) {
    __syncthreads();
    prow[tid] = rowbuf[tid];
  }
  if (I == J) return;  // the diagonal block is stored whole: its row sums are complete
  P::finish_columns(colbuf, colC, colS, wave, lane, tid, a.pcol + (size_t)blk * kSymT);
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
