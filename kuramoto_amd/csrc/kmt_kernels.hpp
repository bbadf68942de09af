// HIP device code of the integrator (gfx950 / CDNA4, wave64).
//
// One DP45 attempt is 9 launches on one stream (captured in a hipGraph):
//   stage_kernel<2..7>  A streamed from HBM once per stage for both cos and
//                       sin; the epilogue writes k_s and the next stage's
//                       (cos, sin) of y + h*sum a_{s+1,j} k_j (stage 6 also
//                       keeps the 5th-order solution; stage 7 forms the
//                       embedded error and a per-workgroup sum of squares)
//   control_kernel      error norm, accept/reject, step-size update,
//                       accepted-step log, output times crossed
//   emit_commit_kernel  dense output at the crossed output times + order
//                       parameter partial sums, then commit (ylast <- y,
//                       y <- y5, k1 <- k7) and the next attempt's stage-2
//                       (cos, sin)
//   finalize_kernel     order parameter of each emitted row, ring and
//                       output bookkeeping, end-of-run detection
// Reference counterparts: dp45.cu:34-164 (stages, error, controller, FSAL,
// interpolation) and kuramoto.cu:42-55, :104-126, :164-249 (RHS, order
// parameter, output grid).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kmt {

constexpr int kBlock = 256;  // threads per workgroup (4 waves of 64)
constexpr int kWave = 64;

// ---------------------------------------------------------------------
// Dormand–Prince 4(5) coefficients (dp45.cu:5-21).  Each entry is the same
// double expression as the reference, rounded to the working precision.
// ---------------------------------------------------------------------
namespace tab {
constexpr double A[6][6] = {
    {1.0 / 5, 0, 0, 0, 0, 0},
    {3.0 / 40, 9.0 / 40, 0, 0, 0, 0},
    {44.0 / 45, -56.0 / 15, 32.0 / 9, 0, 0, 0},
    {19372.0 / 6561, -25360.0 / 2187, 64448.0 / 6561, -212.0 / 729, 0, 0},
    {9017.0 / 3168, -355.0 / 33, 46732.0 / 5247, 49.0 / 176, -5103.0 / 18656, 0},
    {35.0 / 384, 0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84}};
constexpr double E[7] = {-71.0 / 57600, 0, 71.0 / 16695, -71.0 / 1920, 17253.0 / 339200, -22.0 / 525, 1.0 / 40};
constexpr double P[7][4] = {
    {1.0, -8048581381.0 / 2820520608, 8663915743.0 / 2820520608, -12715105075.0 / 11282082432},
    {0, 0, 0, 0},
    {0, 131558114200.0 / 32700410799, -68118460800.0 / 10900136933, 87487479700.0 / 32700410799},
    {0, -1754552775.0 / 470086768, 14199869525.0 / 1410260304, -10690763975.0 / 1880347072},
    {0, 127303824393.0 / 49829197408, -318862633887.0 / 49829197408, 701980252875.0 / 199316789632},
    {0, -282668133.0 / 205662961, 2019193451.0 / 616988883, -1453857185.0 / 822651844},
    {0, 40617522.0 / 29380423, -110615467.0 / 29380423, 69997945.0 / 29380423}};
}  // namespace tab

// ---------------------------------------------------------------------
// Device-resident integrator state (one per context).
// ---------------------------------------------------------------------
template <class R> struct State {
  R t, h, t_last, t1;
  R atl, rtl;
  int active;          // run in progress
  int commit_pending;  // accepted step awaiting output emission + commit
  int pend_lo, pend_hi;
  int eval_i, n_eval;
  int ring_count, ring_cap;
  int times_count, times_cap;
  int fixed;
  int last_accept;
  int error;
  int pad_;
  long long accepted, rejected, attempts, rhs_evals;
  double norm;
};

// Storage tags for the adjacency.
struct bf16_t { uint16_t bits; };
struct f16_t { uint16_t bits; };

// VEC: stored values per 16-byte load; Val: the type a stored value unpacks
// to; Acc: the type the coupling sums accumulate in.  Native FP32 storage
// accumulates in double: the sums are then the correctly rounded float of
// the exact sums up to the double rounding (the oracle rounds its double
// sums the same way), so FP32 runs do not depend on the summation order.
// 16-bit storage keeps FP32 accumulation.
template <class S> struct Store;
template <> struct Store<float> { static constexpr int VEC = 4; using Val = float; using Acc = double; };
template <> struct Store<double> { static constexpr int VEC = 2; using Val = double; using Acc = double; };
template <> struct Store<bf16_t> { static constexpr int VEC = 8; using Val = float; using Acc = float; };
template <> struct Store<f16_t> { static constexpr int VEC = 8; using Val = float; using Acc = float; };

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
// Note: reinterpret a loaded 16-byte vector as a whole (u32x4 -> f32x4) or
// copy an element into a scalar first.  hipcc 7.2 mis-compiles
// __builtin_bit_cast(float, v[q]) applied directly to a vector element (it
// reads lane 0 for every q).

__device__ __forceinline__ float bf16_lo(uint32_t w) { return __builtin_bit_cast(float, w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __builtin_bit_cast(float, w & 0xFFFF0000u); }
__device__ __forceinline__ float f16_lo(uint32_t w) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xFFFFu)); }
__device__ __forceinline__ float f16_hi(uint32_t w) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16)); }

// 16 bytes of the adjacency -> VEC values.
template <class S> __device__ __forceinline__ void unpack(const u32x4 v, typename Store<S>::Val *out);
template <> __device__ __forceinline__ void unpack<float>(const u32x4 v, float *o) {
  const f32x4 f = __builtin_bit_cast(f32x4, v);
  o[0] = f.x; o[1] = f.y; o[2] = f.z; o[3] = f.w;
}
template <> __device__ __forceinline__ void unpack<double>(const u32x4 v, double *o) {
  const uint32_t w0 = v.x, w1 = v.y, w2 = v.z, w3 = v.w;
  o[0] = __builtin_bit_cast(double, ((uint64_t)w1 << 32) | w0);
  o[1] = __builtin_bit_cast(double, ((uint64_t)w3 << 32) | w2);
}
template <> __device__ __forceinline__ void unpack<bf16_t>(const u32x4 v, float *o) {
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int q = 0; q < 4; ++q) { o[2 * q] = bf16_lo(w[q]); o[2 * q + 1] = bf16_hi(w[q]); }
}
template <> __device__ __forceinline__ void unpack<f16_t>(const u32x4 v, float *o) {
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int q = 0; q < 4; ++q) { o[2 * q] = f16_lo(w[q]); o[2 * q + 1] = f16_hi(w[q]); }
}

// VEC consecutive reals (16- or 32-byte aligned).
template <class R, int VEC> __device__ __forceinline__ void load_reals(const R *p, R *o) {
  constexpr int W = (int)(VEC * sizeof(R) / 16);
  static_assert(W >= 1, "vector load width");
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const u32x4 v = *reinterpret_cast<const u32x4 *>(reinterpret_cast<const char *>(p) + 16 * w);
    unpack<R>(v, o + (16 / sizeof(R)) * w);
  }
}

// Working-precision math.  FP32 cos / sin are the float roundings of the
// double functions (within 1 ulp of the exact value), i.e. the correctly
// rounded float except within 2^-29 relative of a rounding boundary: FP32
// runs then do not depend on one float libm's last bit (the oracle rounds
// the same way).  With Cauchy frequencies the fast oscillators' phases
// reach 1e4, where one float ulp is 1e-3 rad, and a last-bit difference of
// a cos or sin moves the whole population by ~1e-4 within t = 1.
template <class R> __device__ __forceinline__ R rcos(R x) {
  if constexpr (sizeof(R) == 4) return (float)cos((double)x); else return cos(x);
}
template <class R> __device__ __forceinline__ R rsin(R x) {
  if constexpr (sizeof(R) == 4) return (float)sin((double)x); else return sin(x);
}
template <class R> __device__ __forceinline__ R rabs(R x) {
  if constexpr (sizeof(R) == 4) return fabsf(x); else return fabs(x);
}
template <class R> __device__ __forceinline__ R rfma(R a, R b, R c) {
  if constexpr (sizeof(R) == 4) return __builtin_fmaf(a, b, c); else return __builtin_fma(a, b, c);
}
template <class R> __device__ __forceinline__ R rmax(R a, R b) {
  if constexpr (sizeof(R) == 4) return fmaxf(a, b); else return fmax(a, b);
}

template <class T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Deterministic block sum of one double per thread (fixed tree order).
__device__ __forceinline__ double block_sum(double v, double *sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
#pragma unroll
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (tid < off) sh[tid] += sh[tid + off];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------
// Shared epilogue of every RHS evaluation.
// ---------------------------------------------------------------------
template <class R> struct Vecs {
  const R *cs_in;  // cos at [0, ldv), sin at [ldv, 2 ldv)
  R *cs_out;
  long ldv;        // padded vector length
  long N;
  const R *omega;
  R *y;
  R *k[7];
  R *ynew;         // 5th-order solution of the current attempt
  R *f_out;        // KIND_RHS output
  R *cs_pairs;     // interleaved (C_i, S_i), or nullptr
  R c1;
  State<R> *st;
  double *err_part;  // one per workgroup (stage 7)
  int check_state;   // stage kernels: skip unless an attempt is live
};

constexpr int KIND_RHS = 0;       // f = RHS(theta), optional couplings
constexpr int KIND_COUPLING = 1;  // couplings only
// KIND 2..7 = DP45 stage s

// kuramoto.cu:45: f = w + c1 (sin th C - cos th S), contracted as
// fma(c1, fma(sin, C, -(cos S)), w).
template <class R> __device__ __forceinline__ R rhs_value(R om, R c1, R cs, R sn, R C, R S) {
  return rfma<R>(c1, rfma<R>(sn, C, -(cs * S)), om);
}

// Row epilogue for oscillator i.  Returns the squared scaled error of
// stage 7 (0 otherwise).  Stage combinations follow dp45.cu:34-69 with the
// contraction acc = k1*a0; acc = fma(k_j, a_j, acc); y + h*acc -> fma(acc, h, y).
template <class R, int KIND>
__device__ __forceinline__ double row_epilogue(const Vecs<R> &a, long i, R C, R S, R h) {
  const R cs = a.cs_in[i], sn = a.cs_in[a.ldv + i];
  const R f = rhs_value<R>(a.omega[i], a.c1, cs, sn, C, S);
  if constexpr (KIND == KIND_RHS) {
    a.f_out[i] = f;
    if (a.cs_pairs) { a.cs_pairs[2 * i] = C; a.cs_pairs[2 * i + 1] = S; }
    return 0.0;
  } else if constexpr (KIND == KIND_COUPLING) {
    a.cs_pairs[2 * i] = C;
    a.cs_pairs[2 * i + 1] = S;
    return 0.0;
  } else {
    constexpr int s = KIND;  // stage 2..7, k_s = f
    a.k[s - 1][i] = f;
    if constexpr (s < 7) {
      // coefficients of stage s+1 over k_1..k_s
      R acc = a.k[0][i] * (R)tab::A[s - 1][0];
#pragma unroll
      for (int j = 1; j < s; ++j) {
        const R kj = (j == s - 1) ? f : a.k[j][i];
        acc = rfma<R>(kj, (R)tab::A[s - 1][j], acc);
      }
      const R yt = rfma<R>(acc, h, a.y[i]);
      if constexpr (s + 1 == 7) a.ynew[i] = yt;
      a.cs_out[i] = rcos<R>(yt);
      a.cs_out[a.ldv + i] = rsin<R>(yt);
      return 0.0;
    } else {
      // embedded error, dp45.cu:79-84
      R acc = a.k[0][i] * (R)tab::E[0];
#pragma unroll
      for (int j = 1; j < 7; ++j) {
        const R kj = (j == 6) ? f : a.k[j][i];
        acc = rfma<R>(kj, (R)tab::E[j], acc);
      }
      const R st = a.st->rtl;
      const R num = h * acc;
      const R den = rfma<R>(st, rmax<R>(rabs<R>(a.y[i]), rabs<R>(a.ynew[i])), a.st->atl);
      const R e = num / den;
      return (double)e * (double)e;
    }
  }
}

// ---------------------------------------------------------------------
// Fused gemv-pair stage kernel (dense adjacency).
// Each workgroup owns ROWS consecutive rows; all 256 threads sweep the
// columns with 16-byte loads (two chunks in flight per row), reuse each
// (cos, sin) chunk for all ROWS rows from registers, then reduce across
// the workgroup (wave shuffles + LDS) in a fixed order.
// ---------------------------------------------------------------------
template <class R, class S> struct DenseArgs {
  const S *A;
  long lda;     // row stride (elements), multiple of 64
  R mscale;     // M = mscale * stored
  Vecs<R> v;
};

// Row sums of one group of ROWS rows [row0, row_end): wave shuffles, then
// the 4 wave sums in wave order (fixed tree), scale, row epilogue.  Returns
// (in thread 0) the group's sum of squared scaled errors for stage 7.
template <class R, class Acc, int ROWS, int KIND>
__device__ __forceinline__ double finish_group(Acc (&accC)[ROWS], Acc (&accS)[ROWS], long row0, long row_end,
                                               R mscale, const Vecs<R> &v) {
  __shared__ Acc red[kBlock / kWave][2 * ROWS];
  __shared__ double errsq[ROWS];
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1), wave = tid / kWave;
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    accC[r] = wave_sum(accC[r]);
    accS[r] = wave_sum(accS[r]);
  }
  if (lane == 0) {
#pragma unroll
    for (int r = 0; r < ROWS; ++r) { red[wave][r] = accC[r]; red[wave][ROWS + r] = accS[r]; }
  }
  __syncthreads();
  if (tid < ROWS) {
    const long i = row0 + tid;
    double sq = 0.0;
    if (i < row_end) {
      Acc C = red[0][tid], Sv = red[0][ROWS + tid];
#pragma unroll
      for (int w = 1; w < kBlock / kWave; ++w) { C += red[w][tid]; Sv += red[w][ROWS + tid]; }
      const R h = (KIND >= 2) ? v.st->h : (R)0;
      sq = row_epilogue<R, KIND>(v, i, (R)C * mscale, (R)Sv * mscale, h);
    }
    if constexpr (KIND == 7) errsq[tid] = sq;
  }
  double s = 0.0;
  if constexpr (KIND == 7) {
    __syncthreads();
    if (tid == 0) {
#pragma unroll
      for (int r = 0; r < ROWS; ++r) s += errsq[r];
    }
  }
  return s;
}

// One group per workgroup: rows [row0, min(row0 + ROWS, N)); stage 7 writes
// one error partial per workgroup.
template <class R, class Acc, int ROWS, int KIND>
__device__ __forceinline__ void finish_rows(Acc (&accC)[ROWS], Acc (&accS)[ROWS], long row0, R mscale,
                                            const Vecs<R> &v) {
  const double s = finish_group<R, Acc, ROWS, KIND>(accC, accS, row0, v.N, mscale, v);
  if constexpr (KIND == 7) {
    if (threadIdx.x == 0) v.err_part[blockIdx.x] = s;
  }
}

// BAL = false: workgroup b owns rows [b ROWS, b ROWS + ROWS).
// BAL = true ("balanced"): the grid is one round of resident workgroups and
// workgroup b owns the contiguous rows [b N / G, (b + 1) N / G), walked in
// groups of ROWS, so every workgroup streams the same number of rows (±1)
// and the launch has no tail round.  Rows past the end of a short group
// re-read the group's last row (cache hits) and are discarded.  Every row is
// summed in the same order in every variant.
template <class R, class S, int ROWS, int KIND, bool NT, bool BAL>
__global__ __launch_bounds__(kBlock) void stage_kernel(const DenseArgs<R, S> a) {
  if constexpr (KIND >= 2) {
    if (a.v.check_state && (!a.v.st->active || a.v.st->commit_pending)) return;
  }
  using Acc = typename Store<S>::Acc;
  using Val = typename Store<S>::Val;
  constexpr int VEC = Store<S>::VEC;
  constexpr long STRIDE = (long)kBlock * VEC;
  const int tid = threadIdx.x;
  const long N = a.v.N;
  long rb, re;
  if constexpr (BAL) {
    rb = (long)blockIdx.x * N / gridDim.x;
    re = (long)(blockIdx.x + 1) * N / gridDim.x;
  } else {
    rb = (long)blockIdx.x * ROWS;
    re = rb + ROWS < N ? rb + ROWS : N;
  }
  const R *cosv = a.v.cs_in;
  const R *sinv = a.v.cs_in + a.v.ldv;
  auto load_row = [&](const S *p) -> u32x4 {
    if constexpr (NT) return __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p));
    else return *reinterpret_cast<const u32x4 *>(p);
  };
  double esum = 0.0;
  for (long row0 = rb; row0 < re; row0 += ROWS) {
    const long nr = re - row0 < ROWS ? re - row0 : ROWS;
    // rows past N of the last group (non-balanced grid) are zero padding rows
    // of the allocation; the balanced grid re-reads the group's last row
    const S *rowp[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) rowp[r] = a.A + (row0 + (BAL ? (r < nr ? r : nr - 1) : r)) * a.lda;

    Acc accC[ROWS], accS[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) { accC[r] = 0; accS[r] = 0; }

    long c = (long)tid * VEC;
    for (; c + STRIDE < a.lda; c += 2 * STRIDE) {
      R cv0[VEC], sv0[VEC], cv1[VEC], sv1[VEC];
      u32x4 raw0[ROWS], raw1[ROWS];
      // issue order: the (cos, sin) chunks (L2 hits, needed by every row)
      // first, then all 2 ROWS adjacency loads, then the first use.  The
      // loads complete in issue order, so row r of the first chunk waits
      // only for its own load.  Without the scheduling barriers the
      // scheduler sinks the second chunk's loads next to their uses (one
      // load in flight per wave) to save registers.
      load_reals<R, VEC>(cosv + c, cv0);
      load_reals<R, VEC>(sinv + c, sv0);
      load_reals<R, VEC>(cosv + c + STRIDE, cv1);
      load_reals<R, VEC>(sinv + c + STRIDE, sv1);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int r = 0; r < ROWS; ++r) raw0[r] = load_row(rowp[r] + c);
#pragma unroll
      for (int r = 0; r < ROWS; ++r) raw1[r] = load_row(rowp[r] + c + STRIDE);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        Val m[VEC];
        unpack<S>(raw0[r], m);
#pragma unroll
        for (int q = 0; q < VEC; ++q) {
          accC[r] = rfma<Acc>((Acc)m[q], (Acc)cv0[q], accC[r]);
          accS[r] = rfma<Acc>((Acc)m[q], (Acc)sv0[q], accS[r]);
        }
        unpack<S>(raw1[r], m);
#pragma unroll
        for (int q = 0; q < VEC; ++q) {
          accC[r] = rfma<Acc>((Acc)m[q], (Acc)cv1[q], accC[r]);
          accS[r] = rfma<Acc>((Acc)m[q], (Acc)sv1[q], accS[r]);
        }
      }
    }
    if (c < a.lda) {
      R cv0[VEC], sv0[VEC];
      u32x4 raw0[ROWS];
#pragma unroll
      for (int r = 0; r < ROWS; ++r) raw0[r] = load_row(rowp[r] + c);
      load_reals<R, VEC>(cosv + c, cv0);
      load_reals<R, VEC>(sinv + c, sv0);
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        Val m[VEC];
        unpack<S>(raw0[r], m);
#pragma unroll
        for (int q = 0; q < VEC; ++q) {
          accC[r] = rfma<Acc>((Acc)m[q], (Acc)cv0[q], accC[r]);
          accS[r] = rfma<Acc>((Acc)m[q], (Acc)sv0[q], accS[r]);
        }
      }
    }
    esum += finish_group<R, Acc, ROWS, KIND>(accC, accS, row0, row0 + nr, a.mscale, a.v);
    if constexpr (BAL) __syncthreads();  // the next group reuses the reduction buffers
  }
  if constexpr (KIND == 7) {
    if (tid == 0) a.v.err_part[blockIdx.x] = esum;
  }
}

// The same group pass with the adjacency streamed through a ring of two
// chunk buffers: chunk k + 1 of every row loads while chunk k is reduced,
// so each wave always has one to two chunks of loads in flight (the loop
// above issues a pair of chunks and then waits for all of it) with the
// same two register sets.  Rows and vectors are buffer resources of lda
// elements: the lanes of the last, partial chunk read zeros (no memory
// access) instead of branching, every load in the loop is unconditional,
// and each chunk's loads are issued before the previous chunk's products.
// Every row is summed in the same order as stage_kernel: chunk by chunk, VEC
// values each; a zero chunk adds +0 to sums that are never -0, so the
// results are the same bits.  One group per workgroup.
template <class R, class S, int ROWS, int KIND, bool NT>
__global__ __launch_bounds__(kBlock) void stage_ring_kernel(const DenseArgs<R, S> a) {
  if constexpr (KIND >= 2) {
    if (a.v.check_state && (!a.v.st->active || a.v.st->commit_pending)) return;
  }
  using Acc = typename Store<S>::Acc;
  using Val = typename Store<S>::Val;
  constexpr int VEC = Store<S>::VEC;
  constexpr long STRIDE = (long)kBlock * VEC;
  constexpr int WV = (int)(VEC * sizeof(R) / 16);  // 16-byte loads per vector chunk
  const int tid = threadIdx.x;
  const long row0 = (long)blockIdx.x * ROWS;
  const long nch = (a.lda + STRIDE - 1) / STRIDE;
  __amdgpu_buffer_rsrc_t rrow[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r)
    rrow[r] = __builtin_amdgcn_make_buffer_rsrc(const_cast<S *>(a.A + (row0 + r) * a.lda), (short)0,
                                                (int)(a.lda * (long)sizeof(S)), 0x00020000);
  const __amdgpu_buffer_rsrc_t rcos = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<R *>(a.v.cs_in), (short)0, (int)(a.lda * (long)sizeof(R)), 0x00020000);
  const __amdgpu_buffer_rsrc_t rsin = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<R *>(a.v.cs_in + a.v.ldv), (short)0, (int)(a.lda * (long)sizeof(R)), 0x00020000);
  constexpr int AUX = NT ? 2 : 0;
  struct Chunk {
    u32x4 raw[ROWS];
    u32x4 cv[WV], sv[WV];
  };
  auto load = [&](long k, Chunk &ch) {
    const long c = k * STRIDE + (long)tid * VEC;
    const int vo = (int)(c * (long)sizeof(R));
#pragma unroll
    for (int w = 0; w < WV; ++w) {
      ch.cv[w] = __builtin_amdgcn_raw_buffer_load_b128(rcos, vo + 16 * w, 0, 0);
      ch.sv[w] = __builtin_amdgcn_raw_buffer_load_b128(rsin, vo + 16 * w, 0, 0);
    }
    const int mo = (int)(c * (long)sizeof(S));
#pragma unroll
    for (int r = 0; r < ROWS; ++r) ch.raw[r] = __builtin_amdgcn_raw_buffer_load_b128(rrow[r], mo, 0, AUX);
  };
  Acc accC[ROWS], accS[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) { accC[r] = 0; accS[r] = 0; }
  auto reduce = [&](const Chunk &ch) {
    R cv[VEC], sv[VEC];
#pragma unroll
    for (int w = 0; w < WV; ++w) {
      unpack<R>(ch.cv[w], cv + (16 / sizeof(R)) * w);
      unpack<R>(ch.sv[w], sv + (16 / sizeof(R)) * w);
    }
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      Val m[VEC];
      unpack<S>(ch.raw[r], m);
#pragma unroll
      for (int q = 0; q < VEC; ++q) {
        accC[r] = rfma<Acc>((Acc)m[q], (Acc)cv[q], accC[r]);
        accS[r] = rfma<Acc>((Acc)m[q], (Acc)sv[q], accS[r]);
      }
    }
  };
  Chunk c0, c1;
  load(0, c0);
  // chunk pairs; the loads past the last chunk return zeros without memory
  // traffic, so the loop needs no condition around them
  long k = 0;
#pragma unroll 1
  for (; k + 1 < nch; k += 2) {
    __builtin_amdgcn_sched_barrier(0);
    load(k + 1, c1);
    __builtin_amdgcn_sched_barrier(0);
    reduce(c0);
    __builtin_amdgcn_sched_barrier(0);
    load(k + 2, c0);
    __builtin_amdgcn_sched_barrier(0);
    reduce(c1);
  }
  reduce(c0);  // chunk nch - 1 for an odd count, zeros otherwise
  finish_rows<R, Acc, ROWS, KIND>(accC, accS, row0, a.mscale, a.v);
}

// ---------------------------------------------------------------------
// -A with register-rebuilt adjacency tiles: M_ij = sum_m w_m (X_im X_jm)
// is formed in registers for every (row, column) pair of a stage from the
// transposed factor XT (K x ldv, zero columns past N) — no N x N matrix in
// HBM, (2K + 2) N^2 VALU operations per RHS (K products and K fmas per
// entry, two fmas into the double sums; FP32 adds the entry's conversion to
// double: 2K + 3).  Each entry uses the dense
// generator's formula and the columns are visited in the dense stage
// kernel's order (VEC per 16-byte chunk, chunks strided by kBlock VEC), so
// the result equals the dense FP32/FP64 path on the generated matrix bit
// for bit.
// ---------------------------------------------------------------------
template <class R> struct RebuildArgs {
  const R *XT;  // [K][ldv]
  const R *w;   // [K]
  int K;
  Vecs<R> v;
};

// KT: rank bound of the register arrays; EXACT: K == KT (no per-term
// guards in the unrolled inner loop).
template <class R, int ROWS, int KIND, int KT, bool EXACT>
__global__ __launch_bounds__(kBlock) void rebuild_stage_kernel(const RebuildArgs<R> a) {
  if constexpr (KIND >= 2) {
    if (a.v.check_state && (!a.v.st->active || a.v.st->commit_pending)) return;
  }
  constexpr int VEC = 16 / sizeof(R);
  constexpr long STRIDE = (long)kBlock * VEC;
  const long ldv = a.v.ldv;
  const long row0 = (long)blockIdx.x * ROWS;
  const int K = EXACT ? KT : a.K;
  // The row factors are the same in every lane; a lane shuffle gives them a
  // per-lane register copy (otherwise they sit in scalar registers that
  // spill, and every use in the inner loop reloads them).
  const int lane = threadIdx.x & (kWave - 1);
  R xi[ROWS][KT], wm[KT];
#pragma unroll
  for (int m = 0; m < KT; ++m) {
    wm[m] = __shfl(m < K ? a.w[m] : (R)0, lane, kWave);
#pragma unroll
    for (int r = 0; r < ROWS; ++r) xi[r][m] = __shfl(m < K ? a.XT[m * ldv + row0 + r] : (R)0, lane, kWave);
  }
  const R *cosv = a.v.cs_in;
  const R *sinv = a.v.cs_in + ldv;
  using Acc = typename Store<R>::Acc;  // the dense path's accumulation
  Acc accC[ROWS], accS[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) { accC[r] = 0; accS[r] = 0; }
  for (long c = (long)threadIdx.x * VEC; c < ldv; c += STRIDE) {
    R cv[VEC], sv[VEC], xj[KT][VEC];
    load_reals<R, VEC>(cosv + c, cv);
    load_reals<R, VEC>(sinv + c, sv);
#pragma unroll
    for (int m = 0; m < KT; ++m)
      if (m < K) load_reals<R, VEC>(a.XT + m * ldv + c, xj[m]);
#pragma unroll
    for (int r = 0; r < ROWS; ++r)
#pragma unroll
      for (int q = 0; q < VEC; ++q) {
        R mij = 0;
#pragma unroll
        for (int m = 0; m < KT; ++m)
          if (m < K) mij = rfma<R>(wm[m], xi[r][m] * xj[m][q], mij);
        accC[r] = rfma<Acc>((Acc)mij, (Acc)cv[q], accC[r]);
        accS[r] = rfma<Acc>((Acc)mij, (Acc)sv[q], accS[r]);
      }
  }
  finish_rows<R, Acc, ROWS, KIND>(accC, accS, row0, (R)1, a.v);
}

// ---------------------------------------------------------------------
// Factored (-A) coupling: M = X diag(w) X^T (makecoupling's m1/m2 branch,
// kuramoto.cu:62-108).  G_m = w_m sum_n X_nm (cos_n, sin_n); C_n =
// sum_m X_nm G_m.  Three small launches per RHS; all sums in double.
// ---------------------------------------------------------------------
constexpr int kRankTile = 8;

template <class R>
__global__ __launch_bounds__(kBlock) void lowrank_partial_kernel(const R *X, int K, long N, const R *cs, long ldv,
                                                                 double *part, const State<R> *st, int check_state) {
  if (check_state && (!st->active || st->commit_pending)) return;
  __shared__ double sh[kBlock];
  const int m0 = blockIdx.y * kRankTile;
  double ac[kRankTile], as[kRankTile];
#pragma unroll
  for (int q = 0; q < kRankTile; ++q) { ac[q] = 0; as[q] = 0; }
  for (long n = (long)blockIdx.x * kBlock + threadIdx.x; n < N; n += (long)gridDim.x * kBlock) {
    const double cv = cs[n], sv = cs[ldv + n];
#pragma unroll
    for (int q = 0; q < kRankTile; ++q) {
      if (m0 + q < K) {
        const double x = X[n * K + m0 + q];
        ac[q] = __builtin_fma(x, cv, ac[q]);
        as[q] = __builtin_fma(x, sv, as[q]);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kRankTile; ++q) {
    const double tc = block_sum(ac[q], sh);
    const double ts = block_sum(as[q], sh);
    if (threadIdx.x == 0 && m0 + q < K) {
      part[((size_t)blockIdx.x * K + m0 + q) * 2] = tc;
      part[((size_t)blockIdx.x * K + m0 + q) * 2 + 1] = ts;
    }
  }
}

template <class R>
__global__ __launch_bounds__(kBlock) void lowrank_reduce_kernel(const double *part, int nblk, int K, const R *w,
                                                                double *G, const State<R> *st, int check_state) {
  if (check_state && (!st->active || st->commit_pending)) return;
  __shared__ double sh[kBlock];
  for (int v = 0; v < 2 * K; ++v) {
    double s = 0;
    for (int b = threadIdx.x; b < nblk; b += kBlock) s += part[(size_t)b * 2 * K + v];
    const double tot = block_sum(s, sh);
    if (threadIdx.x == 0) G[v] = (double)w[v / 2] * tot;
  }
}

template <class R, int KIND>
__global__ __launch_bounds__(kBlock) void lowrank_expand_kernel(const R *X, int K, const double *G, Vecs<R> v) {
  if constexpr (KIND >= 2) {
    if (v.check_state && (!v.st->active || v.st->commit_pending)) return;
  }
  __shared__ double errsq[kBlock];
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  double sq = 0.0;
  if (i < v.N) {
    double C = 0, S = 0;
    for (int m = 0; m < K; ++m) {
      const double x = X[i * K + m];
      C = __builtin_fma(x, G[2 * m], C);
      S = __builtin_fma(x, G[2 * m + 1], S);
    }
    const R h = (KIND >= 2) ? v.st->h : (R)0;
    sq = row_epilogue<R, KIND>(v, i, (R)C, (R)S, h);
  }
  if constexpr (KIND == 7) {
    const double tot = block_sum(sq, errsq);
    if (threadIdx.x == 0) v.err_part[blockIdx.x] = tot;
  }
}

// ---------------------------------------------------------------------
// Low-rank coupling, one launch per RHS (K <= kLrMaxRank).
//   G = w .* X^T (cos, sin) is formed in two halves: the launch that writes
//   a stage vector also writes one partial sum of X^T (cos, sin) per
//   workgroup (wave shuffles, then the wave sums in order); the next launch
//   starts with every workgroup adding those partials in workgroup order
//   (the same order in each, so all of them hold the same G), then computes
//   C_i = sum_m X_im G_c[m], S_i likewise and the row epilogue of its stage.
//   lowrank_gram_kernel writes the partials of a given vector (the first
//   stage of an attempt, a stand-alone RHS).  No cross-workgroup
//   synchronisation inside a launch; partial buffers alternate between
//   consecutive launches.
// Rows are strided over a grid of at most kLrMaxBlocks workgroups.
// ---------------------------------------------------------------------
constexpr int kLrMaxRank = 16;
constexpr int kLrMaxBlocks = 512;

template <int KT>
__device__ __forceinline__ void lr_write_partials(double (&acc)[2 * KT], int K, double *part) {
  __shared__ double red[kBlock / kWave][2 * KT];
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int nv = 2 * K;
#pragma unroll
  for (int q = 0; q < 2 * KT; ++q)
    if (q < nv) {
      const double t = wave_sum(acc[q]);
      if (lane == 0) red[wave][q] = t;
    }
  __syncthreads();
  if (threadIdx.x < nv) {
    double t = red[0][threadIdx.x];
#pragma unroll
    for (int q = 1; q < kBlock / kWave; ++q) t += red[q][threadIdx.x];
    part[(size_t)blockIdx.x * nv + threadIdx.x] = t;
  }
}

// G (2K values, interleaved c/s) from nparts <= 2 kBlock partial vectors,
// into g (LDS).  Thread t adds partials t and t + kBlock (all loads issued
// at once), then the sums go through the wave shuffles and the wave order.
static_assert(kLrMaxBlocks <= 2 * kBlock, "lr_reduce covers two partials per thread");
template <class R, int KT>
__device__ __forceinline__ void lr_reduce(const double *part, int nparts, int K, const R *w, double *g) {
  __shared__ double red[kBlock / kWave][2 * KT];
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int nv = 2 * K;
  const int b0 = threadIdx.x, b1 = threadIdx.x + kBlock;
  double t0[2 * KT], t1[2 * KT];
#pragma unroll
  for (int q = 0; q < 2 * KT; ++q) {
    t0[q] = (q < nv && b0 < nparts) ? part[(size_t)b0 * nv + q] : 0.0;
    t1[q] = (q < nv && b1 < nparts) ? part[(size_t)b1 * nv + q] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < 2 * KT; ++q)
    if (q < nv) {
      const double t = wave_sum(t0[q] + t1[q]);
      if (lane == 0) red[wave][q] = t;
    }
  __syncthreads();
  if (threadIdx.x < 2 * KT) {
    const int q = threadIdx.x;
    double t = 0.0;
    if (q < nv) {
      t = red[0][q];
#pragma unroll
      for (int r = 1; r < kBlock / kWave; ++r) t += red[r][q];
      t *= (double)w[q >> 1];
    }
    g[q] = t;
  }
  __syncthreads();
}

template <class R, int KT>
__global__ __launch_bounds__(kBlock) void lowrank_gram_kernel(const R *X, int K, long N, const R *cs, long ldv,
                                                              double *part, const State<R> *st, int check_state) {
  if (check_state && (!st->active || st->commit_pending)) return;
  double acc[2 * KT];
#pragma unroll
  for (int q = 0; q < 2 * KT; ++q) acc[q] = 0.0;
  for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < N; i += (long)gridDim.x * kBlock) {
    const double c = cs[i], s = cs[ldv + i];
#pragma unroll
    for (int m = 0; m < KT; ++m)
      if (m < K) {
        const double x = X[i * K + m];
        acc[2 * m] = __builtin_fma(x, c, acc[2 * m]);
        acc[2 * m + 1] = __builtin_fma(x, s, acc[2 * m + 1]);
      }
  }
  lr_write_partials<KT>(acc, K, part);
}

template <class R, int KIND, int KT, bool NEXT>
__global__ __launch_bounds__(kBlock) void lowrank_stage_kernel(const R *X, int K, const double *part_in, int nparts,
                                                               const R *w, Vecs<R> v, double *part_out) {
  if constexpr (KIND >= 2) {
    if (v.check_state && (!v.st->active || v.st->commit_pending)) return;
  }
  __shared__ double errsq[kBlock];
  __shared__ double gs[2 * KT];
  lr_reduce<R, KT>(part_in, nparts, K, w, gs);
  double g[2 * KT];
#pragma unroll
  for (int q = 0; q < 2 * KT; ++q) g[q] = gs[q];
  double acc[2 * KT];
#pragma unroll
  for (int q = 0; q < 2 * KT; ++q) acc[q] = 0.0;
  const R h = (KIND >= 2) ? v.st->h : (R)0;
  double sq = 0.0;
  for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < v.N; i += (long)gridDim.x * kBlock) {
    double x[KT];
    double C = 0, S = 0;
#pragma unroll
    for (int m = 0; m < KT; ++m)
      if (m < K) {
        x[m] = X[i * K + m];
        C = __builtin_fma(x[m], g[2 * m], C);
        S = __builtin_fma(x[m], g[2 * m + 1], S);
      }
    sq += row_epilogue<R, KIND>(v, i, (R)C, (R)S, h);
    if constexpr (NEXT) {
      const double c = v.cs_out[i], s = v.cs_out[v.ldv + i];
#pragma unroll
      for (int m = 0; m < KT; ++m)
        if (m < K) {
          acc[2 * m] = __builtin_fma(x[m], c, acc[2 * m]);
          acc[2 * m + 1] = __builtin_fma(x[m], s, acc[2 * m + 1]);
        }
    }
  }
  if constexpr (KIND == 7) {
    const double tot = block_sum(sq, errsq);
    if (threadIdx.x == 0) v.err_part[blockIdx.x] = tot;
  }
  if constexpr (NEXT) lr_write_partials<KT>(acc, K, part_out);
}

// ---------------------------------------------------------------------
// Controller (dp45.cu:117-138): one workgroup.
// ---------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(kBlock) void control_kernel(State<R> *st, const double *err_part, int nparts, long N,
                                                         const R *t_eval, R *times) {
  if (!st->active || st->commit_pending) return;
  __shared__ double sh[kBlock];
  double s = 0.0;
  for (int b = threadIdx.x; b < nparts; b += kBlock) s += err_part[b];
  const double sum = block_sum(s, sh);
  if (threadIdx.x != 0) return;
  st->attempts += 1;
  st->rhs_evals += 6;
  const R t = st->t;
  R h = st->h;
  int accept;
  R hnext = h;
  if (st->fixed) {
    accept = 1;
  } else {
    const R nrm = (R)sqrt(sum);                                // {S,D}nrm2
    const R norm = (R)((double)nrm / sqrt((double)N));        // dp45.cu:119
    R factor = (R)(0.9 * pow((double)norm, -0.2));            // dp45.cu:120
    st->norm = (double)norm;
    if (norm < (R)1) {
      accept = 1;
      if (factor > (R)10) factor = (R)10;
      hnext = h * factor;
    } else {
      accept = 0;
      if (factor < (R)1) {
        if ((double)factor < 0.2) factor = (R)0.2;
        hnext = h * factor;
      }
      if (isnan((double)norm)) {  // the reference would loop forever here
        st->error = 1;
        st->active = 0;
      }
    }
  }
  st->last_accept = accept;
  if (!accept) {
    st->rejected += 1;
    st->h = hnext;
    return;
  }
  const R tn = t + h;
  st->t_last = t;
  st->t = tn;
  st->h = hnext;
  st->accepted += 1;
  if (st->times_count < st->times_cap) {
    times[2 * st->times_count] = tn;
    times[2 * st->times_count + 1] = hnext;
    st->times_count += 1;
  }
  // output times crossed by this step (kuramoto.cu:205-212)
  int j = st->eval_i;
  while (j < st->n_eval && tn >= t_eval[j]) ++j;
  st->pend_lo = st->eval_i;
  st->pend_hi = j;
  st->commit_pending = 1;
}

// ---------------------------------------------------------------------
// Dense output + order-parameter partials + commit + next stage-2 vector.
// ---------------------------------------------------------------------
template <class R> struct EmitArgs {
  State<R> *st;
  long N, ldv;
  R *y, *ylast, *ynew;
  R *k[7];
  R *cs_out;          // stage-2 (cos, sin) for the next attempt
  const R *t_eval;
  R *ring;            // ring_cap x ldv rows (or nullptr)
  double *ord_part;   // ring_cap x gridDim.x x 2
};

template <class R> __device__ __forceinline__ void dense_coeffs(const R *kv, R Q[4]) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    R acc = kv[0] * (R)tab::P[0][q];
#pragma unroll
    for (int j = 1; j < 7; ++j) acc = rfma<R>(kv[j], (R)tab::P[j][q], acc);
    Q[q] = acc;
  }
}

// ylast + (hh x) (Q0 + x (Q1 + x (Q2 + x Q3))),  x = (te - t_last) / hh
template <class R> __device__ __forceinline__ R dense_value(const R Q[4], R yl, R hh, R x) {
  R in = rfma<R>(x, Q[3], Q[2]);
  in = rfma<R>(x, in, Q[1]);
  in = rfma<R>(x, in, Q[0]);
  return rfma<R>(hh * x, in, yl);
}

template <class R> __device__ __forceinline__ void stage2_vector(const EmitArgs<R> &a, long i, R y, R k1, R h) {
  const R yt = rfma<R>(k1 * (R)tab::A[0][0], h, y);
  a.cs_out[i] = rcos<R>(yt);
  a.cs_out[a.ldv + i] = rsin<R>(yt);
}

template <class R>
__global__ __launch_bounds__(kBlock) void emit_commit_kernel(const EmitArgs<R> a) {
  const State<R> *st = a.st;
  if (!st->active) return;
  __shared__ double sh[kBlock];
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  const bool valid = i < a.N;
  const R h = st->h;
  if (!st->commit_pending) {
    if (valid) stage2_vector<R>(a, i, a.y[i], a.k[0][i], h);
    return;
  }
  const int pend = st->pend_hi - st->pend_lo;
  const int room = st->ring_cap - st->ring_count;
  const int n_emit = pend < room ? pend : room;
  if (n_emit > 0) {
    const R t = st->t, tl = st->t_last;
    const R hh = t - tl;
    R Q[4] = {0, 0, 0, 0};
    R yl = 0;
    if (valid) {
      R kv[7];
#pragma unroll
      for (int j = 0; j < 7; ++j) kv[j] = a.k[j][i];
      dense_coeffs<R>(kv, Q);
      yl = a.y[i];
    }
    for (int q = 0; q < n_emit; ++q) {
      const int slot = st->ring_count + q;
      const R te = a.t_eval[st->pend_lo + q];
      const R x = (te - tl) / hh;
      double cx = 0.0, sx = 0.0;
      if (valid) {
        const R val = dense_value<R>(Q, yl, hh, x);
        if (a.ring) a.ring[(size_t)slot * a.ldv + i] = val;
        cx = (double)rcos<R>(val);
        sx = (double)rsin<R>(val);
      }
      const double bc = block_sum(cx, sh);
      const double bs = block_sum(sx, sh);
      if (threadIdx.x == 0) {
        a.ord_part[((size_t)slot * gridDim.x + blockIdx.x) * 2] = bc;
        a.ord_part[((size_t)slot * gridDim.x + blockIdx.x) * 2 + 1] = bs;
      }
    }
  }
  if (n_emit == pend && valid) {
    // commit (dp45.cu:123-125, FSAL copy :160) and the next stage-2 vector
    const R yo = a.y[i], yn = a.ynew[i], k7 = a.k[6][i];
    a.ylast[i] = yo;
    a.y[i] = yn;
    a.k[0][i] = k7;
    stage2_vector<R>(a, i, yn, k7, h);
  }
}

// r = sqrt((X/N)^2 + (Y/N)^2) in the working precision (kuramoto.cu:220:
// pow((X/N*X/N + Y/N*Y/N), 0.5) with X, Y the sums of cos and sin).
template <class R> __device__ __forceinline__ R order_value(double X, double Y, long N) {
  const R XR = (R)X, YR = (R)Y, NR = (R)N;
  return (R)pow((double)(XR / NR * XR / NR + YR / NR * YR / NR), 0.5);
}

template <class R>
__global__ __launch_bounds__(kBlock) void finalize_kernel(State<R> *st, const double *ord_part, int nblk, long N,
                                                          R *r_ring) {
  if (!st->active || !st->commit_pending) return;
  __shared__ double sh[kBlock];
  const int pend = st->pend_hi - st->pend_lo;
  const int room = st->ring_cap - st->ring_count;
  const int n_emit = pend < room ? pend : room;
  const int base = st->ring_count;
  for (int q = 0; q < n_emit; ++q) {
    const int slot = base + q;
    double sx = 0, sy = 0;
    for (int b = threadIdx.x; b < nblk; b += kBlock) {
      sx += ord_part[((size_t)slot * nblk + b) * 2];
      sy += ord_part[((size_t)slot * nblk + b) * 2 + 1];
    }
    const double X = block_sum(sx, sh);
    const double Y = block_sum(sy, sh);
    if (threadIdx.x == 0) r_ring[slot] = order_value<R>(X, Y, N);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    st->ring_count = base + n_emit;
    st->pend_lo += n_emit;
    if (st->pend_lo == st->pend_hi) {
      st->commit_pending = 0;
      st->eval_i = st->pend_hi;
      if (!(st->t < st->t1)) st->active = 0;
    }
  }
}

// ---------------------------------------------------------------------
// Small helpers: sin/cos of a phase vector, stand-alone order parameter,
// interpolation over the last accepted step, adjacency generation and
// conversion.
// ---------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(kBlock) void sincos_kernel(const R *y, long N, long ldv, R *cs) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < N) {
    cs[i] = rcos<R>(y[i]);
    cs[ldv + i] = rsin<R>(y[i]);
  }
}

// The same over the padded length: zeros at [N, ldv) of both halves, for a
// (cos, sin) buffer that other calls with other N share.
template <class R>
__global__ __launch_bounds__(kBlock) void sincos_pad_kernel(const R *y, long N, long ldv, R *cs) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < ldv) {
    const bool in = i < N;
    cs[i] = in ? rcos<R>(y[i]) : (R)0;
    cs[ldv + i] = in ? rsin<R>(y[i]) : (R)0;
  }
}

template <class R>
__global__ __launch_bounds__(kBlock) void order_partial_kernel(const R *y, long N, double *part) {
  __shared__ double sh[kBlock];
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  double c = 0, s = 0;
  if (i < N) { c = (double)rcos<R>(y[i]); s = (double)rsin<R>(y[i]); }
  const double bc = block_sum(c, sh);
  const double bs = block_sum(s, sh);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = bc; part[2 * blockIdx.x + 1] = bs; }
}

template <class R>
__global__ __launch_bounds__(kBlock) void order_final_kernel(const double *part, int nblk, long N, double *out) {
  __shared__ double sh[kBlock];
  double sx = 0, sy = 0;
  for (int b = threadIdx.x; b < nblk; b += kBlock) { sx += part[2 * b]; sy += part[2 * b + 1]; }
  const double X = block_sum(sx, sh);
  const double Y = block_sum(sy, sh);
  if (threadIdx.x == 0) out[0] = (double)order_value<R>(X, Y, N);
}

// dp45_eval (dp45.cu:143-146) over the last accepted step.
template <class R>
__global__ __launch_bounds__(kBlock) void interp_kernel(const State<R> *st, long N, const R *ylast,
                                                        const R *k0, const R *k1, const R *k2, const R *k3,
                                                        const R *k4, const R *k5, const R *k6, R t_eval, R *out) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  const R tl = st->t_last;
  const R hh = st->t - tl;
  const R x = (t_eval - tl) / hh;
  const R kv[7] = {k0[i], k1[i], k2[i], k3[i], k4[i], k5[i], k6[i]};
  R Q[4];
  dense_coeffs<R>(kv, Q);
  out[i] = dense_value<R>(Q, ylast[i], hh, x);
}

__device__ __forceinline__ uint16_t to_bf16_rne(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)0x7FC0u;  // NaN
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

template <class R, class S> __device__ __forceinline__ S store_value(R v, R inv_scale);
template <> __device__ __forceinline__ float store_value<float, float>(float v, float) { return v; }
template <> __device__ __forceinline__ double store_value<double, double>(double v, double) { return v; }
template <> __device__ __forceinline__ bf16_t store_value<float, bf16_t>(float v, float s) {
  return bf16_t{to_bf16_rne(v * s)};
}
template <> __device__ __forceinline__ f16_t store_value<float, f16_t>(float v, float s) {
  return f16_t{__builtin_bit_cast(uint16_t, (_Float16)(v * s))};
}

// FP64 working precision with 16-bit storage: the scaled value is rounded
// to FP32 first, then to 16 bits (two roundings; the stored matrix is what
// the integration uses and what kmt_get_adjacency_rows returns).
template <> __device__ __forceinline__ bf16_t store_value<double, bf16_t>(double v, double s) {
  return bf16_t{to_bf16_rne((float)(v * s))};
}
template <> __device__ __forceinline__ f16_t store_value<double, f16_t>(double v, double s) {
  return f16_t{__builtin_bit_cast(uint16_t, (_Float16)(float)(v * s))};
}

template <class R, class S> __device__ __forceinline__ R load_value(S v, R scale);
template <> __device__ __forceinline__ float load_value<float, float>(float v, float) { return v; }
template <> __device__ __forceinline__ double load_value<double, double>(double v, double) { return v; }
template <> __device__ __forceinline__ float load_value<float, bf16_t>(bf16_t v, float s) {
  return __builtin_bit_cast(float, (uint32_t)v.bits << 16) * s;
}
template <> __device__ __forceinline__ float load_value<float, f16_t>(f16_t v, float s) {
  return (float)__builtin_bit_cast(_Float16, v.bits) * s;
}
template <> __device__ __forceinline__ double load_value<double, bf16_t>(bf16_t v, double s) {
  return (double)__builtin_bit_cast(float, (uint32_t)v.bits << 16) * s;
}
template <> __device__ __forceinline__ double load_value<double, f16_t>(f16_t v, double s) {
  return (double)(float)__builtin_bit_cast(_Float16, v.bits) * s;
}

// getadj (kuramoto.cu:145-161) from the stored factor: one workgroup per
// row, M_ij = sum_m fma(w_m, X_im X_jm, acc) (exactly symmetric).
template <class R, class S>
__global__ __launch_bounds__(kBlock) void gen_adj_kernel(const R *X, const R *w, int K, long N, long row0, S *A,
                                                         long lda, R inv_scale) {
  const long i = row0 + blockIdx.x;
  if (i >= N) return;
  const R *Xi = X + i * K;
  S *Ai = A + (size_t)blockIdx.x * lda;
  for (long j = threadIdx.x; j < N; j += kBlock) {
    const R *Xj = X + j * K;
    R acc = 0;
    for (int m = 0; m < K; ++m) acc = rfma<R>(w[m], Xi[m] * Xj[m], acc);
    Ai[j] = store_value<R, S>(acc, inv_scale);
  }
}

// Host-supplied rows (real) -> stored rows.
template <class R, class S>
__global__ __launch_bounds__(kBlock) void convert_rows_kernel(const R *src, long ld_src, long ncols, long nrows, S *dst,
                                                              long lda, R inv_scale) {
  const long r = blockIdx.y;
  if (r >= nrows) return;
  for (long j = (long)blockIdx.x * kBlock + threadIdx.x; j < ncols; j += (long)gridDim.x * kBlock)
    dst[(size_t)r * lda + j] = store_value<R, S>(src[(size_t)r * ld_src + j], inv_scale);
}

// Stored rows -> real rows (dequantized).
template <class R, class S>
__global__ __launch_bounds__(kBlock) void read_rows_kernel(const S *A, long lda, long ncols, long nrows, R scale,
                                                           R *dst) {
  const long r = blockIdx.y;
  if (r >= nrows) return;
  for (long j = (long)blockIdx.x * kBlock + threadIdx.x; j < ncols; j += (long)gridDim.x * kBlock)
    dst[(size_t)r * ncols + j] = load_value<R, S>(A[(size_t)r * lda + j], scale);
}

}  // namespace kmt
