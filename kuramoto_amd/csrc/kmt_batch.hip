// Batched integration: B independent DP45 runs advanced in lockstep stages
// on one GPU, each with its own step size, accept/reject decision, output
// grid cursor and order-parameter ring (the in-process replacement for the
// reference's SLURM arrays of single runs, job1.sh:13-25).  FP32 state.
//
// Two coupling layouts:
//  * shared   — all instances have the same N and one adjacency A (BF16 or
//               FP16 in HBM).  A stage is the product Z^T = V^T A^T on MFMA
//               (v_mfma_f32_32x32x16_{bf16,f16}), V^T = the 2B rows
//               (cos theta_b, sin theta_b) split into hi + lo 16-bit parts
//               so the vectors keep ~16 bits; split-K over workgroups with
//               fixed-order partial sums.  A is read once per stage for all
//               instances.
//  * grouped  — instance b has its own N_b and its own adjacency A_b; a
//               stage is one grouped gemv-pair launch over (instance,
//               row-block) work items (the plain stage kernel's loop).
// After the coupling launch, one epilogue launch applies the RHS
// (kuramoto.cu:45) and the DP45 stage combination per (instance, oscillator);
// controller / dense output / order parameter follow the single-run kernels
// with an instance index.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/kuramoto_amd.h"
#include "kmt_common.hpp"
#include "kmt_kernels.hpp"
#include "kmt_mlp.hpp"

namespace kmt {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

// Per-instance device state.
struct BInst {
  float t, h, t_last, c1;
  float mscale;         // M = mscale * stored (per instance in grouped mode)
  int active, commit_pending, pend_lo, pend_hi;
  int eval_i, ring_count, error, last_accept;
  int n;                // oscillators of this instance
  int proposed;         // the current h came from the learned proposer
  long long accepted, rejected, attempts, rhs_evals;
  long long proposer_accepts, proposer_rejects;
  double norm;
  float logh[3], logerr[3];  // history for the proposer (most recent first)
  float r_now, r_prev;       // order parameter at the last two emitted outputs
  float tr_now, tr_prev;     // and their output times
  float h_std;               // the standard controller's proposal behind the current h
  int hist_n, r_n;           // accepted steps (up to 3) and outputs (up to 2) in the history
  int vrow;             // shared mode: row of cos in V^T and in the product (sin: vrow + half); -1 = none
};

// ---------------------------------------------------------------------
// sin/cos of each instance's phases; optionally the hi/lo 16-bit rows of V^T
// ---------------------------------------------------------------------
template <class T> __device__ __forceinline__ uint16_t to16(float v);
template <> __device__ __forceinline__ uint16_t to16<bf16_t>(float v) { return to_bf16_rne(v); }
template <> __device__ __forceinline__ uint16_t to16<f16_t>(float v) { return __builtin_bit_cast(uint16_t, (_Float16)v); }
template <class T> __device__ __forceinline__ float from16(uint16_t u);
template <> __device__ __forceinline__ float from16<bf16_t>(uint16_t u) { return __builtin_bit_cast(float, (uint32_t)u << 16); }
template <> __device__ __forceinline__ float from16<f16_t>(uint16_t u) { return (float)__builtin_bit_cast(_Float16, u); }

struct VtOut {
  uint16_t *hi, *lo;  // [Mpad][ldk], row vrow = cos, row half + vrow = sin of one instance
  long ldk;
  int half;
};

template <class T> __device__ __forceinline__ void put_vt(const VtOut &v, int row, long n, float c, float s) {
  if (row < 0) return;
  const uint16_t ch = to16<T>(c), sh = to16<T>(s);
  v.hi[(size_t)row * v.ldk + n] = ch;
  v.hi[(size_t)(v.half + row) * v.ldk + n] = sh;
  v.lo[(size_t)row * v.ldk + n] = to16<T>(c - from16<T>(ch));
  v.lo[(size_t)(v.half + row) * v.ldk + n] = to16<T>(s - from16<T>(sh));
}

template <class T>
__global__ __launch_bounds__(kBlock) void bsincos_kernel(const float *y, const BInst *inst, long ldn, float *cs,
                                                         VtOut vt, int write_vt) {
  const int b = blockIdx.y;
  const long n = (long)blockIdx.x * kBlock + threadIdx.x;
  if (n >= inst[b].n) return;
  const float v = y[(size_t)b * ldn + n];
  const float c = rcos<float>(v), s = rsin<float>(v);
  cs[((size_t)b * 2) * ldn + n] = c;
  cs[((size_t)b * 2 + 1) * ldn + n] = s;
  if (write_vt) put_vt<T>(vt, inst[b].vrow, n, c, s);
}

// Rewrite the V^T rows of every active instance from its float (cos, sin)
// stage vector (after the rows were reassigned): the same conversion as
// put_vt at the vector's producer, so the product sees identical bits.
template <class T>
__global__ __launch_bounds__(kBlock) void bvt_rows_kernel(const float *cs, const BInst *inst, long ldn, VtOut vt) {
  const int b = blockIdx.y;
  const long n = (long)blockIdx.x * kBlock + threadIdx.x;
  if (!inst[b].active || n >= inst[b].n) return;
  put_vt<T>(vt, inst[b].vrow, n, cs[((size_t)b * 2) * ldn + n], cs[((size_t)b * 2 + 1) * ldn + n]);
}

// ---------------------------------------------------------------------
// Shared adjacency: Z^T (Mpad x N) = V^T (Mpad x K) . A^T (K x N) on MFMA.
//
// A is stored in MFMA fragment order: 32-row x 64-column tiles, each 4 KB
// (bf16/f16) and laid out as 4 k-steps x 64 lanes x 8 values, so that one
// wave's fragment load for a k-step is 1 KB of consecutive bytes and a wave
// streams one contiguous run of HBM along its rows (tiles of one 32-row
// band are consecutive in k).  Element (i, j) sits at tiled_offset(i, j).
//
// Workgroup = WAVES waves; wave w owns the NT 32-row bands starting at band
// (rb WAVES + w) NT and all MT 32-row tiles of V^T.  K is split into ksplit
// slabs; blockIdx.x = rb ksplit + slab, so with ksplit a multiple or
// divisor of 8 the workgroups of one XCD (dispatched round-robin, block % 8)
// work on the same slab(s) and share its V^T rows in their L2.  V^T chunks
// (MT*32 x 64, hi and lo) are double-buffered in LDS: the next chunk's A
// fragments and V^T pieces are loaded before the current chunk's MFMAs.
// Operand layouts (verified on gfx950): A-operand lane l holds row l%32,
// k = 8(l/32)..+7; B-operand lane l holds column l%32, the same k; D lane
// l, item v holds row 8(v/4) + 4(l/32) + v%4, column l%32.
// ---------------------------------------------------------------------
constexpr int kKC = 64;        // K per chunk (one tile column)
constexpr int kLdsRow = 72;    // padded LDS row (64 + 8 halves)
constexpr long kTile = 32L * kKC;  // values per tile

__host__ __device__ __forceinline__ size_t tiled_offset(long i, long j, long nkc) {
  const long lane = (i & 31) + 32 * ((j & 15) >> 3);
  return ((((size_t)(i >> 5) * nkc + (j >> 6)) * 4 + ((j & 63) >> 4)) * 64 + lane) * 8 + (j & 7);
}

template <class T> __device__ __forceinline__ f32x16_t mfma32(const u32x4 a, const u32x4 b, f32x16_t c);
template <> __device__ __forceinline__ f32x16_t mfma32<bf16_t>(const u32x4 a, const u32x4 b, f32x16_t c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b), c,
                                                 0, 0, 0);
}
template <> __device__ __forceinline__ f32x16_t mfma32<f16_t>(const u32x4 a, const u32x4 b, f32x16_t c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8_t, a), __builtin_bit_cast(f16x8_t, b), c,
                                                0, 0, 0);
}

template <int S> using Stage = std::integral_constant<int, S>;

// PF = chunks of A in flight ahead of the one being multiplied (register
// sets: PF + 1, used round-robin; the loop is unrolled PF + 1 times so that
// every set is addressed statically).
template <int I, int U, class F> __device__ __forceinline__ void static_for(F &&f) {
  if constexpr (I < U) {
    f(Stage<I>{});
    static_for<I + 1, U>(f);
  }
}

template <class T, int MT, int NT, int WAVES, bool LO, int PF, int VD>
__global__ __launch_bounds__(WAVES * kWave) void bgemm_kernel(const T *At, long nkc, const uint16_t *vthi,
                                                              const uint16_t *vtlo, long ldk, int ksplit,
                                                              long slab_chunks, float *P, long ldn, int mpad) {
  static_assert(PF >= 1 && PF <= 3 && (VD == 1 || VD == 2), "prefetch depth");
  constexpr int NTHR = WAVES * kWave;
  constexpr int NV = LO ? 2 : 1;
  constexpr int VPIECES = MT * 32 * 8;  // 16-B pieces of one V^T chunk (hi or lo)
  constexpr int VPT = (VPIECES + NTHR - 1) / NTHR;
  __shared__ uint16_t sh[2][NV][MT * 32 * kLdsRow];
  // the wave index in a scalar register: the band's buffer resource below
  // must be wave-uniform
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x % kWave;
  const int slab = (int)(blockIdx.x % ksplit);
  const long rb = blockIdx.x / ksplit;
  const long band0 = (rb * WAVES + wave) * NT;
  const long c0 = slab * slab_chunks;
  const long c1 = std::min(c0 + slab_chunks, nkc);
  const uint16_t *vsrc[NV];
  vsrc[0] = vthi;
  if constexpr (LO) vsrc[NV - 1] = vtlo;

  f32x16_t acc[MT][NT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[mt][nt][v] = 0.f;

  u32x4 a[PF + 1][NT][4];
  u32x4 vr[VD][NV][VPT];
  // Every load and LDS store in the chunk loop is unconditional: a load
  // behind a branch makes the compiler wait, where the paths join, for every
  // load in flight, and the A prefetch would drain once per loop period.
  // Band nt of A over this slab's chunks is a buffer resource of exactly
  // those chunks: prefetches past the slab's last chunk return zeros
  // without memory traffic (the chunk offset is in the lane offset, which
  // the hardware range-checks).
  __amdgpu_buffer_rsrc_t ars[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt)
    ars[nt] = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<T *>(At + ((size_t)(band0 + nt) * nkc + c0) * kTile), (short)0,
        (int)(c1 > c0 ? (c1 - c0) * kTile * (long)sizeof(T) : 0), 0x00020000);
  auto load_a = [&](long c, u32x4 (&dst)[NT][4]) {
    const int base = (int)((c - c0) * kTile * (long)sizeof(T)) + lane * 16;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int ks = 0; ks < 4; ++ks)
        dst[nt][ks] = __builtin_amdgcn_raw_buffer_load_b128(ars[nt], base + ks * 1024, 0, 2);
  };
  // V^T pieces: past the last piece a thread repeats the last one (the same
  // bytes to the same LDS place), past the slab the last chunk
  auto vpiece = [&](int u) {
    const int p = threadIdx.x + u * NTHR;
    return VPIECES % NTHR == 0 || p < VPIECES ? p : VPIECES - 1;
  };
  auto load_v = [&](long c, u32x4 (&vr)[NV][VPT]) {
    const long cc = c < c1 ? c : c1 - 1;
#pragma unroll
    for (int q = 0; q < NV; ++q)
#pragma unroll
      for (int u = 0; u < VPT; ++u) {
        const int p = vpiece(u);
        vr[q][u] = *reinterpret_cast<const u32x4 *>(vsrc[q] + (size_t)(p >> 3) * ldk + cc * kKC + (p & 7) * 8);
      }
  };
  auto store_v = [&](const u32x4 (&vr)[NV][VPT], int buf) {
#pragma unroll
    for (int q = 0; q < NV; ++q)
#pragma unroll
      for (int u = 0; u < VPT; ++u) {
        const int p = vpiece(u);
        *reinterpret_cast<u32x4 *>(&sh[buf][q][(p >> 3) * kLdsRow + (p & 7) * 8]) = vr[q][u];
      }
  };
  // The MFMAs of one chunk: 4 k-steps x (hi, lo) x MT tiles of V^T from
  // LDS.  With one band per wave (NT = 1) the V^T fragments of k-step ks + 1
  // are read from LDS while the MFMAs of k-step ks run (two fragment sets),
  // so the LDS latency is not exposed once per MFMA pair.
  auto multiply = [&](const u32x4 (&af)[NT][4], int cur) {
    if constexpr (NT == 1) {
      u32x4 vb[2][NV][MT];
      auto frag = [&](int ks, u32x4 (&dst)[NV][MT]) {
        const int off = (lane % 32) * kLdsRow + 16 * ks + 8 * (lane / 32);
#pragma unroll
        for (int q = 0; q < NV; ++q)
#pragma unroll
          for (int mt = 0; mt < MT; ++mt)
            dst[q][mt] = *reinterpret_cast<const u32x4 *>(&sh[cur][q][mt * 32 * kLdsRow + off]);
      };
      frag(0, vb[0]);
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        if (ks + 1 < 4) frag(ks + 1, vb[(ks + 1) & 1]);
#pragma unroll
        for (int q = 0; q < NV; ++q)
#pragma unroll
          for (int mt = 0; mt < MT; ++mt) acc[mt][0] = mfma32<T>(vb[ks & 1][q][mt], af[0][ks], acc[mt][0]);
      }
      // scheduling: the first k-step's reads, then one read of the next
      // k-step per MFMA of the current one (mask 0x100 = LDS read, 0x008 =
      // MFMA), then the last k-step's MFMAs
      constexpr int F = NV * MT;
      __builtin_amdgcn_sched_group_barrier(0x100, F, 0);
#pragma unroll
      for (int i = 0; i < 3 * F; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      }
      __builtin_amdgcn_sched_group_barrier(0x008, F, 0);
    } else {
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const int off = (lane % 32) * kLdsRow + 16 * ks + 8 * (lane / 32);
#pragma unroll
        for (int q = 0; q < NV; ++q)
#pragma unroll
          for (int mt = 0; mt < MT; ++mt) {
            const u32x4 vv = *reinterpret_cast<const u32x4 *>(&sh[cur][q][mt * 32 * kLdsRow + off]);
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = mfma32<T>(vv, af[nt][ks], acc[mt][nt]);
          }
      }
    }
  };
  // One chunk c held in A register set S (= (c - c0) mod (PF + 1)).  V^T is
  // loaded VD chunks ahead into register set (c - c0) mod VD and written to
  // LDS one chunk ahead: the loads of chunk c + VD are issued first, then A
  // for chunk c + PF, then the MFMAs of chunk c, then the LDS store of chunk
  // c + 1 (whose wait leaves every later load in flight).  The loop is
  // unrolled over a common period of both register rotations so that every
  // set is addressed statically.
  constexpr int U = (PF + 1) % VD == 0 ? PF + 1 : VD * (PF + 1);
  auto step = [&](auto stage, long c) {
    constexpr int I = decltype(stage)::value;
    constexpr int S = I % (PF + 1);
    constexpr int NX = (S + PF) % (PF + 1);
    constexpr int SV = I % VD;
    const int cur = (int)((c - c0) & 1);
    // the order is fixed: a load issued after the LDS store would make the
    // store's wait cover the A chunk that is still a step ahead
    __builtin_amdgcn_sched_barrier(0);
    load_v(c + VD, vr[SV]);
    load_a(c + PF, a[NX]);
    __builtin_amdgcn_sched_barrier(0);
    multiply(a[S], cur);
    __builtin_amdgcn_sched_barrier(0);
    store_v(vr[(SV + 1) % VD], cur ^ 1);  // (after the last chunk: into the buffer no one reads again)
    __syncthreads();
  };

  if (c0 < c1) {
    load_a(c0, a[0]);
    if constexpr (PF >= 2)
      if (c0 + 1 < c1) load_a(c0 + 1, a[1]);
    if constexpr (PF >= 3)
      if (c0 + 2 < c1) load_a(c0 + 2, a[2]);
    load_v(c0, vr[0]);
    store_v(vr[0], 0);
    if constexpr (VD == 2)
      if (c0 + 1 < c1) load_v(c0 + 1, vr[1]);
  }
  // Wait for every prologue load here: the loop below then only waits for
  // the prefetch it issues itself (without this, the wait inserted for the
  // prologue's fragments inside the loop also drains the later chunks' loads).
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  long c = c0;
  for (; c + U <= c1; c += U) static_for<0, U>([&](auto i) { step(i, c + decltype(i)::value); });
  static_for<0, U - 1>([&](auto i) {
    if (c + decltype(i)::value < c1) step(i, c + decltype(i)::value);
  });

  float *Pk = P + (size_t)slab * mpad * ldn;
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int m = mt * 32 + 8 * (v / 4) + 4 * (lane / 32) + (v % 4);
        Pk[(size_t)m * ldn + (band0 + nt) * 32 + lane % 32] = acc[mt][nt][v];
      }
}

// Shared adjacency from a factor, written in the tiled order (one
// workgroup per 32 x 64 tile; the same arithmetic as gen_adj_kernel).
template <class S>
__global__ __launch_bounds__(kBlock) void gen_adj_tiled_kernel(const float *X, const float *w, int K, long N, long nkc,
                                                               long tile0, S *At, float inv_scale) {
  const long tile = tile0 + blockIdx.x;
  const long band = tile / nkc, kc = tile % nkc;
  S *dst = At + (size_t)tile * kTile;
  for (int o = threadIdx.x; o < kTile; o += kBlock) {
    const int lane = (o >> 3) & 63, ks = o >> 9;
    const long i = band * 32 + (lane & 31);
    const long j = kc * kKC + ks * 16 + 8 * (lane >> 5) + (o & 7);
    float acc = 0.f;
    if (i < N && j < N) {
      const float *Xi = X + i * K, *Xj = X + j * K;
      for (int m = 0; m < K; ++m) acc = rfma<float>(w[m], Xi[m] * Xj[m], acc);
    }
    dst[o] = store_value<float, S>(acc, inv_scale);
  }
}

// Tiled shared adjacency -> real rows (dequantized).
template <class S>
__global__ __launch_bounds__(kBlock) void read_tiled_rows_kernel(const S *At, long nkc, long ncols, long row0,
                                                                 long nrows, float scale, float *dst) {
  const long r = blockIdx.y;
  if (r >= nrows) return;
  for (long j = (long)blockIdx.x * kBlock + threadIdx.x; j < ncols; j += (long)gridDim.x * kBlock)
    dst[(size_t)r * ncols + j] = load_value<float, S>(At[tiled_offset(row0 + r, j, nkc)], scale);
}

// ---------------------------------------------------------------------
// Grouped adjacency: one gemv-pair per instance, work item = (b, row0).
// The loop matches the single-run stage kernel (same summation order).
// ---------------------------------------------------------------------
struct GroupItem {
  int b;
  int row0;
};

template <class S, int ROWS>
__global__ __launch_bounds__(kBlock) void bgemv_kernel(const S *Abase, const size_t *a_off, const long *a_ld,
                                                       const GroupItem *items, const BInst *inst, const float *cs,
                                                       long ldn, float *P, int B, int check_state) {
  using Acc = typename Store<S>::Acc;
  using Val = typename Store<S>::Val;
  constexpr int VEC = Store<S>::VEC;
  constexpr long STRIDE = (long)kBlock * VEC;
  const GroupItem it = items[blockIdx.x];
  const BInst *in = inst + it.b;
  if (check_state && (!in->active || in->commit_pending)) return;
  const long lda = a_ld[it.b];
  const S *Ab = Abase + a_off[it.b] + (size_t)it.row0 * lda;
  const float *cosv = cs + (size_t)it.b * 2 * ldn;
  const float *sinv = cosv + ldn;
  const int tid = threadIdx.x;
  Acc accC[ROWS], accS[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) { accC[r] = 0; accS[r] = 0; }
  long c = (long)tid * VEC;
  for (; c + STRIDE < lda; c += 2 * STRIDE) {
    float cv0[VEC], sv0[VEC], cv1[VEC], sv1[VEC];
    u32x4 raw0[ROWS], raw1[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) raw0[r] = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(Ab + r * lda + c));
#pragma unroll
    for (int r = 0; r < ROWS; ++r)
      raw1[r] = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(Ab + r * lda + c + STRIDE));
    load_reals<float, VEC>(cosv + c, cv0);
    load_reals<float, VEC>(sinv + c, sv0);
    load_reals<float, VEC>(cosv + c + STRIDE, cv1);
    load_reals<float, VEC>(sinv + c + STRIDE, sv1);
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
  if (c < lda) {
    float cv0[VEC], sv0[VEC];
    u32x4 raw0[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) raw0[r] = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(Ab + r * lda + c));
    load_reals<float, VEC>(cosv + c, cv0);
    load_reals<float, VEC>(sinv + c, sv0);
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
  __shared__ Acc red[kBlock / kWave][2 * ROWS];
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
    const long i = it.row0 + tid;
    if (i < in->n) {
      Acc C = red[0][tid], Sv = red[0][ROWS + tid];
#pragma unroll
      for (int w = 1; w < kBlock / kWave; ++w) { C += red[w][tid]; Sv += red[w][ROWS + tid]; }
      P[(size_t)it.b * ldn + i] = (float)C;
      P[(size_t)(B + it.b) * ldn + i] = (float)Sv;
    }
  }
}

// ---------------------------------------------------------------------
// Epilogue: RHS + DP45 stage combination per (instance, oscillator).
// ---------------------------------------------------------------------
struct BVecs {
  const float *cs_in;
  float *cs_out;
  long ldn;
  int B;
  const float *omega;
  float *y, *ynew;
  float *kbuf;        // [7][B][ldn]
  float *f_out;       // RHS kind output, [B][ldn]
  const float *P;     // [ksplit][mrows][ldn]
  int ksplit, mrows;
  int phalf;          // row of an instance's S = its row of C + phalf
  BInst *inst;
  double *err_part;   // [B][nblk]
  int nblk;
  float atl, rtl;
  VtOut vt;
  int write_vt;
};

template <class T, int KIND>
__global__ __launch_bounds__(kBlock) void bepilogue_kernel(const BVecs a) {
  const int b = blockIdx.y;
  BInst *in = a.inst + b;
  if (KIND >= 2 && (!in->active || in->commit_pending)) return;
  __shared__ double sh[kBlock];
  const long n = (long)blockIdx.x * kBlock + threadIdx.x;
  const size_t vb = (size_t)b * a.ldn;
  const size_t kst = (size_t)a.B * a.ldn;  // stride between k_j slabs
  double sq = 0.0;
  if (n < in->n) {
    float C = 0.f, S = 0.f;
    for (int ks = 0; ks < a.ksplit; ++ks) {
      C += a.P[((size_t)ks * a.mrows + in->vrow) * a.ldn + n];
      S += a.P[((size_t)ks * a.mrows + a.phalf + in->vrow) * a.ldn + n];
    }
    C *= in->mscale;
    S *= in->mscale;
    const float cs = a.cs_in[2 * vb + n], sn = a.cs_in[2 * vb + a.ldn + n];
    const float f = rhs_value<float>(a.omega[vb + n], in->c1, cs, sn, C, S);
    const float h = in->h;
    if constexpr (KIND == KIND_RHS) {
      a.f_out[vb + n] = f;
    } else if constexpr (KIND < 7) {
      constexpr int s = KIND;
      a.kbuf[(s - 1) * kst + vb + n] = f;
      float acc = a.kbuf[vb + n] * (float)tab::A[s - 1][0];
#pragma unroll
      for (int j = 1; j < s; ++j) {
        const float kj = (j == s - 1) ? f : a.kbuf[j * kst + vb + n];
        acc = rfma<float>(kj, (float)tab::A[s - 1][j], acc);
      }
      const float yt = rfma<float>(acc, h, a.y[vb + n]);
      if constexpr (s + 1 == 7) a.ynew[vb + n] = yt;
      const float c = rcos<float>(yt), si = rsin<float>(yt);
      a.cs_out[2 * vb + n] = c;
      a.cs_out[2 * vb + a.ldn + n] = si;
      if (a.write_vt) put_vt<T>(a.vt, in->vrow, n, c, si);
    } else {
      a.kbuf[6 * kst + vb + n] = f;
      float acc = a.kbuf[vb + n] * (float)tab::E[0];
#pragma unroll
      for (int j = 1; j < 7; ++j) {
        const float kj = (j == 6) ? f : a.kbuf[j * kst + vb + n];
        acc = rfma<float>(kj, (float)tab::E[j], acc);
      }
      const float num = h * acc;
      const float den = rfma<float>(a.rtl, fmaxf(fabsf(a.y[vb + n]), fabsf(a.ynew[vb + n])), a.atl);
      const float e = num / den;
      sq = (double)e * (double)e;
    }
  }
  if constexpr (KIND == 7) {
    const double tot = block_sum(sq, sh);
    if (threadIdx.x == 0) a.err_part[(size_t)b * a.nblk + blockIdx.x] = tot;
  }
}

// ---------------------------------------------------------------------
// Learned step-size proposer: the network of kmt_mlp.hpp, whose output is a
// correction log(h_next / h_standard) of the standard controller's
// proposal, clamped to [-0.7, 0.7].  Evaluated by wave 0 of the
// controller's block (one lane per hidden unit), with the same function as
// the training kernels' forward pass.
// ---------------------------------------------------------------------
constexpr float kProposerClamp = 0.7f;

// Inputs: the last 3 log h relative to the standard proposal, the last 3 log
// error norms, r at the last output and its slope over the last two outputs.
__device__ void proposer_features(const BInst *in, float h_std, float (&x)[kIn]) {
  const float lh = logf(h_std);
  for (int q = 0; q < 3; ++q) x[q] = in->hist_n > q ? in->logh[q] - lh : 0.f;
  for (int q = 0; q < 3; ++q) x[3 + q] = in->hist_n > q ? in->logerr[q] : 0.f;
  x[6] = in->r_now;
  const float dtr = in->tr_now - in->tr_prev;
  x[7] = (in->r_n > 1 && dtr > 0.f) ? fminf(fmaxf((in->r_now - in->r_prev) / dtr, -10.f), 10.f) : 0.f;
}

struct LogRec {   // one accepted/rejected attempt for training the proposer
  float x[kIn];
  float h_used, h_std, err_norm;
  int accepted, b;
  int after_accept;  // h came from the proposal made after an accepted step
  int attempt;       // the instance's attempt number (orders the log)
};

// The controller's decision for one instance (dp45.cu:117-138), run by one
// thread: accept or reject the attempt, the standard controller's next step,
// the proposer's history, the outputs an accepted step covers.  Returns 1
// when the proposer is to replace the standard next step (x: its inputs,
// h_std: the standard step it corrects).
__device__ int control_decide(BInst *in, double sum, int b, int fixed, const float *t_eval, int n_eval,
                              bool propose, LogRec *log, int *log_count, int log_cap, float (&x)[kIn],
                              float &h_std) {
  in->attempts += 1;
  in->rhs_evals += 6;
  const float t = in->t, h = in->h;
  int accept;
  float hstd = h, norm = 0.f;
  if (fixed) {
    accept = 1;
  } else {
    const float nrm = (float)sqrt(sum);
    norm = (float)((double)nrm / sqrt((double)in->n));
    float factor = (float)(0.9 * pow((double)norm, -0.2));
    in->norm = (double)norm;
    if (norm < 1.f) {
      accept = 1;
      if (factor > 10.f) factor = 10.f;
      hstd = h * factor;
    } else {
      accept = 0;
      if (factor < 1.f) {
        if ((double)factor < 0.2) factor = 0.2f;
        hstd = h * factor;
      }
      if (isnan(norm)) {
        in->error = 1;
        in->active = 0;
      }
    }
  }
  // training log of every attempt (inputs as seen before this attempt,
  // relative to the step it used)
  if (log && !fixed) {
    const int slot = atomicAdd(log_count, 1);
    if (slot < log_cap) {
      LogRec &L = log[slot];
      proposer_features(in, h, L.x);
      L.h_used = h;
      L.h_std = in->h_std;
      L.err_norm = norm;
      L.accepted = accept;
      L.b = b;
      L.after_accept = in->last_accept;
      L.attempt = (int)in->attempts;
    }
  }
  if (in->proposed) {
    if (accept) in->proposer_accepts += 1;
    else in->proposer_rejects += 1;
  }
  in->last_accept = accept;
  in->h = hstd;
  in->h_std = hstd;
  in->proposed = 0;
  if (!accept) {
    in->rejected += 1;
    return 0;
  }
  in->logh[2] = in->logh[1];
  in->logh[1] = in->logh[0];
  in->logh[0] = logf(h);
  in->logerr[2] = in->logerr[1];
  in->logerr[1] = in->logerr[0];
  in->logerr[0] = logf(fmaxf(norm, 1e-12f));
  in->hist_n = min(in->hist_n + 1, 3);
  const float tn = t + h;
  in->t_last = t;
  in->t = tn;
  in->accepted += 1;
  int j = in->eval_i;
  while (j < n_eval && tn >= t_eval[j]) ++j;
  in->pend_lo = in->eval_i;
  in->pend_hi = j;
  in->commit_pending = 1;
  if (!propose || fixed || in->hist_n < 3) return 0;
  proposer_features(in, hstd, x);
  h_std = hstd;
  return 1;
}

// Controller per instance, block b: the error norm from the stage-7
// partials (fixed-order block sum), the decision by thread 0, and with the
// proposer on (PROPOSE) the next h after an accepted step = h_std
// exp(clamp(mlp(x))), the network evaluated by wave 0; after a rejection
// the standard controller's step is kept (fallback).
template <bool PROPOSE>
__global__ __launch_bounds__(kBlock) void bcontrol_kernel(BInst *inst, const double *err_part, int nblk, float t1,
                                                          int fixed, const float *t_eval, int n_eval,
                                                          const uint16_t *mlp_w, LogRec *log, int *log_count,
                                                          int log_cap) {
  const int b = blockIdx.x;
  BInst *in = inst + b;
  if (!in->active || in->commit_pending) return;
  __shared__ double sh[kBlock];
  double s = 0.0;
  for (int q = threadIdx.x; q < nblk; q += kBlock) s += err_part[(size_t)b * nblk + q];
  const double sum = block_sum(s, sh);
  if (threadIdx.x >= kWave) return;
  float x[kIn] = {};
  float hstd = 0.f;
  int use = 0;
  if (threadIdx.x == 0)
    use = control_decide(in, sum, b, fixed, t_eval, n_eval, PROPOSE, log, log_count, log_cap, x, hstd);
  if constexpr (PROPOSE) {
    if (__shfl(use, 0, kWave)) {
#pragma unroll
      for (int j = 0; j < kIn; ++j) x[j] = __shfl(x[j], 0, kWave);
      float h1, h2;
      const float o = mlp_wave_forward(mlp_w, x, h1, h2);
      if (threadIdx.x == 0) {
        in->h = hstd * expf(fminf(fmaxf(o, -kProposerClamp), kProposerClamp));
        in->proposed = 1;
      }
    }
  }
}

// Dense output (order parameter only), commit, next stage-2 vector.
struct BEmit {
  BInst *inst;
  long ldn;
  int B;
  float *y, *ylast, *ynew, *kbuf;
  float *cs_out;
  const float *t_eval;
  double *ord_part;   // [B][ring_cap][nblk][2]
  int ring_cap, nblk;
  VtOut vt;
  int write_vt;
};

template <class T>
__global__ __launch_bounds__(kBlock) void bemit_kernel(const BEmit a) {
  const int b = blockIdx.y;
  const BInst *in = a.inst + b;
  if (!in->active) return;
  __shared__ double sh[kBlock];
  const long n = (long)blockIdx.x * kBlock + threadIdx.x;
  const bool valid = n < in->n;
  const size_t vb = (size_t)b * a.ldn;
  const size_t kst = (size_t)a.B * a.ldn;
  const float h = in->h;
  auto stage2 = [&](float yv, float k1) {
    const float yt = rfma<float>(k1 * (float)tab::A[0][0], h, yv);
    const float c = rcos<float>(yt), s = rsin<float>(yt);
    a.cs_out[2 * vb + n] = c;
    a.cs_out[2 * vb + a.ldn + n] = s;
    if (a.write_vt) put_vt<T>(a.vt, in->vrow, n, c, s);
  };
  if (!in->commit_pending) {
    if (valid) stage2(a.y[vb + n], a.kbuf[vb + n]);
    return;
  }
  const int pend = in->pend_hi - in->pend_lo;
  const int room = a.ring_cap - in->ring_count;
  const int n_emit = pend < room ? pend : room;
  if (n_emit > 0) {
    const float t = in->t, tl = in->t_last, hh = t - tl;
    float Q[4] = {0, 0, 0, 0}, yl = 0;
    if (valid) {
      float kv[7];
#pragma unroll
      for (int j = 0; j < 7; ++j) kv[j] = a.kbuf[j * kst + vb + n];
      dense_coeffs<float>(kv, Q);
      yl = a.y[vb + n];
    }
    for (int q = 0; q < n_emit; ++q) {
      const int slot = in->ring_count + q;
      const float x = (a.t_eval[in->pend_lo + q] - tl) / hh;
      double cx = 0, sx = 0;
      if (valid) {
        const float val = dense_value<float>(Q, yl, hh, x);
        cx = (double)rcos<float>(val);
        sx = (double)rsin<float>(val);
      }
      const double bc = block_sum(cx, sh), bs = block_sum(sx, sh);
      if (threadIdx.x == 0) {
        const size_t o = (((size_t)b * a.ring_cap + slot) * a.nblk + blockIdx.x) * 2;
        a.ord_part[o] = bc;
        a.ord_part[o + 1] = bs;
      }
    }
  }
  if (n_emit == pend && valid) {
    const float yo = a.y[vb + n], yn = a.ynew[vb + n], k7 = a.kbuf[6 * kst + vb + n];
    a.ylast[vb + n] = yo;
    a.y[vb + n] = yn;
    a.kbuf[vb + n] = k7;
    stage2(yn, k7);
  }
}

// r for the emitted rows, ring bookkeeping, end of run per instance; also
// the last two emitted outputs' r and times for the proposer's inputs.
__global__ __launch_bounds__(kBlock) void bfinalize_kernel(BInst *inst, const double *ord_part, int nblk, int ring_cap,
                                                           float t1, const float *t_eval, float *r_ring) {
  const int b = blockIdx.x;
  BInst *in = inst + b;
  if (!in->active || !in->commit_pending) return;
  __shared__ double sh[kBlock];
  const int pend = in->pend_hi - in->pend_lo;
  const int room = ring_cap - in->ring_count;
  const int n_emit = pend < room ? pend : room;
  const int base = in->ring_count;
  float r_last = in->r_now, r_before = in->r_prev;
  for (int q = 0; q < n_emit; ++q) {
    const int slot = base + q;
    double sx = 0, sy = 0;
    for (int x = threadIdx.x; x < nblk; x += kBlock) {
      const size_t o = (((size_t)b * ring_cap + slot) * nblk + x) * 2;
      sx += ord_part[o];
      sy += ord_part[o + 1];
    }
    const double X = block_sum(sx, sh), Y = block_sum(sy, sh);
    const float r = order_value<float>(X, Y, in->n);
    if (threadIdx.x == 0) r_ring[(size_t)b * ring_cap + slot] = r;
    r_before = r_last;
    r_last = r;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (n_emit > 0) {
      const int last = in->pend_lo + n_emit - 1;
      in->tr_prev = n_emit > 1 ? t_eval[last - 1] : in->tr_now;
      in->tr_now = t_eval[last];
      in->r_prev = r_before;
      in->r_now = r_last;
      in->r_n = min(in->r_n + n_emit, 2);
    }
    in->ring_count = base + n_emit;
    in->pend_lo += n_emit;
    if (in->pend_lo == in->pend_hi) {
      in->commit_pending = 0;
      in->eval_i = in->pend_hi;
      if (!(in->t < t1)) in->active = 0;
    }
  }
}

}  // namespace kmt

using namespace kmt;

static inline long rup(long x, long m) { return (x + m - 1) / m * m; }

// ---------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------
struct kmt_batch {
  int device = 0, B = 0, shared = 0, storage = KMT_ADJ_BF16, fixed = 0;
  float atl = 1e-6f, rtl = 0.f;
  std::vector<long> Nb;
  long Nmax = 0, ldn = 0, lda_shared = 0, rows_shared = 0;
  int mpad = 32, mt = 1, nblk = 0;
  // shared mode: the product runs on the rows of the active instances only.
  // Instance b's cos sits in row inst[b].vrow of V^T and of the product, its
  // sin in row vhalf + vrow; mt_act 32-row tiles (mpad_act rows) are used.
  // kmt_batch_begin starts from vrow = b, vhalf = B; when finished instances
  // leave fewer tiles' worth of rows, compact() renumbers the active ones.
  int mt_act = 1, mpad_act = 32, vhalf = 0, compaction = 1;
  hipStream_t stream = nullptr;
  float *y = nullptr, *ylast = nullptr, *ynew = nullptr, *om = nullptr, *kbuf = nullptr, *cs[2] = {nullptr, nullptr};
  float *P = nullptr;
  size_t P_elems = 0;
  uint16_t *vthi = nullptr, *vtlo = nullptr;
  double *err_part = nullptr, *ord_part = nullptr;
  float *r_ring = nullptr, *t_eval = nullptr;
  BInst *inst = nullptr;
  std::vector<BInst> host_inst;
  void *A = nullptr;
  size_t A_bytes = 0;
  std::vector<size_t> a_off;
  std::vector<long> a_ld;
  std::vector<float> mscale;
  size_t *a_off_d = nullptr;
  long *a_ld_d = nullptr;
  GroupItem *items = nullptr;
  int n_items = 0;
  int n_eval = 0, ring_cap = 0;
  float t1 = 0.f;
  int begun = 0;
  // tiling (shared) / rows (grouped)
  int tile_var = 3, ksplit = 8, lo = 1;
  double tune_ms[96] = {0};  // (tile_var 0..5) x (split 1..16)
  hipGraphExec_t g1 = nullptr, g8 = nullptr;
  // proposer
  uint16_t *mlp_w = nullptr;
  int mlp_on = 0;
  LogRec *log = nullptr;
  int *log_count = nullptr;
  int log_cap = 0;

  ~kmt_batch() { release(); }
  // Destroy the captured attempt graphs (after the stream has drained, as
  // every caller already ensures).
  void drop_graphs() {
    if ((g1 || g8) && stream) (void)hipStreamSynchronize(stream);
    if (g1) (void)hipGraphExecDestroy(g1);
    if (g8) (void)hipGraphExecDestroy(g8);
    g1 = g8 = nullptr;
  }
  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graphs();
    void *ptrs[] = {y, ylast, ynew, om, kbuf, cs[0], cs[1], P, vthi, vtlo, err_part, ord_part, r_ring, t_eval,
                    inst, A, a_off_d, a_ld_d, items, mlp_w, log, log_count};
    for (void *p : ptrs)
      if (p) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }
  template <class T> int alloc(T **p, size_t n) {
    HIP_TRY(hipMalloc((void **)p, sizeof(T) * std::max<size_t>(n, 1)));
    HIP_TRY(hipMemsetAsync(*p, 0, sizeof(T) * std::max<size_t>(n, 1), stream));
    return KMT_OK;
  }

  int create(int dev, int nb, const long *N, double a, double r, int shr, int stg, int fx) {
    device = dev;
    B = nb;
    shared = shr;
    storage = stg;
    fixed = fx;
    atl = (float)a;
    rtl = (float)r;
    Nb.assign(N, N + B);
    Nmax = *std::max_element(Nb.begin(), Nb.end());
    if (shared)
      for (long n : Nb)
        if (n != Nb[0]) return kmt_fail(KMT_ERR_ARG, "shared adjacency needs equal N for all instances");
    if (shared && storage == KMT_ADJ_NATIVE) return kmt_fail(KMT_ERR_ARG, "shared (MFMA) batches store A in bf16 or f16");
    if (shared && 2 * B > 128) return kmt_fail(KMT_ERR_ARG, "shared batches hold at most 64 instances per context");
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    ldn = rup(Nmax, shr ? 512 : 256);
    nblk = (int)(ldn / kBlock);
    KMT_TRY(alloc(&y, (size_t)B * ldn));
    KMT_TRY(alloc(&ylast, (size_t)B * ldn));
    KMT_TRY(alloc(&ynew, (size_t)B * ldn));
    KMT_TRY(alloc(&om, (size_t)B * ldn));
    KMT_TRY(alloc(&kbuf, (size_t)7 * B * ldn));
    KMT_TRY(alloc(&cs[0], (size_t)2 * B * ldn));
    KMT_TRY(alloc(&cs[1], (size_t)2 * B * ldn));
    KMT_TRY(alloc(&err_part, (size_t)B * nblk));
    KMT_TRY(alloc(&inst, (size_t)B));
    host_inst.assign(B, BInst{});
    for (int b = 0; b < B; ++b) {
      std::memset(&host_inst[b], 0, sizeof(BInst));
      host_inst[b].n = (int)Nb[b];
      host_inst[b].mscale = 1.f;
      host_inst[b].vrow = b;
    }
    mscale.assign(B, 1.f);
    if (shared) {
      mt = (2 * B + 31) / 32;  // 32-row MFMA tiles of V^T: 1..4
      mpad = 32 * mt;
      mt_act = mt;
      mpad_act = mpad;
      vhalf = B;
      lda_shared = rup(Nmax, 64);
      rows_shared = ldn;  // multiple of 512: whole workgroups of every product variant
      KMT_TRY(alloc(&vthi, (size_t)mpad * lda_shared));
      KMT_TRY(alloc(&vtlo, (size_t)mpad * lda_shared));
    }
    HIP_TRY(hipStreamSynchronize(stream));
    return KMT_OK;
  }

  int set_instance(int b, const float *omega, const float *theta, double c1, double t0, double h0) {
    if (b < 0 || b >= B) return kmt_fail(KMT_ERR_ARG, "instance %d out of range", b);
    HIP_TRY(hipMemcpyAsync(om + (size_t)b * ldn, omega, sizeof(float) * Nb[b], hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(y + (size_t)b * ldn, theta, sizeof(float) * Nb[b], hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    host_inst[b].c1 = (float)c1;
    host_inst[b].t = (float)t0;
    host_inst[b].h = (float)h0;
    begun = 0;
    return KMT_OK;
  }

  size_t ebytes() const { return storage == KMT_ADJ_NATIVE ? 4 : 2; }

  int ensure_A() {
    if (A) return KMT_OK;
    a_off.assign(B, 0);
    a_ld.assign(B, 0);
    size_t total = 0;
    if (shared) {
      total = (size_t)rows_shared * lda_shared;
      for (int b = 0; b < B; ++b) a_ld[b] = lda_shared;
    } else {
      for (int b = 0; b < B; ++b) {
        a_off[b] = total;
        a_ld[b] = rup(Nb[b], 64);
        total += (size_t)rup(Nb[b], 64) * a_ld[b];
      }
    }
    A_bytes = total * ebytes();
    hipError_t e = hipMalloc(&A, A_bytes);
    if (e != hipSuccess) {
      A = nullptr;
      return kmt_fail(KMT_ERR_MEMORY, "cannot allocate %zu bytes of adjacency: %s", A_bytes, hipGetErrorString(e));
    }
    HIP_TRY(hipMemsetAsync(A, 0, A_bytes, stream));
    KMT_TRY(alloc(&a_off_d, (size_t)B));
    KMT_TRY(alloc(&a_ld_d, (size_t)B));
    HIP_TRY(hipMemcpyAsync(a_off_d, a_off.data(), sizeof(size_t) * B, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(a_ld_d, a_ld.data(), sizeof(long) * B, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return KMT_OK;
  }

  // dense adjacency of instance b (or the shared one, b = -1) from a factor
  int set_factor(int b, const float *Xh, const float *wh, int K) {
    if (shared ? b != -1 : (b < 0 || b >= B)) return kmt_fail(KMT_ERR_ARG, "instance %d: use -1 for the shared adjacency", b);
    KMT_TRY(ensure_A());
    const long N = shared ? Nmax : Nb[b];
    double mx = 0;
    for (long n = 0; n < N; ++n) {
      double s = 0;
      for (int m = 0; m < K; ++m) s += std::fabs((double)wh[m]) * (double)Xh[n * K + m] * (double)Xh[n * K + m];
      mx = std::max(mx, s);
    }
    const double sc = (storage == KMT_ADJ_NATIVE || !(mx > 0)) ? 1.0 : std::ldexp(1.0, (int)std::ceil(std::log2(mx)));
    float *Xd = nullptr, *wd = nullptr;
    KMT_TRY(alloc(&Xd, (size_t)N * std::max(K, 1)));
    KMT_TRY(alloc(&wd, (size_t)std::max(K, 1)));
    hipError_t e = hipMemcpyAsync(Xd, Xh, sizeof(float) * N * K, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(wd, wh, sizeof(float) * K, hipMemcpyHostToDevice, stream);
    const int bi = shared ? 0 : b;
    const long lda = a_ld[bi];
    const float inv = (float)(1.0 / sc);
    if (shared && e == hipSuccess) {
      const long nkc = lda_shared / kKC, ntiles = (rows_shared / 32) * nkc;
      for (long t0 = 0; t0 < ntiles && e == hipSuccess; t0 += 65536) {
        const unsigned nb = (unsigned)std::min<long>(65536, ntiles - t0);
        if (storage == KMT_ADJ_BF16)
          gen_adj_tiled_kernel<bf16_t><<<nb, kBlock, 0, stream>>>(Xd, wd, K, N, nkc, t0, (bf16_t *)A, inv);
        else
          gen_adj_tiled_kernel<f16_t><<<nb, kBlock, 0, stream>>>(Xd, wd, K, N, nkc, t0, (f16_t *)A, inv);
        e = hipGetLastError();
      }
    }
    for (long r0 = 0; !shared && r0 < N && e == hipSuccess; r0 += 4096) {
      const long nr = std::min<long>(4096, N - r0);
      if (storage == KMT_ADJ_BF16)
        gen_adj_kernel<float, bf16_t><<<(unsigned)nr, kBlock, 0, stream>>>(Xd, wd, K, N, r0, (bf16_t *)A + a_off[bi] + (size_t)r0 * lda, lda, inv);
      else if (storage == KMT_ADJ_F16)
        gen_adj_kernel<float, f16_t><<<(unsigned)nr, kBlock, 0, stream>>>(Xd, wd, K, N, r0, (f16_t *)A + a_off[bi] + (size_t)r0 * lda, lda, inv);
      else
        gen_adj_kernel<float, float><<<(unsigned)nr, kBlock, 0, stream>>>(Xd, wd, K, N, r0, (float *)A + a_off[bi] + (size_t)r0 * lda, lda, 1.f);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    (void)hipFree(Xd);
    (void)hipFree(wd);
    if (e != hipSuccess) return kmt_fail(KMT_ERR_HIP, "adjacency build: %s", hipGetErrorString(e));
    if (shared)
      for (int q = 0; q < B; ++q) { mscale[q] = (float)sc; host_inst[q].mscale = (float)sc; }
    else { mscale[b] = (float)sc; host_inst[b].mscale = (float)sc; }
    drop_graphs();
    return KMT_OK;
  }

  int get_adjacency_rows(int b, long i0, long i1, float *out) {
    if (!A) return kmt_fail(KMT_ERR_STATE, "no adjacency");
    const int bi = shared ? 0 : b;
    const long N = shared ? Nmax : Nb[b];
    if (i0 < 0 || i1 > N || i0 > i1) return kmt_fail(KMT_ERR_ARG, "bad row range");
    const long lda = a_ld[bi];
    float *stage = nullptr;
    KMT_TRY(alloc(&stage, (size_t)(i1 - i0) * N));
    dim3 grid((unsigned)std::min<long>((N + kBlock - 1) / kBlock, 64), (unsigned)(i1 - i0));
    if (shared) {
      if (storage == KMT_ADJ_BF16)
        read_tiled_rows_kernel<bf16_t><<<grid, kBlock, 0, stream>>>((const bf16_t *)A, lda_shared / kKC, N, i0, i1 - i0, mscale[0], stage);
      else
        read_tiled_rows_kernel<f16_t><<<grid, kBlock, 0, stream>>>((const f16_t *)A, lda_shared / kKC, N, i0, i1 - i0, mscale[0], stage);
    } else if (storage == KMT_ADJ_BF16)
      read_rows_kernel<float, bf16_t><<<grid, kBlock, 0, stream>>>((const bf16_t *)A + a_off[bi] + (size_t)i0 * lda, lda, N, i1 - i0, mscale[bi], stage);
    else if (storage == KMT_ADJ_F16)
      read_rows_kernel<float, f16_t><<<grid, kBlock, 0, stream>>>((const f16_t *)A + a_off[bi] + (size_t)i0 * lda, lda, N, i1 - i0, mscale[bi], stage);
    else
      read_rows_kernel<float, float><<<grid, kBlock, 0, stream>>>((const float *)A + a_off[bi] + (size_t)i0 * lda, lda, N, i1 - i0, 1.f, stage);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, stage, sizeof(float) * (i1 - i0) * N, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    (void)hipFree(stage);
    if (e != hipSuccess) return kmt_fail(KMT_ERR_HIP, "adjacency read: %s", hipGetErrorString(e));
    return KMT_OK;
  }

  VtOut vt_out() const {
    VtOut v;
    v.hi = vthi;
    v.lo = vtlo;
    v.ldk = lda_shared;
    v.half = vhalf;
    return v;
  }

  // ---- coupling launches ----
  // shared product variants (tile_var): 0 = 4 waves x 1 band (128 rows of A
  // per workgroup), 1 = 8 waves x 1 band (256), 2 = 8 waves x 2 bands (512),
  // 3 / 4 = 8 waves x 1 band with two / three chunks of A in flight,
  // 5 = 4 waves x 1 band with two chunks in flight
  template <class T, int MT, int NT, int WAVES, bool LO, int PF, int VD> void gemm_v() {
    const long nkc = lda_shared / kKC;
    const long slab_chunks = (nkc + ksplit - 1) / ksplit;
    const long nrb = rows_shared / (32L * NT * WAVES);
    bgemm_kernel<T, MT, NT, WAVES, LO, PF, VD><<<(unsigned)(nrb * ksplit), WAVES * kWave, 0, stream>>>(
        (const T *)A, nkc, vthi, vtlo, lda_shared, ksplit, slab_chunks, P, ldn, mpad_act);
  }
  template <class T, int MT, bool LO> void gemm_lo() {
    switch (tile_var) {
      case 0: gemm_v<T, MT, 1, 4, LO, 1, 2>(); break;
      case 1: gemm_v<T, MT, 1, 8, LO, 1, 2>(); break;
      case 3: gemm_v<T, MT, 1, 8, LO, 2, 2>(); break;
      case 4: gemm_v<T, MT, 1, 8, LO, 3, 2>(); break;
      case 5: gemm_v<T, MT, 1, 4, LO, 2, 1>(); break;
      default: gemm_v<T, MT, 2, 8, LO, 1, 1>(); break;
    }
  }
  template <class T, int MT> void gemm_mt() {
    if (lo) gemm_lo<T, MT, true>(); else gemm_lo<T, MT, false>();
  }
  template <class T> void gemm_t() {
    if (mt_act == 1) gemm_mt<T, 1>();
    else if (mt_act == 2) gemm_mt<T, 2>();
    else if (mt_act == 3) gemm_mt<T, 3>();
    else gemm_mt<T, 4>();
  }
  void coupling(int check) {
    if (shared) {
      if (storage == KMT_ADJ_BF16) gemm_t<bf16_t>(); else gemm_t<f16_t>();
    } else {
      if (storage == KMT_ADJ_BF16)
        bgemv_kernel<bf16_t, 8><<<n_items, kBlock, 0, stream>>>((const bf16_t *)A, a_off_d, a_ld_d, items, inst, cs_cur, ldn, P, B, check);
      else if (storage == KMT_ADJ_F16)
        bgemv_kernel<f16_t, 8><<<n_items, kBlock, 0, stream>>>((const f16_t *)A, a_off_d, a_ld_d, items, inst, cs_cur, ldn, P, B, check);
      else
        bgemv_kernel<float, 8><<<n_items, kBlock, 0, stream>>>((const float *)A, a_off_d, a_ld_d, items, inst, cs_cur, ldn, P, B, check);
    }
  }
  const float *cs_cur = nullptr;

  BVecs bvecs(const float *in, float *out) {
    BVecs v;
    v.cs_in = in;
    v.cs_out = out;
    v.ldn = ldn;
    v.B = B;
    v.omega = om;
    v.y = y;
    v.ynew = ynew;
    v.kbuf = kbuf;
    v.f_out = kbuf;  // the RHS kind writes k1
    v.P = P;
    v.ksplit = shared ? ksplit : 1;
    v.mrows = shared ? mpad_act : 2 * B;
    v.phalf = shared ? vhalf : B;
    v.inst = inst;
    v.err_part = err_part;
    v.nblk = nblk;
    v.atl = atl;
    v.rtl = rtl;
    v.vt = vt_out();
    v.write_vt = shared;
    return v;
  }

  template <class T> void epi(int kind, const BVecs &v) {
    dim3 grid((unsigned)nblk, (unsigned)B);
    switch (kind) {
      case 0: bepilogue_kernel<T, 0><<<grid, kBlock, 0, stream>>>(v); break;
      case 2: bepilogue_kernel<T, 2><<<grid, kBlock, 0, stream>>>(v); break;
      case 3: bepilogue_kernel<T, 3><<<grid, kBlock, 0, stream>>>(v); break;
      case 4: bepilogue_kernel<T, 4><<<grid, kBlock, 0, stream>>>(v); break;
      case 5: bepilogue_kernel<T, 5><<<grid, kBlock, 0, stream>>>(v); break;
      case 6: bepilogue_kernel<T, 6><<<grid, kBlock, 0, stream>>>(v); break;
      default: bepilogue_kernel<T, 7><<<grid, kBlock, 0, stream>>>(v); break;
    }
  }
  void epilogue(int kind, const BVecs &v) {
    if (storage == KMT_ADJ_F16) epi<f16_t>(kind, v);
    else epi<bf16_t>(kind, v);
  }

  BEmit bemit_args() {
    BEmit e;
    e.inst = inst;
    e.ldn = ldn;
    e.B = B;
    e.y = y;
    e.ylast = ylast;
    e.ynew = ynew;
    e.kbuf = kbuf;
    e.cs_out = cs[0];
    e.t_eval = t_eval;
    e.ord_part = ord_part;
    e.ring_cap = ring_cap;
    e.nblk = nblk;
    e.vt = vt_out();
    e.write_vt = shared;
    return e;
  }
  void emit() {
    dim3 grid((unsigned)nblk, (unsigned)B);
    if (storage == KMT_ADJ_F16) bemit_kernel<f16_t><<<grid, kBlock, 0, stream>>>(bemit_args());
    else bemit_kernel<bf16_t><<<grid, kBlock, 0, stream>>>(bemit_args());
  }

  void launch_attempt(std::vector<std::pair<hipEvent_t, hipEvent_t>> *ev = nullptr) {
    for (int s = 2; s <= 7; ++s) {
      const float *in = cs[(s - 2) & 1];
      float *out = cs[(s - 1) & 1];
      cs_cur = in;
      if (ev) (void)hipEventRecord((*ev)[0].first, stream);
      coupling(1);
      if (ev) {
        (void)hipEventRecord((*ev)[0].second, stream);
        ev->erase(ev->begin());
      }
      epilogue(s, bvecs(in, out));
    }
    if (mlp_on && mlp_w)
      bcontrol_kernel<true><<<B, kBlock, 0, stream>>>(inst, err_part, nblk, t1, fixed, t_eval, n_eval, mlp_w, log,
                                                      log_count, log_cap);
    else
      bcontrol_kernel<false><<<B, kBlock, 0, stream>>>(inst, err_part, nblk, t1, fixed, t_eval, n_eval, nullptr, log,
                                                       log_count, log_cap);
    emit();
    bfinalize_kernel<<<B, kBlock, 0, stream>>>(inst, ord_part, nblk, ring_cap, t1, t_eval, r_ring);
  }

  int build_graphs() {
    drop_graphs();
    for (int which = 0; which < 2; ++which) {
      hipGraph_t g = nullptr;
      HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      for (int a = 0; a < (which ? 8 : 1); ++a) launch_attempt();
      HIP_TRY(hipStreamEndCapture(stream, &g));
      hipGraphExec_t ex = nullptr;
      hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return kmt_fail(KMT_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
      (which ? g8 : g1) = ex;
    }
    return KMT_OK;
  }

  int ensure_P() {
    const size_t need = shared ? (size_t)16 * mpad * ldn : (size_t)2 * B * ldn;
    if (P && P_elems >= need) return KMT_OK;
    if (P) (void)hipFree(P);
    P = nullptr;
    KMT_TRY(alloc(&P, need));
    P_elems = need;
    return KMT_OK;
  }

  int build_items() {
    if (shared || items) return KMT_OK;
    std::vector<GroupItem> it;
    for (int b = 0; b < B; ++b)
      for (long r = 0; r < Nb[b]; r += 8) it.push_back(GroupItem{b, (int)r});
    n_items = (int)it.size();
    KMT_TRY(alloc(&items, it.size()));
    HIP_TRY(hipMemcpyAsync(items, it.data(), sizeof(GroupItem) * it.size(), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return KMT_OK;
  }

  int begin(const double *te, int ne, double t1_, int rcap) {
    if (!A) return kmt_fail(KMT_ERR_STATE, "no adjacency: call kmt_batch_set_factor first");
    if (rcap < 1) return kmt_fail(KMT_ERR_ARG, "ring_rows must be >= 1");
    HIP_TRY(hipStreamSynchronize(stream));
    drop_graphs();
    KMT_TRY(ensure_P());
    KMT_TRY(build_items());
    if (t_eval) (void)hipFree(t_eval);
    if (ord_part) (void)hipFree(ord_part);
    if (r_ring) (void)hipFree(r_ring);
    t_eval = nullptr; ord_part = nullptr; r_ring = nullptr;
    n_eval = ne;
    ring_cap = rcap;
    t1 = (float)t1_;
    std::vector<float> tf(std::max(ne, 1));
    for (int q = 0; q < ne; ++q) tf[q] = (float)te[q];
    KMT_TRY(alloc(&t_eval, tf.size()));
    HIP_TRY(hipMemcpyAsync(t_eval, tf.data(), sizeof(float) * tf.size(), hipMemcpyHostToDevice, stream));
    KMT_TRY(alloc(&ord_part, (size_t)B * ring_cap * nblk * 2));
    KMT_TRY(alloc(&r_ring, (size_t)B * ring_cap));
    for (int b = 0; b < B; ++b) {
      BInst &in = host_inst[b];
      const float t0 = in.t, h0 = in.h, c1 = in.c1, ms = in.mscale;
      const int n = in.n;
      std::memset(&in, 0, sizeof(BInst));
      in.t = t0; in.h = h0; in.c1 = c1; in.mscale = ms; in.n = n;
      in.t_last = t0; in.h_std = h0;
      in.active = t0 < t1 ? 1 : 0;
      in.rhs_evals = 1;
      in.last_accept = 1;
      in.vrow = b;
    }
    if (shared) {
      mt_act = mt;
      mpad_act = mpad;
      vhalf = B;
    }
    HIP_TRY(hipMemcpyAsync(inst, host_inst.data(), sizeof(BInst) * B, hipMemcpyHostToDevice, stream));
    // k1 = f(y) for every instance, then the first stage-2 vectors
    dim3 g2((unsigned)nblk, (unsigned)B);
    if (storage == KMT_ADJ_F16) bsincos_kernel<f16_t><<<g2, kBlock, 0, stream>>>(y, inst, ldn, cs[1], vt_out(), shared);
    else bsincos_kernel<bf16_t><<<g2, kBlock, 0, stream>>>(y, inst, ldn, cs[1], vt_out(), shared);
    cs_cur = cs[1];
    coupling(0);
    epilogue(KIND_RHS, bvecs(cs[1], nullptr));
    emit();
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(stream));
    begun = 1;
    return KMT_OK;
  }

  int fetch() {
    HIP_TRY(hipMemcpyAsync(host_inst.data(), inst, sizeof(BInst) * B, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return KMT_OK;
  }

  void fill(kmt_batch_status *s) {
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    s->t_min = 1e300;
    s->t_max = -1e300;
    for (const BInst &in : host_inst) {
      s->accepted += in.accepted;
      s->rejected += in.rejected;
      s->attempts = std::max<long>(s->attempts, in.attempts);
      s->rhs_evals += in.rhs_evals;
      s->active += in.active;
      s->errors += in.error;
      s->rows_ready = std::max(s->rows_ready, in.ring_count);
      s->t_min = std::min(s->t_min, (double)in.t);
      s->t_max = std::max(s->t_max, (double)in.t);
      s->osc_steps += (double)in.accepted * in.n;
      s->proposer_accepts += in.proposer_accepts;
      s->proposer_rejects += in.proposer_rejects;
    }
  }

  int advance(int n, kmt_batch_status *s) {
    if (!begun) return kmt_fail(KMT_ERR_STATE, "kmt_batch_advance before kmt_batch_begin");
    KMT_TRY(fetch());
    int any = 0;
    for (const BInst &in : host_inst) any |= in.active;
    if (any && n > 0 && shared && compaction) KMT_TRY(compact());
    if (any && n > 0) {
      if (!g1) KMT_TRY(build_graphs());
      for (; n >= 8; n -= 8) HIP_TRY(hipGraphLaunch(g8, stream));
      for (; n > 0; --n) HIP_TRY(hipGraphLaunch(g1, stream));
    }
    KMT_TRY(fetch());
    fill(s);
    return KMT_OK;
  }

  // Renumber the active instances' rows when they fit in fewer 32-row tiles
  // (called between host synchronisations, nothing in flight).  Each output
  // row of the product depends only on its own row of V^T, so results are
  // bit-identical to the uncompacted run; only the finished instances' rows
  // stop being multiplied.  Graphs are rebuilt for the smaller product.
  int compact() {
    int act = 0;
    for (const BInst &in : host_inst) act += in.active ? 1 : 0;
    const int need = std::max(1, (2 * act + 31) / 32);
    if (act == 0 || need >= mt_act) return KMT_OK;
    int row = 0;
    for (int b = 0; b < B; ++b) {
      host_inst[b].vrow = host_inst[b].active ? row++ : -1;
      HIP_TRY(hipMemcpyAsync(&inst[b].vrow, &host_inst[b].vrow, sizeof(int), hipMemcpyHostToDevice, stream));
    }
    vhalf = act;
    mt_act = need;
    mpad_act = 32 * need;
    // the next attempt's first product reads the stage-2 vectors, which
    // cs[0] holds in FP32: write them into the new rows
    dim3 g2((unsigned)nblk, (unsigned)B);
    if (storage == KMT_ADJ_F16) bvt_rows_kernel<f16_t><<<g2, kBlock, 0, stream>>>(cs[0], inst, ldn, vt_out());
    else bvt_rows_kernel<bf16_t><<<g2, kBlock, 0, stream>>>(cs[0], inst, ldn, vt_out());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(stream));
    drop_graphs();
    ++compactions;
    return KMT_OK;
  }
  int compactions = 0;

  int read_order(int b, int max_n, float *r, int *nout) {
    if (b < 0 || b >= B) return kmt_fail(KMT_ERR_ARG, "instance out of range");
    KMT_TRY(fetch());
    const int ready = host_inst[b].ring_count;
    const int n = std::min(std::max(max_n, 0), ready);
    std::vector<float> buf(std::max(ready, 1));
    if (ready) {
      HIP_TRY(hipMemcpyAsync(buf.data(), r_ring + (size_t)b * ring_cap, sizeof(float) * ready, hipMemcpyDeviceToHost, stream));
      HIP_TRY(hipStreamSynchronize(stream));
    }
    for (int q = 0; q < n; ++q) r[q] = buf[q];
    if (n < ready) HIP_TRY(hipMemcpyAsync(r_ring + (size_t)b * ring_cap, buf.data() + n, sizeof(float) * (ready - n), hipMemcpyHostToDevice, stream));
    const int left = ready - n;
    HIP_TRY(hipMemcpyAsync(&inst[b].ring_count, &left, sizeof(int), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    host_inst[b].ring_count = left;
    if (nout) *nout = n;
    return KMT_OK;
  }

  int get_instance(int b, float *theta, double *st) {
    if (b < 0 || b >= B) return kmt_fail(KMT_ERR_ARG, "instance out of range");
    KMT_TRY(fetch());
    const BInst &in = host_inst[b];
    if (theta) {
      const float *src = (begun && in.commit_pending ? ynew : y) + (size_t)b * ldn;
      HIP_TRY(hipMemcpyAsync(theta, src, sizeof(float) * Nb[b], hipMemcpyDeviceToHost, stream));
      HIP_TRY(hipStreamSynchronize(stream));
    }
    if (st) {
      st[0] = in.t; st[1] = in.h; st[2] = (double)in.accepted; st[3] = (double)in.rejected;
      st[4] = in.active; st[5] = in.error; st[6] = (double)in.proposer_accepts; st[7] = (double)in.proposer_rejects;
    }
    return KMT_OK;
  }

  int time_attempts(int attempts, double *region_ms, double *coupling_ms, kmt_batch_status *s) {
    if (!begun) return kmt_fail(KMT_ERR_STATE, "kmt_batch_time_attempts before kmt_batch_begin");
    std::vector<std::pair<hipEvent_t, hipEvent_t>> all(6 * attempts);
    for (auto &p : all) {
      HIP_TRY(hipEventCreate(&p.first));
      HIP_TRY(hipEventCreate(&p.second));
    }
    hipEvent_t r0, r1;
    HIP_TRY(hipEventCreate(&r0));
    HIP_TRY(hipEventCreate(&r1));
    auto work = all;
    HIP_TRY(hipEventRecord(r0, stream));
    for (int a = 0; a < attempts; ++a) launch_attempt(&work);
    HIP_TRY(hipEventRecord(r1, stream));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventSynchronize(r1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, r0, r1));
    double tot = 0;
    for (auto &p : all) {
      float m = 0;
      HIP_TRY(hipEventElapsedTime(&m, p.first, p.second));
      tot += m;
      (void)hipEventDestroy(p.first);
      (void)hipEventDestroy(p.second);
    }
    (void)hipEventDestroy(r0);
    (void)hipEventDestroy(r1);
    if (region_ms) *region_ms = ms;
    if (coupling_ms) *coupling_ms = tot / all.size();
    KMT_TRY(fetch());
    fill(s);
    return KMT_OK;
  }

  // Tile autotuner of the shared (MFMA) product: the six workgroup shapes
  // (tile_var) x split-K 1..16.  The split changes the order in which
  // partial sums are added (results agree to FP32 rounding); the workgroup
  // shape does not change results.  Splits that are not powers of two
  // matter: the grid is (row blocks x split) workgroups, and the split that
  // makes it just fill a whole number of rounds of resident workgroups
  // leaves the least idle tail (at B = 64, N = 20000, 8-wave shapes: split
  // 6, 474 workgroups, beats split 8, 632).  Each candidate is scored by
  // its measured time plus the time to move its split-K partials (written
  // here, read back by the epilogue: 2 x ksplit x mpad x ldn x 4 bytes at
  // 6 TB/s), so a wider split has to pay for its extra traffic.
  int autotune(int reps, int *best_nt, int *best_split, double *best_ms) {
    if (!shared) return KMT_OK;
    if (!A) return kmt_fail(KMT_ERR_STATE, "no adjacency");
    KMT_TRY(ensure_P());
    if (!begun) {
      // Time the candidates on the vectors of the instances' initial phases,
      // not on the zeroed buffers: the MFMA pipes draw less power on zero
      // operands and then run at higher clocks, which makes every candidate
      // look faster than it is inside a run.  (After kmt_batch_begin the
      // buffers hold the current stage's vectors; begin rewrites them.)
      HIP_TRY(hipMemcpyAsync(inst, host_inst.data(), sizeof(BInst) * B, hipMemcpyHostToDevice, stream));
      dim3 g2((unsigned)nblk, (unsigned)B);
      if (storage == KMT_ADJ_F16) bsincos_kernel<f16_t><<<g2, kBlock, 0, stream>>>(y, inst, ldn, cs[1], vt_out(), 1);
      else bsincos_kernel<bf16_t><<<g2, kBlock, 0, stream>>>(y, inst, ldn, cs[1], vt_out(), 1);
      HIP_TRY(hipGetLastError());
    }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    double best = 1e30;
    int bnt = 3, bks = 8;
    int idx = 0;
    for (int nt = 0; nt < 6; ++nt)
      for (int si = 1; si <= 16; ++si, ++idx) {
        tile_var = nt;
        ksplit = si;
        gemm_t_any();
        HIP_TRY(hipEventRecord(e0, stream));
        for (int r = 0; r < reps; ++r) gemm_t_any();
        HIP_TRY(hipEventRecord(e1, stream));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        tune_ms[idx] = ms / reps;
        const double partial_ms = 2.0 * ksplit * mpad_act * (double)ldn * 4.0 / 6e12 * 1e3;
        const double score = ms / reps + partial_ms;
        if (score < best) { best = score; bnt = nt; bks = ksplit; }
      }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    tile_var = bnt;
    ksplit = bks;
    drop_graphs();
    if (best_nt) *best_nt = bnt;
    if (best_split) *best_split = bks;
    if (best_ms) *best_ms = tune_ms[bnt * 16 + bks - 1];
    return KMT_OK;
  }
  int select_tile(int variant, int split) {
    if (!shared) return kmt_fail(KMT_ERR_STATE, "tile selection applies to shared (MFMA) batches");
    if (variant < 0 || variant > 5) return kmt_fail(KMT_ERR_ARG, "workgroup variant must be 0..5 (got %d)", variant);
    if (split < 1 || split > 16) return kmt_fail(KMT_ERR_ARG, "split_k must be 1..16 (got %d)", split);
    tile_var = variant;
    ksplit = split;
    drop_graphs();
    return KMT_OK;
  }
  void gemm_t_any() {
    if (storage == KMT_ADJ_BF16) gemm_t<bf16_t>(); else gemm_t<f16_t>();
  }

  int set_proposer(const uint16_t *w, int nparams, int on) {
    if (w) {
      if (nparams != kNParams) return kmt_fail(KMT_ERR_ARG, "proposer needs %d parameters (got %d)", kNParams, nparams);
      if (!mlp_w) KMT_TRY(alloc(&mlp_w, (size_t)kNParams));
      HIP_TRY(hipMemcpyAsync(mlp_w, w, sizeof(uint16_t) * kNParams, hipMemcpyHostToDevice, stream));
      HIP_TRY(hipStreamSynchronize(stream));
    }
    if (on && !mlp_w) return kmt_fail(KMT_ERR_STATE, "no proposer weights");
    mlp_on = on;
    drop_graphs();
    return KMT_OK;
  }

  int set_log(int cap) {
    HIP_TRY(hipStreamSynchronize(stream));  // the log below may be in use by queued work
    drop_graphs();
    if (log) { (void)hipFree(log); log = nullptr; }
    if (log_count) { (void)hipFree(log_count); log_count = nullptr; }
    log_cap = std::max(cap, 0);
    if (log_cap) {
      KMT_TRY(alloc(&log, (size_t)log_cap));
      KMT_TRY(alloc(&log_count, 1));
    }
    return KMT_OK;
  }

  int read_log(int max_n, float *rec, int *nout) {
    if (!log) return kmt_fail(KMT_ERR_STATE, "attempt log not enabled");
    int cnt = 0;
    HIP_TRY(hipMemcpyAsync(&cnt, log_count, sizeof(int), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    // The slots are taken in whatever order the instances' controllers run:
    // the records are returned sorted by (instance, attempt), the same on
    // every run.  Past the capacity the kept subset would depend on that
    // order, so an overflowed log is an error.
    if (cnt > log_cap)
      return kmt_fail(KMT_ERR_STATE, "attempt log overflowed (%d attempts, capacity %d)", cnt, log_cap);
    std::vector<LogRec> buf(std::max(cnt, 1));
    if (cnt) {
      HIP_TRY(hipMemcpyAsync(buf.data(), log, sizeof(LogRec) * cnt, hipMemcpyDeviceToHost, stream));
      HIP_TRY(hipStreamSynchronize(stream));
    }
    std::sort(buf.begin(), buf.begin() + cnt, [](const LogRec &u, const LogRec &v) {
      return u.b != v.b ? u.b < v.b : u.attempt < v.attempt;
    });
    const int n = std::min(cnt, std::max(max_n, 0));
    for (int q = 0; q < n; ++q) {
      float *o = rec + (size_t)q * 14;
      for (int j = 0; j < kIn; ++j) o[j] = buf[q].x[j];
      o[8] = buf[q].h_used; o[9] = buf[q].h_std; o[10] = buf[q].err_norm; o[11] = (float)buf[q].accepted;
      o[12] = (float)buf[q].b; o[13] = (float)buf[q].after_accept;
    }
    const int zero = 0;
    HIP_TRY(hipMemcpyAsync(log_count, &zero, sizeof(int), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (nout) *nout = n;
    return KMT_OK;
  }
};

extern "C" {

#define B_CHECK \
  if (!bp) return kmt_fail(KMT_ERR_ARG, "batch is NULL");

int kmt_batch_create(kmt_batch **out, int device, int B, const long *N, double atol, double rtol, int shared,
                     int storage, int fixed_step) {
  if (!out || !N || B < 1) return kmt_fail(KMT_ERR_ARG, "bad batch arguments");
  *out = nullptr;
  for (int b = 0; b < B; ++b)
    if (N[b] < 1) return kmt_fail(KMT_ERR_ARG, "instance %d: N must be >= 1", b);
  if (storage != KMT_ADJ_NATIVE && storage != KMT_ADJ_BF16 && storage != KMT_ADJ_F16)
    return kmt_fail(KMT_ERR_ARG, "unknown storage %d", storage);
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return kmt_fail(KMT_ERR_ARG, "device %d not present", device);
  kmt_batch *bp = new kmt_batch();
  int rc = bp->create(device, B, N, atol, rtol, shared, storage, fixed_step);
  if (rc != KMT_OK) {
    delete bp;
    return rc;
  }
  *out = bp;
  return KMT_OK;
}
int kmt_batch_destroy(kmt_batch *bp) {
  delete bp;
  return KMT_OK;
}
int kmt_batch_set_instance(kmt_batch *bp, int b, const float *omega, const float *theta0, double c1, double t0,
                           double h0) {
  B_CHECK;
  if (!omega || !theta0) return kmt_fail(KMT_ERR_ARG, "NULL omega or theta0");
  return bp->set_instance(b, omega, theta0, c1, t0, h0);
}
int kmt_batch_set_factor(kmt_batch *bp, int b, const float *X, const float *w, int K) {
  B_CHECK;
  if (K < 0 || (K > 0 && (!X || !w))) return kmt_fail(KMT_ERR_ARG, "bad factor");
  return bp->set_factor(b, X, w, K);
}
int kmt_batch_get_adjacency_rows(kmt_batch *bp, int b, long i0, long i1, float *out) {
  B_CHECK;
  return bp->get_adjacency_rows(b, i0, i1, out);
}
int kmt_batch_begin(kmt_batch *bp, const double *t_eval, int n_eval, double t1, int ring_rows) {
  B_CHECK;
  if (n_eval > 0 && !t_eval) return kmt_fail(KMT_ERR_ARG, "t_eval is NULL");
  return bp->begin(t_eval, n_eval, t1, ring_rows);
}
int kmt_batch_advance(kmt_batch *bp, int max_attempts, kmt_batch_status *st) {
  B_CHECK;
  return bp->advance(max_attempts, st);
}
int kmt_batch_read_order(kmt_batch *bp, int b, int max_n, float *r, int *n) {
  B_CHECK;
  return bp->read_order(b, max_n, r, n);
}
int kmt_batch_get_instance(kmt_batch *bp, int b, float *theta, double *state8) {
  B_CHECK;
  return bp->get_instance(b, theta, state8);
}
int kmt_batch_autotune(kmt_batch *bp, int reps, int *n_tiles, int *split_k, double *best_ms) {
  B_CHECK;
  return bp->autotune(reps, n_tiles, split_k, best_ms);
}
int kmt_batch_time_attempts(kmt_batch *bp, int attempts, double *region_ms, double *coupling_ms,
                            kmt_batch_status *st) {
  B_CHECK;
  if (attempts < 1) return kmt_fail(KMT_ERR_ARG, "attempts < 1");
  return bp->time_attempts(attempts, region_ms, coupling_ms, st);
}
int kmt_batch_select_tile(kmt_batch *bp, int variant, int split_k) {
  if (!bp) return kmt_fail(KMT_ERR_ARG, "batch is NULL");
  return bp->select_tile(variant, split_k);
}
int kmt_batch_set_compaction(kmt_batch *bp, int enabled) {
  B_CHECK;
  bp->compaction = enabled ? 1 : 0;
  return KMT_OK;
}
int kmt_batch_product_rows(kmt_batch *bp, int *rows, int *compactions) {
  B_CHECK;
  if (rows) *rows = bp->shared ? bp->mpad_act : 2 * bp->B;
  if (compactions) *compactions = bp->compactions;
  return KMT_OK;
}
int kmt_batch_tune_table(kmt_batch *bp, double *ms96) {
  B_CHECK;
  if (ms96) for (int q = 0; q < 96; ++q) ms96[q] = bp->tune_ms[q];
  return KMT_OK;
}
int kmt_batch_set_vector_split(kmt_batch *bp, int hi_lo) {
  B_CHECK;
  bp->lo = hi_lo ? 1 : 0;
  bp->drop_graphs();
  return KMT_OK;
}
int kmt_batch_set_proposer(kmt_batch *bp, const uint16_t *weights_f16, int n_params, int enabled) {
  B_CHECK;
  return bp->set_proposer(weights_f16, n_params, enabled);
}
int kmt_batch_set_attempt_log(kmt_batch *bp, int capacity) {
  B_CHECK;
  return bp->set_log(capacity);
}
int kmt_batch_read_attempt_log(kmt_batch *bp, int max_n, float *records, int *n) {
  B_CHECK;
  if (!records) return kmt_fail(KMT_ERR_ARG, "records is NULL");
  return bp->read_log(max_n, records, n);
}
int kmt_proposer_param_count(void) { return kNParams; }

}  // extern "C"
