// Single-network engine (device buffers, launch sequences, hipGraph of one /
// eight DP45 attempts, output rings) shared by the per-precision
// translation units kmt_engine_f32.hip / kmt_engine_f64.hip; the C ABI in
// kmt_api.hip reaches it through EngineBase.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/kuramoto_amd.h"
#include "kmt_common.hpp"
#include "kmt_kernels.hpp"
#include "kmt_sign.hpp"
#include "kmt_sym.hpp"

namespace kmt {

#define fail kmt_fail

static inline long round_up(long x, long m) { return (x + m - 1) / m * m; }

// ---------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------
struct EngineBase {
  virtual ~EngineBase() {}
  virtual int create(int, long, double, double, int, double) = 0;
  virtual int set_omega(const void *) = 0;
  virtual int set_state(const void *, double, double) = 0;
  virtual int get_state(void *, double *, double *) = 0;
  virtual int set_adjacency(const void *, int, long, int) = 0;
  virtual int set_factor(const void *, const void *, int) = 0;
  virtual int set_factor_sign(int, uint32_t) = 0;
  virtual int build_adjacency(int) = 0;
  virtual int use_factored(int) = 0;
  virtual int get_adjacency_rows(long, long, void *) = 0;
  virtual int rhs(const void *, void *, void *) = 0;
  virtual int rhs_device(const void *, void *, void *, void *, hipStream_t) = 0;
  virtual size_t rhs_scratch_bytes() const = 0;
  virtual int device_index() const = 0;
  virtual int order(const void *, double *) = 0;
  virtual int begin(const double *, int, double, int, int) = 0;
  virtual int advance(int, kmt_status *) = 0;
  virtual int read_outputs(int, void *, void *, void *, int *) = 0;
  virtual int read_times(int, void *, void *, int *) = 0;
  virtual int step(int *, double *, double *) = 0;
  virtual int eval(double, void *) = 0;
  virtual int time_stage(int, double *) = 0;
  virtual int time_attempts(int, double *, double *, kmt_status *) = 0;
  virtual int layout(long *, int *, int *, int *) = 0;
  virtual int tune(int, double *) = 0;
  virtual int select_variant(int, int) = 0;
  double mscale_ = 1.0;
};

template <class R> struct Engine : EngineBase {
  int device = 0;
  long N = 0, ldv = 0, rows_alloc = 0, lda = 0;
  R atl = 0, rtl = 0, c1 = 0;
  int fixed = 0;
  hipStream_t stream = nullptr;

  // vectors (ldv each unless noted)
  R *y = nullptr, *ylast = nullptr, *ynew = nullptr, *omega = nullptr, *kbuf = nullptr, *k[7] = {};
  R *cs[2] = {nullptr, nullptr};  // 2*ldv each
  R *cs_scr = nullptr, *vec_scr = nullptr, *f_scr = nullptr, *pairs = nullptr, *k_scr = nullptr;  // scratch
  double *err_part = nullptr;
  int n_err_part = 0;
  double *ord_scr = nullptr;  // stand-alone order parameter partials + result
  State<R> *st = nullptr;
  State<R> *st_host = nullptr;  // pinned
  State<R> shadow{};

  // adjacency
  int storage = -1;  // KMT_ADJ_*
  void *A = nullptr;
  size_t A_bytes = 0;
  R *X = nullptr, *w = nullptr;
  R *XT = nullptr;  // transposed factor [K][ldv] for the register-rebuild form
  int K = 0;
  int factored = 0;  // 0 dense A, KMT_LOWRANK_CONTRACT, KMT_LOWRANK_REBUILD
  // KMT_ADJ_SYM: (I, J) of each stored block, and the row / column partials
  long sym_nb = 0, sym_blocks = 0;
  int2 *sym_ij = nullptr;
  SymPair<double> *sym_prow = nullptr, *sym_pcol = nullptr;
  double *lr_part = nullptr, *lr_G = nullptr;
  int lr_nblk = 0, lr_grid = 0;
  // packed +-1 factor (kmt_set_factor_sign): per pass (1: rows m over
  // k = n, 2: rows n over k = m) the sign bits in the pass's tile order
  // (ldp rows of 32 npairs words: the bit rows, the ones row, zero rows),
  // the B fragments of the digits and the int32 partial sums
  int sign = 0;
  struct SignPass {
    long ldp = 0, nsb = 0, npairs = 0, pairs_per = 1;  // nsb = 2 npairs super-blocks of 512
    int nsplit = 1;
    uint32_t *tiles = nullptr;
    uint8_t *bfrag = nullptr;
    int32_t *part = nullptr;
    uint32_t *ones = nullptr;  // the ones row: bits set for k < kdim, 16 nsb words
  } s_pass[2];
  std::vector<R> X_host, w_host;

  // run
  int begun = 0;
  R *t_eval = nullptr;
  int n_eval = 0;
  R *ring = nullptr, *r_ring = nullptr, *times = nullptr;
  double *ord_part = nullptr;
  int ring_cap = 0, keep_rows = 0, times_cap = 4096, nblk_emit = 0;
  hipGraphExec_t g1 = nullptr, g8 = nullptr;
  bool graphs_valid = false;
  int tune_rows = 8, tune_nt = 1, tune_bal = 0, tune_ring = 0, tuned = 0;
  double tune_ms[12] = {0};
  int bal_grid = 0;  // workgroups of the balanced variant (one resident round)
  R t0_ = 0, h0_ = 0;

  ~Engine() override { release(); }

  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graphs();
    R *vecs[] = {y, ylast, ynew, omega, kbuf, cs[0], cs[1], cs_scr, vec_scr, f_scr, pairs, k_scr, X, w, XT, t_eval, ring, r_ring, times};
    for (R *p : vecs)
      if (p) (void)hipFree(p);
    double *dv[] = {err_part, ord_scr, lr_part, lr_G, ord_part};
    drop_sign();
    for (double *p : dv)
      if (p) (void)hipFree(p);
    if (st) (void)hipFree(st);
    if (sym_ij) (void)hipFree(sym_ij);
    if (sym_prow) (void)hipFree(sym_prow);
    if (sym_pcol) (void)hipFree(sym_pcol);
    if (st_host) (void)hipHostFree(st_host);
    if (A) (void)hipFree(A);
    if (stream) (void)hipStreamDestroy(stream);
    y = ylast = ynew = omega = kbuf = nullptr;
    stream = nullptr;
  }

  // Destroy the captured attempt graphs.  Every entry point that launches
  // them synchronises the stream before it returns, so none is in flight
  // here; the synchronisation keeps that true for any later caller.
  void drop_graphs() {
    if ((g1 || g8) && stream) (void)hipStreamSynchronize(stream);
    if (g1) (void)hipGraphExecDestroy(g1);
    if (g8) (void)hipGraphExecDestroy(g8);
    g1 = g8 = nullptr;
    graphs_valid = false;
  }

  int alloc_vec(R **p, long n) {
    HIP_TRY(hipMalloc((void **)p, sizeof(R) * n));
    HIP_TRY(hipMemsetAsync(*p, 0, sizeof(R) * n, stream));
    return KMT_OK;
  }

  int create(int dev, long n, double a, double r, int fx, double c) override {
    device = dev;
    N = n;
    atl = (R)a;
    rtl = (R)r;
    fixed = fx;
    c1 = (R)c;
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    ldv = round_up(N, 64);
    lda = ldv;
    rows_alloc = round_up(N, 64);
    KMT_TRY(alloc_vec(&y, ldv));
    KMT_TRY(alloc_vec(&ylast, ldv));
    KMT_TRY(alloc_vec(&ynew, ldv));
    KMT_TRY(alloc_vec(&omega, ldv));
    KMT_TRY(alloc_vec(&kbuf, 7 * ldv));
    for (int j = 0; j < 7; ++j) k[j] = kbuf + j * ldv;
    KMT_TRY(alloc_vec(&cs[0], 2 * ldv));
    KMT_TRY(alloc_vec(&cs[1], 2 * ldv));
    KMT_TRY(alloc_vec(&cs_scr, 2 * ldv));
    KMT_TRY(alloc_vec(&vec_scr, ldv));
    KMT_TRY(alloc_vec(&f_scr, ldv));
    KMT_TRY(alloc_vec(&pairs, 2 * ldv));
    // one error partial per stage-7 workgroup: at most one per row (every
    // grid - row groups, balanced, low-rank, rebuilt rows - has <= N workgroups)
    n_err_part = (int)rows_alloc;
    HIP_TRY(hipMalloc((void **)&err_part, sizeof(double) * n_err_part));
    HIP_TRY(hipMalloc((void **)&ord_scr, sizeof(double) * (2 * ((ldv + kBlock - 1) / kBlock) + 2)));
    HIP_TRY(hipMalloc((void **)&st, sizeof(State<R>)));
    HIP_TRY(hipHostMalloc((void **)&st_host, sizeof(State<R>), hipHostMallocDefault));
    std::memset(&shadow, 0, sizeof(shadow));
    shadow.atl = atl;
    shadow.rtl = rtl;
    shadow.fixed = fixed;
    shadow.times_cap = times_cap;
    HIP_TRY(hipMemcpyAsync(st, &shadow, sizeof(shadow), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return KMT_OK;
  }

  int layout(long *plda, int *eb, int *rows, int *flags) override {
    if (plda) *plda = sym() ? kSymT : lda;
    if (eb) *eb = storage < 0 ? (int)sizeof(R) : (int)elem_bytes(storage);
    if (rows) *rows = tune_rows;
    if (flags) *flags = (tune_nt ? 1 : 0) | (tune_bal ? 2 : 0) | (tune_ring ? 4 : 0);
    return KMT_OK;
  }
  // The stage-7 error partials are summed per workgroup, so their grouping,
  // and with it the last bits of the error norm and of h, follows the
  // variant.  kmt_begin therefore picks it by a fixed rule, not by timing,
  // so that runs and restarts of one problem give the same bits in every
  // context and process: from N = 4096 on, a matrix that fits the 256 MB
  // Infinity Cache twice over takes 4 rows per group with plain loads, a
  // larger one 8 rows with non-temporal loads (the autotuner's picks at
  // N=5000 FP32 and N=50000 FP32/FP64); smaller N keep 8 rows, the grouping
  // of the FP64 register-rebuild kernel.  A larger 16-bit matrix takes 16
  // rows: each (cos, sin) chunk is reused for every row of the group from
  // registers, and at 8 rows the vectors' L2 bytes would match the 2-byte
  // entries' (FP64 state, BF16, N=50000: 0.754 against 0.811 ms per stage).
  // A matrix in the working precision takes the chunk ring (the same sums
  // and grouping; 0.6% faster at N=50000 FP64 and FP32 from HBM, 4.4% at
  // N=5000 FP32 from the Infinity Cache; slower with 16-bit entries at 16
  // rows, where it needs 231 registers).
  // kmt_autotune / kmt_select_variant override it.
  void default_variant() {
    const bool cached = N >= 4096 && (double)N * (double)N * (double)elem_bytes(storage) <= 128e6;
    tune_rows = cached ? 4 : (elem_bytes(storage) == 2 ? 16 : 8);
    tune_nt = cached ? 0 : 1;
    tune_bal = 0;
    tune_ring = elem_bytes(storage) != 2;
    bal_grid = 0;
    tuned = 1;
    drop_graphs();
  }

  int select_variant(int rows, int flags) override {
    if (rows != 4 && rows != 8 && rows != 16) return fail(KMT_ERR_ARG, "rows per group must be 4, 8 or 16 (got %d)", rows);
    if (flags & ~7) return fail(KMT_ERR_ARG, "unknown variant flags 0x%x", flags);
    if ((flags & 6) == 6) return fail(KMT_ERR_ARG, "the chunk ring has one workgroup per row group (not balanced)");
    tune_rows = rows;
    tune_nt = flags & 1;
    tune_bal = (flags >> 1) & 1;
    tune_ring = (flags >> 2) & 1;
    bal_grid = 0;
    tuned = 1;
    drop_graphs();
    return KMT_OK;
  }
  int tune(int reps, double *ms12) override {
    KMT_TRY(autotune(reps));
    if (ms12) for (int q = 0; q < 12; ++q) ms12[q] = tune_ms[q];
    return KMT_OK;
  }

  int set_omega(const void *om) override {
    HIP_TRY(hipMemcpyAsync(omega, om, sizeof(R) * N, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return KMT_OK;
  }

  int set_state(const void *th, double t, double h) override {
    HIP_TRY(hipMemcpyAsync(y, th, sizeof(R) * N, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    t0_ = (R)t;
    h0_ = (R)h;
    begun = 0;
    return KMT_OK;
  }

  int fetch_state() {
    HIP_TRY(hipMemcpyAsync(st_host, st, sizeof(State<R>), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    shadow = *st_host;
    return KMT_OK;
  }

  int push_state() {
    *st_host = shadow;
    HIP_TRY(hipMemcpyAsync(st, st_host, sizeof(State<R>), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return KMT_OK;
  }

  int get_state(void *th, double *t, double *h) override {
    if (!begun) {
      HIP_TRY(hipMemcpyAsync(th, y, sizeof(R) * N, hipMemcpyDeviceToHost, stream));
      HIP_TRY(hipStreamSynchronize(stream));
      if (t) *t = (double)t0_;
      if (h) *h = (double)h0_;
      return KMT_OK;
    }
    KMT_TRY(fetch_state());
    const R *src = shadow.commit_pending ? ynew : y;
    HIP_TRY(hipMemcpyAsync(th, src, sizeof(R) * N, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (t) *t = (double)shadow.t;
    if (h) *h = (double)shadow.h;
    return KMT_OK;
  }

  // ---------------- adjacency ----------------
  size_t elem_bytes(int stg) const {
    const int el = stg & ~KMT_ADJ_SYM;
    return (el == KMT_ADJ_BF16 || el == KMT_ADJ_F16) ? 2 : sizeof(R);
  }
  static bool is_sym(int stg) { return stg >= 0 && (stg & KMT_ADJ_SYM); }
  bool sym() const { return is_sym(storage); }

  // block layout of KMT_ADJ_SYM: the (I, J) table and the partial buffers
  int ensure_sym() {
    if (sym_ij) return KMT_OK;
    sym_nb = (N + kSymT - 1) / kSymT;
    sym_blocks = sym_nb * (sym_nb + 1) / 2;
    std::vector<int2> ij((size_t)sym_blocks);
    for (long I = 0; I < sym_nb; ++I)
      for (long J = I; J < sym_nb; ++J) ij[(size_t)sym_index(I, J, sym_nb)] = make_int2((int)I, (int)J);
    HIP_TRY(hipMalloc((void **)&sym_ij, sizeof(int2) * ij.size()));
    HIP_TRY(hipMemcpyAsync(sym_ij, ij.data(), sizeof(int2) * ij.size(), hipMemcpyHostToDevice, stream));
    // sized for the wider of the accumulation types (double for native
    // storage, float for 16-bit storage)
    HIP_TRY(hipMalloc((void **)&sym_prow, sizeof(SymPair<double>) * (size_t)sym_blocks * kSymT));
    HIP_TRY(hipMalloc((void **)&sym_pcol, sizeof(SymPair<double>) * (size_t)sym_blocks * kSymT));
    HIP_TRY(hipMemsetAsync(sym_prow, 0, sizeof(SymPair<double>) * (size_t)sym_blocks * kSymT, stream));
    HIP_TRY(hipMemsetAsync(sym_pcol, 0, sizeof(SymPair<double>) * (size_t)sym_blocks * kSymT, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return KMT_OK;
  }

  int ensure_A(int stg) {
    const int el = stg & ~KMT_ADJ_SYM;
    if (stg < 0 || (el != KMT_ADJ_NATIVE && el != KMT_ADJ_BF16 && el != KMT_ADJ_F16))
      return fail(KMT_ERR_ARG, "unknown adjacency storage %d", stg);
    if (is_sym(stg)) {
      KMT_TRY(ensure_sym());
    } else if (sym_ij) {  // back to row-major storage: drop the block table and partials
      drop_graphs();
      HIP_TRY(hipStreamSynchronize(stream));
      (void)hipFree(sym_ij);
      (void)hipFree(sym_prow);
      (void)hipFree(sym_pcol);
      sym_ij = nullptr;
      sym_prow = sym_pcol = nullptr;
    }
    const size_t bytes = is_sym(stg) ? (size_t)sym_blocks * kSymT * kSymT * elem_bytes(stg)
                                     : (size_t)rows_alloc * (size_t)lda * elem_bytes(stg);
    if (A && A_bytes == bytes && storage == stg) return KMT_OK;
    drop_graphs();
    if (A) {
      HIP_TRY(hipStreamSynchronize(stream));
      HIP_TRY(hipFree(A));
      A = nullptr;
    }
    hipError_t e = hipMalloc(&A, bytes);
    if (e != hipSuccess) {
      A = nullptr;
      return fail(KMT_ERR_MEMORY, "cannot allocate %zu bytes for the adjacency: %s", bytes, hipGetErrorString(e));
    }
    A_bytes = bytes;
    storage = stg;
    tuned = 0;
    bal_grid = 0;
    HIP_TRY(hipMemsetAsync(A, 0, bytes, stream));
    return KMT_OK;
  }

  static double pow2_at_least(double v) {
    if (!(v > 0) || !std::isfinite(v)) return 1.0;
    return std::ldexp(1.0, (int)std::ceil(std::log2(v)));
  }

  template <class S> int upload_rows(const R *M, long ld, R inv_scale) {
    // stage rows through a device buffer, convert into the stored layout
    const long chunk = std::max<long>(1, std::min<long>(N, (64l << 20) / (long)(sizeof(R) * N)));
    R *stage = nullptr;
    HIP_TRY(hipMalloc((void **)&stage, sizeof(R) * chunk * N));
    for (long r0 = 0; r0 < N; r0 += chunk) {
      const long nr = std::min(chunk, N - r0);
      hipError_t e = hipMemcpy2DAsync(stage, sizeof(R) * N, M + (size_t)r0 * ld, sizeof(R) * ld, sizeof(R) * N, nr,
                                      hipMemcpyHostToDevice, stream);
      if (e == hipSuccess) {
        dim3 grid((unsigned)std::min<long>((N + kBlock - 1) / kBlock, 64), (unsigned)nr);
        convert_rows_kernel<R, S><<<grid, kBlock, 0, stream>>>(stage, N, N, nr, (S *)A + (size_t)r0 * lda, lda, inv_scale);
        e = hipGetLastError();
      }
      if (e == hipSuccess) e = hipStreamSynchronize(stream);
      if (e != hipSuccess) {
        (void)hipFree(stage);
        return fail(KMT_ERR_HIP, "adjacency upload failed: %s", hipGetErrorString(e));
      }
    }
    HIP_TRY(hipFree(stage));
    return KMT_OK;
  }

  int set_adjacency(const void *Mv, int src_prec, long ld, int stg) override {
    if (ld < N) return fail(KMT_ERR_ARG, "adjacency row stride %ld < N=%ld", ld, N);
    HIP_TRY(hipStreamSynchronize(stream));
    drop_graphs();  // captured launches hold the old matrix scale
    if (src_prec != (sizeof(R) == 4 ? KMT_F32 : KMT_F64))
      return fail(KMT_ERR_ARG, "adjacency source precision must match the context precision");
    const R *M = (const R *)Mv;
    if (is_sym(stg)) {
      for (long i = 0; i < N; ++i)
        for (long j = i + 1; j < N; ++j)
          if (!(M[(size_t)i * ld + j] == M[(size_t)j * ld + i]) &&
              !(std::isnan((double)M[(size_t)i * ld + j]) && std::isnan((double)M[(size_t)j * ld + i])))
            return fail(KMT_ERR_ARG, "KMT_ADJ_SYM needs a symmetric matrix: M[%ld][%ld] != M[%ld][%ld]", i, j, j, i);
      KMT_TRY(ensure_A(stg));
      factored = 0;
      mscale_ = 1.0;
      if ((stg & ~KMT_ADJ_SYM) == KMT_ADJ_NATIVE) return upload_sym<R>(M, ld, (R)1);
      double mx = 0;
      for (long i = 0; i < N; ++i)
        for (long j = 0; j < N; ++j) mx = std::max(mx, std::fabs((double)M[(size_t)i * ld + j]));
      mscale_ = pow2_at_least(mx);
      const R inv = (R)(1.0 / mscale_);
      if ((stg & ~KMT_ADJ_SYM) == KMT_ADJ_BF16) return upload_sym<bf16_t>(M, ld, inv);
      return upload_sym<f16_t>(M, ld, inv);
    }
    KMT_TRY(ensure_A(stg));
    if (stg == KMT_ADJ_NATIVE) {
      mscale_ = 1.0;
      HIP_TRY(hipMemcpy2DAsync(A, sizeof(R) * lda, M, sizeof(R) * ld, sizeof(R) * N, N, hipMemcpyHostToDevice, stream));
      HIP_TRY(hipStreamSynchronize(stream));
      factored = 0;
      return KMT_OK;
    }
    double mx = 0;
    for (long i = 0; i < N; ++i)
      for (long j = 0; j < N; ++j) mx = std::max(mx, std::fabs((double)M[(size_t)i * ld + j]));
    mscale_ = pow2_at_least(mx);
    factored = 0;
    const R inv = (R)(1.0 / mscale_);
    if (stg == KMT_ADJ_BF16) return upload_rows<bf16_t>(M, ld, inv);
    return upload_rows<f16_t>(M, ld, inv);
  }

  // pack the upper block triangle of a host matrix, a few blocks per copy,
  // through a device staging buffer converted into the stored type
  template <class S> int upload_sym(const R *M, long ld, R inv_scale) {
    const long per = 32;
    std::vector<R> buf((size_t)per * kSymT * kSymT);
    R *stage = nullptr;
    HIP_TRY(hipMalloc((void **)&stage, sizeof(R) * buf.size()));
    for (long b0 = 0; b0 < sym_blocks; b0 += per) {
      const long nbk = std::min(per, sym_blocks - b0);
      std::fill(buf.begin(), buf.end(), (R)0);
      long b = 0;
      for (long I = 0; I < sym_nb && b < b0 + nbk; ++I)
        for (long J = I; J < sym_nb; ++J, ++b) {
          if (b < b0) continue;
          if (b >= b0 + nbk) break;
          R *dst = buf.data() + (size_t)(b - b0) * kSymT * kSymT;
          for (long r = 0; r < kSymT && I * kSymT + r < N; ++r)
            for (long c = 0; c < kSymT && J * kSymT + c < N; ++c)
              dst[(size_t)r * kSymT + c] = M[(size_t)(I * kSymT + r) * ld + J * kSymT + c];
        }
      hipError_t e = hipMemcpyAsync(stage, buf.data(), sizeof(R) * (size_t)nbk * kSymT * kSymT, hipMemcpyHostToDevice,
                                    stream);
      if (e == hipSuccess) {
        dim3 grid(1, (unsigned)(nbk * kSymT));
        convert_rows_kernel<R, S><<<grid, kBlock, 0, stream>>>(stage, kSymT, kSymT, nbk * kSymT,
                                                               (S *)A + (size_t)b0 * kSymT * kSymT, kSymT, inv_scale);
        e = hipGetLastError();
      }
      if (e == hipSuccess) e = hipStreamSynchronize(stream);  // buf is refilled next
      if (e != hipSuccess) {
        (void)hipFree(stage);
        return fail(KMT_ERR_HIP, "adjacency upload failed: %s", hipGetErrorString(e));
      }
    }
    HIP_TRY(hipFree(stage));
    return KMT_OK;
  }

  int set_factor(const void *Xv, const void *wv, int k_) override {
    if (k_ < 0) return fail(KMT_ERR_ARG, "rank K must be >= 0");
    HIP_TRY(hipStreamSynchronize(stream));  // the buffers below may be in use by queued work
    drop_graphs();
    if (X) { (void)hipFree(X); X = nullptr; }
    if (w) { (void)hipFree(w); w = nullptr; }
    if (XT) { (void)hipFree(XT); XT = nullptr; }
    if (lr_part) { (void)hipFree(lr_part); lr_part = nullptr; }
    if (lr_G) { (void)hipFree(lr_G); lr_G = nullptr; }
    drop_sign();
    K = k_;
    const long kk = std::max(K, 1);
    X_host.assign((const R *)Xv, (const R *)Xv + (size_t)N * K);
    w_host.assign((const R *)wv, (const R *)wv + K);
    HIP_TRY(hipMalloc((void **)&X, sizeof(R) * std::max<size_t>((size_t)N * K, 1)));
    HIP_TRY(hipMalloc((void **)&w, sizeof(R) * kk));
    if (K > 0) {
      HIP_TRY(hipMemcpyAsync(X, Xv, sizeof(R) * N * K, hipMemcpyHostToDevice, stream));
      HIP_TRY(hipMemcpyAsync(w, wv, sizeof(R) * K, hipMemcpyHostToDevice, stream));
    }
    lr_nblk = (int)std::min<long>((N + kBlock - 1) / kBlock, 512);
    lr_grid = (int)std::min<long>((N + kBlock - 1) / kBlock, kLrMaxBlocks);
    const int nparts = std::max(lr_nblk, 2 * lr_grid);
    HIP_TRY(hipMalloc((void **)&lr_part, sizeof(double) * 2 * kk * nparts));
    HIP_TRY(hipMemsetAsync(lr_part, 0, sizeof(double) * 2 * kk * nparts, stream));
    HIP_TRY(hipMalloc((void **)&lr_G, sizeof(double) * 2 * kk));
    HIP_TRY(hipMemsetAsync(lr_G, 0, sizeof(double) * 2 * kk, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (factored == KMT_LOWRANK_REBUILD) {  // keep the mode: transpose the new factor
      factored = KMT_LOWRANK_CONTRACT;
      return use_factored(KMT_LOWRANK_REBUILD);
    }
    return KMT_OK;
  }

  void drop_sign() {
    void *b[] = {s_pass[0].tiles, s_pass[0].bfrag, s_pass[0].part, s_pass[0].ones,
                 s_pass[1].tiles, s_pass[1].bfrag, s_pass[1].part, s_pass[1].ones};
    for (void *p : b)
      if (p) (void)hipFree(p);
    s_pass[0] = s_pass[1] = SignPass();
    sign = 0;
  }

  // 16-row tiles per wave of the int8 passes
  static constexpr int kSignNT = 4;
  // Shape of a pass with `rows` output rows (+ the ones row) over kdim:
  // workgroups of 4 row groups of 16 kSignNT rows x splits of the 1024-wide
  // super-block pairs, at least 2 pairs per split.  The split count keeps
  // the waves within two rounds of the `resident` waves the device holds at
  // once and, among those counts, fills the last round best: a wave count
  // just past a whole round leaves most of the GPU idle behind a few waves.
  // N = K = 10^5 on 256 CUs at 3 workgroups per CU (3072 resident waves):
  // 3 splits = 4692 waves = 1.53 rounds, 0.514-0.532 ms per RHS; 2 splits
  // (1.02 rounds) 0.577, 4 (2.04) 0.551, 5 (2.55) 0.535, 6 (3.05) 0.563,
  // 7 (3.56) 0.551; fewer splits also write and re-read fewer partials.
  static constexpr long kSignRowsPerBlock = 16 * kSignNT * (kBlock / kWave);
  static SignPass sign_pass_shape(long rows, long kdim, long resident) {
    SignPass p;
    p.ldp = (rows + 1 + kSignRowsPerBlock - 1) / kSignRowsPerBlock * kSignRowsPerBlock;
    const long rg = p.ldp / (16 * kSignNT);
    p.npairs = std::max<long>(1, (kdim + 2 * kSignSB - 1) / (2 * kSignSB));
    p.nsb = 2 * p.npairs;
    const long spmax = std::min<long>(std::max<long>(1, 2 * resident / rg), std::max<long>(1, p.npairs / 2));
    long sp = 1;
    double best = -1.0;
    for (long c = 1; c <= spmax; ++c) {
      const long w = rg * c, rounds = (w + resident - 1) / resident;
      const double fill = (double)w / (double)(rounds * resident);
      if (fill > best) best = fill, sp = c;
    }
    p.pairs_per = (p.npairs + sp - 1) / sp;
    p.nsplit = (int)((p.npairs + p.pairs_per - 1) / p.pairs_per);
    return p;
  }
  int alloc_sign_pass(SignPass &p, long kdim) {
    HIP_TRY(hipMalloc((void **)&p.bfrag, (size_t)p.nsb * 8 * kWave * 16));
    HIP_TRY(hipMalloc((void **)&p.part, sizeof(int32_t) * 16 * (size_t)p.nsplit * p.ldp));
    std::vector<uint32_t> ones((size_t)16 * p.nsb, 0u);
    for (long q = 0; q < (long)ones.size(); ++q)
      ones[q] = 32 * (q + 1) <= kdim ? 0xFFFFFFFFu : (32 * q < kdim ? (1u << (kdim - 32 * q)) - 1u : 0u);
    HIP_TRY(hipMalloc((void **)&p.ones, sizeof(uint32_t) * ones.size()));
    HIP_TRY(hipMemcpy(p.ones, ones.data(), sizeof(uint32_t) * ones.size(), hipMemcpyHostToDevice));
    return KMT_OK;
  }

  // The volcano factor X_nm = +-1 of the library's sign stream, packed, and
  // w_m = (-1)^m / N (makeu1 / makeu2 / m1 / m2, kuramoto.cu:62-143).
  int set_factor_sign(int k_, uint32_t seed) override {
    if (k_ < 0) return fail(KMT_ERR_ARG, "rank K must be >= 0");
    if (N > kSignMaxDim || k_ > kSignMaxDim)
      return fail(KMT_ERR_ARG, "the packed sign factor needs N, K <= %ld (N=%ld, K=%d)", kSignMaxDim, N, k_);
    HIP_TRY(hipStreamSynchronize(stream));
    drop_graphs();
    if (X) { (void)hipFree(X); X = nullptr; }
    if (w) { (void)hipFree(w); w = nullptr; }
    if (XT) { (void)hipFree(XT); XT = nullptr; }
    if (lr_part) { (void)hipFree(lr_part); lr_part = nullptr; }
    if (lr_G) { (void)hipFree(lr_G); lr_G = nullptr; }
    drop_sign();
    if (factored == KMT_LOWRANK_REBUILD) factored = KMT_LOWRANK_CONTRACT;
    K = k_;
    X_host.clear();
    w_host.assign((size_t)K, (R)0);
    for (int m = 0; m < K; ++m) w_host[m] = (R)(((m & 1) ? -1.0 : 1.0) / (double)N);
    HIP_TRY(hipMalloc((void **)&w, sizeof(R) * std::max(K, 1)));
    if (K > 0) HIP_TRY(hipMemcpyAsync(w, w_host.data(), sizeof(R) * K, hipMemcpyHostToDevice, stream));
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sign_mfma_kernel<R, kSignNT>, kBlock, 0) != hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1) cus = 1;
    const long resident = (long)per_cu * cus * (kBlock / kWave);
    s_pass[0] = sign_pass_shape(K, N, resident);
    s_pass[1] = sign_pass_shape(N, K, resident);
    SignPass &p1 = s_pass[0], &p2 = s_pass[1];
    const size_t tw1 = (size_t)p1.ldp * 32 * p1.npairs, tw2 = (size_t)p2.ldp * 32 * p2.npairs;
    hipError_t e = hipMalloc((void **)&p1.tiles, sizeof(uint32_t) * tw1);
    if (e == hipSuccess) e = hipMalloc((void **)&p2.tiles, sizeof(uint32_t) * tw2);
    if (e != hipSuccess) {
      drop_sign();
      return fail(KMT_ERR_MEMORY, "cannot allocate the packed sign factor (N=%ld, K=%d): %s", N, K, hipGetErrorString(e));
    }
    if (alloc_sign_pass(p1, N) != KMT_OK || alloc_sign_pass(p2, K) != KMT_OK) {
      drop_sign();
      return fail(KMT_ERR_MEMORY, "cannot allocate the sign passes' buffers (N=%ld, K=%d)", N, K);
    }
    // zero rows (and words past each row's k range) stay as cleared here
    HIP_TRY(hipMemsetAsync(p1.tiles, 0, sizeof(uint32_t) * tw1, stream));
    HIP_TRY(hipMemsetAsync(p2.tiles, 0, sizeof(uint32_t) * tw2, stream));
    const size_t words = (size_t)((N + 63) / 64) * p2.npairs * 2048;
    gen_sign_bits_kernel<<<(unsigned)((words + kBlock - 1) / kBlock), kBlock, 0, stream>>>(p2.tiles, N, K, p2.npairs,
                                                                                          seed);
    HIP_TRY(hipGetLastError());
    if (K > 0) {
      const long kw = (K + 31) / 32;
      const size_t waves = (size_t)((N + 63) / 64) * kw;
      transpose_sign_bits_kernel<<<(unsigned)((waves * kWave + kBlock - 1) / kBlock), kBlock, 0, stream>>>(
          p2.tiles, N, p2.npairs, kw, p1.tiles, p1.npairs);
      HIP_TRY(hipGetLastError());
    }
    // the ones rows last: the transpose writes zero words into rows m >= K
    sign_tile_ones_kernel<<<(unsigned)((32 * p1.npairs + kBlock - 1) / kBlock), kBlock, 0, stream>>>(
        p1.tiles, K, p1.ones, p1.npairs);
    sign_tile_ones_kernel<<<(unsigned)((32 * p2.npairs + kBlock - 1) / kBlock), kBlock, 0, stream>>>(
        p2.tiles, N, p2.ones, p2.npairs);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(stream));
    sign = 1;
    return KMT_OK;
  }

  template <class S> int gen_sign_rows(R inv_scale) {
    constexpr int ROWS = 8;
    const long chunk = 4096;
    for (long r0 = 0; r0 < N; r0 += chunk) {
      const long nr = std::min(chunk, N - r0);
      dim3 grid((unsigned)((N + kBlock - 1) / kBlock), (unsigned)((nr + ROWS - 1) / ROWS));
      gen_adj_sign_kernel<R, S, ROWS><<<grid, kBlock, 0, stream>>>(s_pass[1].tiles, s_pass[1].npairs, w, K, N, r0, nr,
                                                                   (S *)A + (size_t)r0 * lda, lda, inv_scale);
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(stream));
    return KMT_OK;
  }

  template <class S> int gen_rows(R inv_scale) {
    const long chunk = 4096;
    for (long r0 = 0; r0 < N; r0 += chunk) {
      const long nr = std::min(chunk, N - r0);
      gen_adj_kernel<R, S><<<(unsigned)nr, kBlock, 0, stream>>>(X, w, K, N, r0, (S *)A + (size_t)r0 * lda, lda, inv_scale);
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(stream));
    return KMT_OK;
  }

  int build_adjacency(int stg) override {
    if (!X && !sign) return fail(KMT_ERR_STATE, "kmt_build_adjacency needs kmt_set_factor or kmt_set_factor_sign first");
    if (sign && is_sym(stg)) return fail(KMT_ERR_ARG, "symmetric storage is not built from a packed sign factor");
    HIP_TRY(hipStreamSynchronize(stream));
    drop_graphs();  // captured launches hold the old matrix scale
    KMT_TRY(ensure_A(stg));
    factored = 0;
    if (sign) {
      // |M_ij| <= sum_m |w_m| (X^2 = 1): scale the 16-bit forms so |stored| <= 1
      double mx = 0;
      for (int m = 0; m < K; ++m) mx += std::fabs((double)w_host[m]);
      mscale_ = stg == KMT_ADJ_NATIVE ? 1.0 : pow2_at_least(mx);
      const R inv = (R)(1.0 / mscale_);
      if (stg == KMT_ADJ_NATIVE) return gen_sign_rows<R>((R)1);
      if (stg == KMT_ADJ_BF16) return gen_sign_rows<bf16_t>(inv);
      return gen_sign_rows<f16_t>(inv);
    }
    if (stg == KMT_ADJ_NATIVE) {
      mscale_ = 1.0;
      return gen_rows<R>((R)1);
    }
    if (stg == KMT_ADJ_SYM) {
      mscale_ = 1.0;
      gen_adj_sym_kernel<R, R><<<(unsigned)sym_blocks, kBlock, 0, stream>>>(X, w, K, N, sym_ij, (R *)A, (R)1);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipStreamSynchronize(stream));
      return KMT_OK;
    }
    // |M_ij| <= max_n sum_m |w_m| X_nm^2 (Cauchy–Schwarz): scale so |stored| <= 1
    double mx = 0;
    for (long n = 0; n < N; ++n) {
      double s = 0;
      for (int m = 0; m < K; ++m) s += std::fabs((double)w_host[m]) * (double)X_host[n * K + m] * (double)X_host[n * K + m];
      mx = std::max(mx, s);
    }
    mscale_ = pow2_at_least(mx);
    const R inv = (R)(1.0 / mscale_);
    if (is_sym(stg)) {
      if ((stg & ~KMT_ADJ_SYM) == KMT_ADJ_BF16)
        gen_adj_sym_kernel<R, bf16_t><<<(unsigned)sym_blocks, kBlock, 0, stream>>>(X, w, K, N, sym_ij, (bf16_t *)A, inv);
      else
        gen_adj_sym_kernel<R, f16_t><<<(unsigned)sym_blocks, kBlock, 0, stream>>>(X, w, K, N, sym_ij, (f16_t *)A, inv);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipStreamSynchronize(stream));
      return KMT_OK;
    }
    if (stg == KMT_ADJ_BF16) return gen_rows<bf16_t>(inv);
    return gen_rows<f16_t>(inv);
  }

  int use_factored(int on) override {
    if (on < 0 || on > KMT_LOWRANK_REBUILD) return fail(KMT_ERR_ARG, "unknown low-rank mode %d", on);
    if (on && !X && !sign) return fail(KMT_ERR_STATE, "-A needs kmt_set_factor or kmt_set_factor_sign first");
    if (on == KMT_LOWRANK_REBUILD && sign)
      return fail(KMT_ERR_ARG, "register-rebuilt tiles need a real factor (kmt_set_factor), not a packed sign factor");
    if (on == KMT_LOWRANK_REBUILD) {
      if (K > kLrMaxRank) return fail(KMT_ERR_ARG, "register-rebuilt tiles support K <= %d (K=%d)", kLrMaxRank, K);
      if (!XT) {
        std::vector<R> t((size_t)std::max(K, 1) * ldv, (R)0);
        for (long i = 0; i < N; ++i)
          for (int m = 0; m < K; ++m) t[(size_t)m * ldv + i] = X_host[(size_t)i * K + m];
        HIP_TRY(hipMalloc((void **)&XT, sizeof(R) * t.size()));
        HIP_TRY(hipMemcpyAsync(XT, t.data(), sizeof(R) * t.size(), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipStreamSynchronize(stream));
      }
    }
    if (factored != on) drop_graphs();
    factored = on;
    return KMT_OK;
  }

  int get_adjacency_rows(long i0, long i1, void *out) override {
    if (!A) return fail(KMT_ERR_STATE, "no dense adjacency on the device");
    if (i0 < 0 || i1 > N || i0 > i1) return fail(KMT_ERR_ARG, "row range [%ld, %ld) outside [0, %ld)", i0, i1, N);
    const long chunk = std::max<long>(1, std::min<long>(i1 - i0, (64l << 20) / (long)(sizeof(R) * N)));
    R *stage = nullptr;
    HIP_TRY(hipMalloc((void **)&stage, sizeof(R) * chunk * N));
    for (long r0 = i0; r0 < i1; r0 += chunk) {
      const long nr = std::min(chunk, i1 - r0);
      dim3 grid((unsigned)std::min<long>((N + kBlock - 1) / kBlock, 64), (unsigned)nr);
      hipError_t e = hipSuccess;
      if (storage == KMT_ADJ_SYM)
        read_sym_rows_kernel<R, R><<<grid, kBlock, 0, stream>>>((const R *)A, sym_nb, N, r0, nr, (R)1, stage);
      else if (storage == (KMT_ADJ_SYM | KMT_ADJ_BF16))
        read_sym_rows_kernel<R, bf16_t><<<grid, kBlock, 0, stream>>>((const bf16_t *)A, sym_nb, N, r0, nr, (R)mscale_, stage);
      else if (storage == (KMT_ADJ_SYM | KMT_ADJ_F16))
        read_sym_rows_kernel<R, f16_t><<<grid, kBlock, 0, stream>>>((const f16_t *)A, sym_nb, N, r0, nr, (R)mscale_, stage);
      else if (storage == KMT_ADJ_NATIVE)
        read_rows_kernel<R, R><<<grid, kBlock, 0, stream>>>((const R *)A + (size_t)r0 * lda, lda, N, nr, (R)1, stage);
      else if (storage == KMT_ADJ_BF16)
        read_rows_kernel<R, bf16_t><<<grid, kBlock, 0, stream>>>((const bf16_t *)A + (size_t)r0 * lda, lda, N, nr, (R)mscale_, stage);
      else
        read_rows_kernel<R, f16_t><<<grid, kBlock, 0, stream>>>((const f16_t *)A + (size_t)r0 * lda, lda, N, nr, (R)mscale_, stage);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipMemcpyAsync((R *)out + (size_t)(r0 - i0) * N, stage, sizeof(R) * nr * N, hipMemcpyDeviceToHost, stream);
      if (e == hipSuccess) e = hipStreamSynchronize(stream);
      if (e != hipSuccess) {
        (void)hipFree(stage);
        return fail(KMT_ERR_HIP, "adjacency read failed: %s", hipGetErrorString(e));
      }
    }
    HIP_TRY(hipFree(stage));
    return KMT_OK;
  }

  // ---------------- launches ----------------
  Vecs<R> vecs(const R *cs_in, R *cs_out, int check) {
    Vecs<R> v;
    v.cs_in = cs_in;
    v.cs_out = cs_out;
    v.ldv = ldv;
    v.N = N;
    v.omega = omega;
    v.y = y;
    for (int j = 0; j < 7; ++j) v.k[j] = k[j];
    v.ynew = ynew;
    v.f_out = f_scr;
    v.cs_pairs = nullptr;
    v.c1 = c1;
    v.st = st;
    v.err_part = err_part;
    v.check_state = check;
    return v;
  }

  int coupling_ready() const { return factored ? (X != nullptr || sign) : (A != nullptr); }

  // workgroups of one resident round of the balanced variant (<= N rows)
  template <class S, int ROWS, bool NT> int balanced_grid() {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, stage_kernel<R, S, ROWS, 3, NT, true>, kBlock, 0) !=
            hipSuccess || per_cu < 1)
      per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1) cus = 1;
    return (int)std::min<long>(std::min<long>((long)per_cu * cus, N), n_err_part);
  }

  template <class S, int KIND, int ROWS, bool NT> void launch_dense_variant(const DenseArgs<R, S> &a) {
    if (tune_bal) {
      if (bal_grid <= 0) bal_grid = balanced_grid<S, ROWS, NT>();
      stage_kernel<R, S, ROWS, KIND, NT, true><<<(unsigned)bal_grid, kBlock, 0, stream>>>(a);
    } else if (tune_ring) {
      const unsigned grid = (unsigned)((N + ROWS - 1) / ROWS);
      stage_ring_kernel<R, S, ROWS, KIND, NT><<<grid, kBlock, 0, stream>>>(a);
    } else {
      const unsigned grid = (unsigned)((N + ROWS - 1) / ROWS);
      stage_kernel<R, S, ROWS, KIND, NT, false><<<grid, kBlock, 0, stream>>>(a);
    }
  }

  template <class S, int KIND> void launch_dense_kind(const Vecs<R> &v) {
    DenseArgs<R, S> a;
    a.A = (const S *)A;
    a.lda = lda;
    a.mscale = (R)mscale_;
    a.v = v;
    switch (tune_rows * 2 + (tune_nt ? 1 : 0)) {
      case 8: launch_dense_variant<S, KIND, 4, false>(a); break;
      case 9: launch_dense_variant<S, KIND, 4, true>(a); break;
      case 16: launch_dense_variant<S, KIND, 8, false>(a); break;
      case 32: launch_dense_variant<S, KIND, 16, false>(a); break;
      case 33: launch_dense_variant<S, KIND, 16, true>(a); break;
      default: launch_dense_variant<S, KIND, 8, true>(a); break;
    }
  }

  // KMT_ADJ_SYM: block kernel over the stored triangle, then the row reduction
  template <class S, int KIND> void launch_sym_as(const Vecs<R> &v) {
    using Acc = typename Store<S>::Acc;
    SymArgs<R, S> a;
    a.A = (const S *)A;
    a.ij = sym_ij;
    a.nb = sym_nb;
    a.prow = reinterpret_cast<SymPair<Acc> *>(sym_prow);
    a.pcol = reinterpret_cast<SymPair<Acc> *>(sym_pcol);
    a.mscale = (R)mscale_;
    a.v = v;
    sym_tile_kernel<R, S><<<(unsigned)sym_blocks, kBlock, 0, stream>>>(a);
    sym_reduce_kernel<R, S, KIND><<<(unsigned)sym_nb, kBlock, 0, stream>>>(a);
  }
  template <int KIND> void launch_sym(const Vecs<R> &v) {
    if (storage == KMT_ADJ_SYM) launch_sym_as<R, KIND>(v);
    else if (storage == (KMT_ADJ_SYM | KMT_ADJ_BF16)) launch_sym_as<bf16_t, KIND>(v);
    else launch_sym_as<f16_t, KIND>(v);
  }

  template <int KIND> void launch_dense(const Vecs<R> &v) {
    if (sym()) launch_sym<KIND>(v);
    else if (storage == KMT_ADJ_NATIVE) launch_dense_kind<R, KIND>(v);
    else if (storage == KMT_ADJ_BF16) launch_dense_kind<bf16_t, KIND>(v);
    else launch_dense_kind<f16_t, KIND>(v);
  }

  bool lr_fused() const { return K <= kLrMaxRank && !sign; }
  // partial-sum buffers of the fused low-rank path (used alternately)
  double *lr_p(int which) const { return lr_part + (size_t)which * 2 * std::max(K, 1) * lr_grid; }

  template <int KT> void launch_gram_t(const R *cs_in, int which, int check) {
    lowrank_gram_kernel<R, KT><<<(unsigned)lr_grid, kBlock, 0, stream>>>(X, K, N, cs_in, ldv, lr_p(which), st, check);
  }
  void launch_gram(const R *cs_in, int which, int check) {
    if (K <= 8) launch_gram_t<8>(cs_in, which, check);
    else launch_gram_t<16>(cs_in, which, check);
  }
  // stage KIND from the partials in lr_p(pin); with NEXT, the partials of
  // its output vector into lr_p(pin ^ 1)
  template <int KIND, int KT, bool NEXT> void launch_lr_stage_t(const Vecs<R> &v, int pin) {
    lowrank_stage_kernel<R, KIND, KT, NEXT><<<(unsigned)lr_grid, kBlock, 0, stream>>>(X, K, lr_p(pin), lr_grid, w, v,
                                                                                       lr_p(pin ^ 1));
  }
  template <int KIND, bool NEXT> void launch_lr_stage(const Vecs<R> &v, int pin) {
    if (K <= 8) launch_lr_stage_t<KIND, 8, NEXT>(v, pin);
    else launch_lr_stage_t<KIND, 16, NEXT>(v, pin);
  }
  // one attempt's six stages on the fused low-rank path: partials of the
  // stage-2 vector, then one launch per stage that also writes the partials
  // of the next stage's vector
  void launch_lr_attempt(std::vector<std::pair<hipEvent_t, hipEvent_t>> *ev) {
    for (int s = 2; s <= 7; ++s) {
      const R *in = cs[(s - 2) & 1];
      R *out = cs[(s - 1) & 1];
      Vecs<R> v = vecs(in, out, 1);
      const int pin = (s - 2) & 1;
      if (ev) (void)hipEventRecord((*ev)[0].first, stream);
      if (s == 2) launch_gram(in, 0, 1);
      switch (s) {
        case 2: launch_lr_stage<2, true>(v, pin); break;
        case 3: launch_lr_stage<3, true>(v, pin); break;
        case 4: launch_lr_stage<4, true>(v, pin); break;
        case 5: launch_lr_stage<5, true>(v, pin); break;
        case 6: launch_lr_stage<6, true>(v, pin); break;
        default: launch_lr_stage<7, false>(v, pin); break;
      }
      if (ev) {
        (void)hipEventRecord((*ev)[0].second, stream);
        ev->erase(ev->begin());
      }
    }
  }

  // packed sign factor: the stage vectors as digits, pass 1 over bitsT
  // (rows m), G as digits, pass 2 over bits (rows n), sums + stage epilogue
  template <int KIND> void launch_sign(const Vecs<R> &v) {
    const int chk = v.check_state && KIND >= 2;
    const SignPass &p1 = s_pass[0], &p2 = s_pass[1];
    if (K > 0) {
      const long n_pad = p1.nsb * kSignSB, k_pad = p2.nsb * kSignSB;
      const long b1 = (p1.ldp / kSignRowsPerBlock) * p1.nsplit, b2 = (p2.ldp / kSignRowsPerBlock) * p2.nsplit;
      sign_limbs_kernel<R><<<(unsigned)((n_pad + kBlock - 1) / kBlock), kBlock, 0, stream>>>(
          v.cs_in, v.cs_in + ldv, N, n_pad, p1.bfrag, st, chk);
      sign_mfma_kernel<R, kSignNT><<<(unsigned)b1, kBlock, 0, stream>>>(
          (const u32x4 *)p1.tiles, (const u32x4 *)p1.bfrag, p1.npairs, p1.pairs_per, p1.nsplit,
          p1.part, p1.ldp, st, chk);
      sign_gram_kernel<R><<<(unsigned)((k_pad + kBlock - 1) / kBlock), kBlock, 0, stream>>>(
          p1.part, p1.nsplit, p1.ldp, w, K, k_pad, p2.bfrag, st, chk);
      sign_mfma_kernel<R, kSignNT><<<(unsigned)b2, kBlock, 0, stream>>>(
          (const u32x4 *)p2.tiles, (const u32x4 *)p2.bfrag, p2.npairs, p2.pairs_per, p2.nsplit, p2.part,
          p2.ldp, st, chk);
    }
    sign_stage_kernel<R, KIND><<<(unsigned)((N + kBlock - 1) / kBlock), kBlock, 0, stream>>>(
        p2.part, K > 0 ? p2.nsplit : 0, p2.ldp, v);
  }

  template <int KIND> void launch_factored(const Vecs<R> &v) {
    if (sign) {
      launch_sign<KIND>(v);
      return;
    }
    if (lr_fused()) {
      launch_gram(v.cs_in, 0, v.check_state && KIND >= 2);
      launch_lr_stage<KIND, false>(v, 0);
      return;
    }
    const int tiles = std::max(1, (K + kRankTile - 1) / kRankTile);
    if (K > 0) {
      lowrank_partial_kernel<R><<<dim3(lr_nblk, tiles), kBlock, 0, stream>>>(X, K, N, v.cs_in, ldv, lr_part, st, v.check_state && KIND >= 2);
      lowrank_reduce_kernel<R><<<1, kBlock, 0, stream>>>(lr_part, lr_nblk, K, w, lr_G, st, v.check_state && KIND >= 2);
    }
    const unsigned grid = (unsigned)((N + kBlock - 1) / kBlock);
    lowrank_expand_kernel<R, KIND><<<grid, kBlock, 0, stream>>>(X, K, lr_G, v);
  }

  int rebuild_rows() const { return K == 5 ? (sizeof(R) == 4 ? 16 : 8) : K <= 8 ? 8 : 4; }
  template <int KIND> void launch_rebuild(const Vecs<R> &v) {
    RebuildArgs<R> a;
    a.XT = XT;
    a.w = w;
    a.K = K;
    a.v = v;
    const int rows = rebuild_rows();
    const unsigned grid = (unsigned)((N + rows - 1) / rows);
    if (K == 5) rebuild_stage_kernel<R, sizeof(R) == 4 ? 16 : 8, KIND, 5, true><<<grid, kBlock, 0, stream>>>(a);
    else if (K <= 8) rebuild_stage_kernel<R, 8, KIND, 8, false><<<grid, kBlock, 0, stream>>>(a);
    else rebuild_stage_kernel<R, 4, KIND, 16, false><<<grid, kBlock, 0, stream>>>(a);
  }

  template <int KIND> void launch_kind(const Vecs<R> &v) {
    if (factored == KMT_LOWRANK_REBUILD) launch_rebuild<KIND>(v);
    else if (factored) launch_factored<KIND>(v);
    else launch_dense<KIND>(v);
  }

  void launch_by_kind(int kind, const Vecs<R> &v) {
    switch (kind) {
      case 0: launch_kind<0>(v); break;
      case 1: launch_kind<1>(v); break;
      case 2: launch_kind<2>(v); break;
      case 3: launch_kind<3>(v); break;
      case 4: launch_kind<4>(v); break;
      case 5: launch_kind<5>(v); break;
      case 6: launch_kind<6>(v); break;
      default: launch_kind<7>(v); break;
    }
  }

  int n_parts_stage7() const {
    if (factored == KMT_LOWRANK_REBUILD) return (int)((N + rebuild_rows() - 1) / rebuild_rows());
    if (factored) return lr_fused() ? lr_grid : (int)((N + kBlock - 1) / kBlock);
    if (sym()) return (int)sym_nb;
    if (tune_bal) return bal_grid;
    return (int)((N + tune_rows - 1) / tune_rows);
  }

  // Autotuner of the dense stage kernel: rows per workgroup {4, 8, 16} x
  // non-temporal adjacency loads {off, on} x grid {one workgroup per row
  // group (with and without the chunk ring), balanced: one resident round
  // with contiguous equal row ranges}, each timed on this context's own
  // matrix (stage 3 on scratch buffers).
  // Every variant sums each row in the same order (the RHS is bit-identical),
  // but the stage-7 error partials follow the variant's workgroups, so the
  // last bits of the error norm and of h can differ from the default rule's:
  // a context tuned here does not restart bit-exactly against one that was
  // not (or that picked another variant).
  int autotune(int reps) {
    if (factored || !A || sym()) return KMT_OK;  // one form of the symmetric kernels
    if (!k_scr) {
      HIP_TRY(hipMalloc((void **)&k_scr, sizeof(R) * 7 * ldv));
      HIP_TRY(hipMemsetAsync(k_scr, 0, sizeof(R) * 7 * ldv, stream));
    }
    Vecs<R> v = vecs(cs[0], cs_scr, 0);
    for (int j = 0; j < 7; ++j) v.k[j] = k_scr + j * ldv;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    const int rows_opts[3] = {4, 8, 16};
    double best = 1e30;
    int best_rows = 8, best_nt = 1, best_bal = 0, best_ring = 0, best_grid = 0;
    // one workgroup per group is timed with and without the chunk ring; its
    // table entry is the faster of the two
    for (int ri = 0; ri < 3; ++ri)
      for (int nt = 0; nt < 2; ++nt)
      for (int bal = 0; bal < 2; ++bal) {
        tune_ms[ri * 4 + nt * 2 + bal] = 1e30;
        for (int ring = 0; ring < (bal ? 1 : 2); ++ring) {
          tune_rows = rows_opts[ri];
          tune_nt = nt;
          tune_bal = bal;
          tune_ring = ring;
          bal_grid = 0;
          launch_by_kind(3, v);
          HIP_TRY(hipEventRecord(e0, stream));
          for (int r = 0; r < reps; ++r) launch_by_kind(3, v);
          HIP_TRY(hipEventRecord(e1, stream));
          HIP_TRY(hipEventSynchronize(e1));
          float ms = 0;
          HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
          tune_ms[ri * 4 + nt * 2 + bal] = std::min(tune_ms[ri * 4 + nt * 2 + bal], (double)ms / reps);
          if (ms < best) {
            best = ms;
            best_rows = tune_rows;
            best_nt = nt;
            best_bal = bal;
            best_ring = ring;
            best_grid = bal_grid;
          }
        }
      }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    tune_rows = best_rows;
    tune_nt = best_nt;
    tune_bal = best_bal;
    tune_ring = best_ring;
    bal_grid = best_grid;
    tuned = 1;
    drop_graphs();
    return KMT_OK;
  }

  EmitArgs<R> emit_args() {
    EmitArgs<R> e;
    e.st = st;
    e.N = N;
    e.ldv = ldv;
    e.y = y;
    e.ylast = ylast;
    e.ynew = ynew;
    for (int j = 0; j < 7; ++j) e.k[j] = k[j];
    e.cs_out = cs[0];
    e.t_eval = t_eval;
    e.ring = keep_rows ? ring : nullptr;
    e.ord_part = ord_part;
    return e;
  }

  // One attempt: stages 2..7 (cs ping-pong 0->1->0->1->0->1), control,
  // emit/commit (writes stage-2 vector into cs[0]), finalize.
  void launch_attempt(std::vector<std::pair<hipEvent_t, hipEvent_t>> *ev = nullptr) {
    if (factored == KMT_LOWRANK_CONTRACT && lr_fused()) launch_lr_attempt(ev);
    else for (int s = 2; s <= 7; ++s) {
      const R *in = cs[(s - 2) & 1];
      R *out = cs[(s - 1) & 1];
      Vecs<R> v = vecs(in, out, 1);
      if (ev) (void)hipEventRecord((*ev)[0].first, stream);
      launch_by_kind(s, v);
      if (ev) {
        (void)hipEventRecord((*ev)[0].second, stream);
        ev->erase(ev->begin());
      }
    }
    control_kernel<R><<<1, kBlock, 0, stream>>>(st, err_part, n_parts_stage7(), N, t_eval, times);
    emit_commit_kernel<R><<<(unsigned)nblk_emit, kBlock, 0, stream>>>(emit_args());
    finalize_kernel<R><<<1, kBlock, 0, stream>>>(st, ord_part, nblk_emit, N, r_ring);
  }

  int build_graphs() {
    drop_graphs();
    for (int which = 0; which < 2; ++which) {
      const int n = which == 0 ? 1 : 8;
      hipGraph_t g = nullptr;
      HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      for (int a = 0; a < n; ++a) launch_attempt();
      HIP_TRY(hipStreamEndCapture(stream, &g));
      hipGraphExec_t ex = nullptr;
      hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return fail(KMT_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
      if (which == 0) g1 = ex;
      else g8 = ex;
    }
    graphs_valid = true;
    return KMT_OK;
  }

  // ---------------- model evaluations ----------------
  int rhs(const void *th, void *f, void *cpl) override {
    if (!coupling_ready()) return fail(KMT_ERR_STATE, "no adjacency: call kmt_set_adjacency, kmt_build_adjacency or kmt_set_factor + kmt_use_factored");
    HIP_TRY(hipMemcpyAsync(vec_scr, th, sizeof(R) * N, hipMemcpyHostToDevice, stream));
    sincos_kernel<R><<<(unsigned)((N + kBlock - 1) / kBlock), kBlock, 0, stream>>>(vec_scr, N, ldv, cs_scr);
    Vecs<R> v = vecs(cs_scr, nullptr, 0);
    v.f_out = f_scr;
    v.cs_pairs = pairs;
    launch_by_kind(f ? KIND_RHS : KIND_COUPLING, v);
    HIP_TRY(hipGetLastError());
    if (f) HIP_TRY(hipMemcpyAsync(f, f_scr, sizeof(R) * N, hipMemcpyDeviceToHost, stream));
    if (cpl) HIP_TRY(hipMemcpyAsync(cpl, pairs, sizeof(R) * 2 * N, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return KMT_OK;
  }

  // kmt_rhs on device vectors: theta, f and the couplings stay where the
  // caller keeps them, the (cos, sin) vector goes to the caller's buffer cs
  // (rhs_scratch_bytes(), zero past N), and the launches go to stream s
  // (nullptr: the context's stream) without a synchronisation.  Nothing is
  // allocated; the factored path uses the context's partial-sum buffers.
  int rhs_device(const void *th, void *f, void *cpl, void *cs, hipStream_t s) override {
    if (!coupling_ready()) return fail(KMT_ERR_STATE, "no adjacency: call kmt_set_adjacency, kmt_build_adjacency or kmt_set_factor + kmt_use_factored");
    const hipStream_t own = stream;
    stream = s ? s : own;
    sincos_pad_kernel<R><<<(unsigned)((ldv + kBlock - 1) / kBlock), kBlock, 0, stream>>>((const R *)th, N, ldv, (R *)cs);
    Vecs<R> v = vecs((const R *)cs, nullptr, 0);
    v.f_out = (R *)f;
    v.cs_pairs = (R *)cpl;
    launch_by_kind(f ? KIND_RHS : KIND_COUPLING, v);
    stream = own;
    HIP_TRY(hipGetLastError());
    return KMT_OK;
  }
  size_t rhs_scratch_bytes() const override { return sizeof(R) * 2 * (size_t)ldv; }
  int device_index() const override { return device; }

  int order(const void *th, double *r) override {
    HIP_TRY(hipMemcpyAsync(vec_scr, th, sizeof(R) * N, hipMemcpyHostToDevice, stream));
    const int nb = (int)((N + kBlock - 1) / kBlock);
    order_partial_kernel<R><<<nb, kBlock, 0, stream>>>(vec_scr, N, ord_scr);
    order_final_kernel<R><<<1, kBlock, 0, stream>>>(ord_scr, nb, N, ord_scr + 2 * nb);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(r, ord_scr + 2 * nb, sizeof(double), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return KMT_OK;
  }

  // ---------------- integration ----------------
  int begin(const double *te, int ne, double t1, int keep, int rcap) override {
    if (!coupling_ready()) return fail(KMT_ERR_STATE, "no adjacency: call kmt_set_adjacency, kmt_build_adjacency or kmt_set_factor + kmt_use_factored");
    if (ne < 0) return fail(KMT_ERR_ARG, "n_eval < 0");
    if (rcap < 1) return fail(KMT_ERR_ARG, "ring_rows must be >= 1");
    HIP_TRY(hipStreamSynchronize(stream));
    drop_graphs();
    if (t_eval) { (void)hipFree(t_eval); t_eval = nullptr; }
    if (ring) { (void)hipFree(ring); ring = nullptr; }
    if (r_ring) { (void)hipFree(r_ring); r_ring = nullptr; }
    if (times) { (void)hipFree(times); times = nullptr; }
    if (ord_part) { (void)hipFree(ord_part); ord_part = nullptr; }
    n_eval = ne;
    keep_rows = keep;
    ring_cap = rcap;
    nblk_emit = (int)((N + kBlock - 1) / kBlock);
    std::vector<R> ter(std::max(ne, 1));
    for (int q = 0; q < ne; ++q) ter[q] = (R)te[q];
    HIP_TRY(hipMalloc((void **)&t_eval, sizeof(R) * ter.size()));
    HIP_TRY(hipMemcpyAsync(t_eval, ter.data(), sizeof(R) * ter.size(), hipMemcpyHostToDevice, stream));
    if (keep) {
      hipError_t e = hipMalloc((void **)&ring, sizeof(R) * (size_t)ring_cap * ldv);
      if (e != hipSuccess) return fail(KMT_ERR_MEMORY, "cannot allocate the output ring (%d rows): %s", ring_cap, hipGetErrorString(e));
    }
    HIP_TRY(hipMalloc((void **)&r_ring, sizeof(R) * ring_cap));
    HIP_TRY(hipMalloc((void **)&times, sizeof(R) * 2 * times_cap));
    HIP_TRY(hipMalloc((void **)&ord_part, sizeof(double) * 2 * (size_t)ring_cap * nblk_emit));
    // state
    std::memset(&shadow, 0, sizeof(shadow));
    shadow.t = t0_;
    shadow.h = h0_;
    shadow.t_last = t0_;
    shadow.t1 = (R)t1;
    shadow.atl = atl;
    shadow.rtl = rtl;
    shadow.fixed = fixed;
    shadow.active = (t0_ < (R)t1) ? 1 : 0;
    shadow.n_eval = ne;
    shadow.ring_cap = ring_cap;
    shadow.times_cap = times_cap;
    shadow.rhs_evals = 1;
    KMT_TRY(push_state());
    if (!tuned && !factored && !sym()) default_variant();
    // k1 = f(y) (dp45.cu:151), then the first stage-2 vector
    sincos_kernel<R><<<(unsigned)((N + kBlock - 1) / kBlock), kBlock, 0, stream>>>(y, N, ldv, cs[1]);
    Vecs<R> v = vecs(cs[1], nullptr, 0);
    v.f_out = k[0];
    launch_by_kind(KIND_RHS, v);
    emit_commit_kernel<R><<<(unsigned)nblk_emit, kBlock, 0, stream>>>(emit_args());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(stream));
    begun = 1;
    return KMT_OK;
  }

  void fill_status(kmt_status *s) {
    if (!s) return;
    s->t = (double)shadow.t;
    s->h = (double)shadow.h;
    s->t_last = (double)shadow.t_last;
    s->accepted = (long)shadow.accepted;
    s->rejected = (long)shadow.rejected;
    s->attempts = (long)shadow.attempts;
    s->rhs_evals = (long)shadow.rhs_evals;
    s->active = shadow.active;
    s->error = shadow.error;
    s->rows_ready = shadow.ring_count;
    s->times_ready = shadow.times_count;
    s->eval_i = shadow.eval_i;
    s->pending = shadow.commit_pending ? shadow.pend_hi - shadow.pend_lo : 0;
  }

  int advance(int max_attempts, kmt_status *s) override {
    if (!begun) return fail(KMT_ERR_STATE, "kmt_advance before kmt_begin");
    if (max_attempts < 0) return fail(KMT_ERR_ARG, "max_attempts < 0");
    KMT_TRY(fetch_state());
    int n = std::min(max_attempts, shadow.times_cap - shadow.times_count);
    if (shadow.active && n > 0) {
      if (!graphs_valid) KMT_TRY(build_graphs());
      for (; n >= 8; n -= 8) HIP_TRY(hipGraphLaunch(g8, stream));
      for (; n > 0; --n) HIP_TRY(hipGraphLaunch(g1, stream));
    }
    KMT_TRY(fetch_state());
    fill_status(s);
    if (shadow.error) return fail(KMT_ERR_NUMERIC, "error norm is not finite at t=%g (h=%g): integration stopped", (double)shadow.t, (double)shadow.h);
    return KMT_OK;
  }

  // cpl (optional): the couplings (C_i, S_i) of every drained row, 2N values
  // a row, computed from the row where it sits in the ring (the launches of
  // kmt_rhs's coupling form, one row at a time through the shared (cos, sin)
  // and pair buffers, each copied out before the next row overwrites them).
  int read_outputs(int max_rows, void *rows, void *rv, void *cpl, int *nout) override {
    if (!begun) return fail(KMT_ERR_STATE, "kmt_read_outputs before kmt_begin");
    KMT_TRY(fetch_state());
    const int ready = shadow.ring_count;
    const int n = std::min(std::max(max_rows, 0), ready);
    if ((rows || cpl) && n > 0 && !keep_rows) return fail(KMT_ERR_STATE, "rows were not kept (kmt_begin keep_rows=0)");
    if (cpl && n > 0 && !coupling_ready()) return fail(KMT_ERR_STATE, "no adjacency for the couplings");
    if (rows && n > 0)
      HIP_TRY(hipMemcpy2DAsync(rows, sizeof(R) * N, ring, sizeof(R) * ldv, sizeof(R) * N, n, hipMemcpyDeviceToHost, stream));
    for (int q = 0; cpl && q < n; ++q) {
      sincos_kernel<R><<<(unsigned)((N + kBlock - 1) / kBlock), kBlock, 0, stream>>>(ring + (size_t)q * ldv, N, ldv, cs_scr);
      Vecs<R> v = vecs(cs_scr, nullptr, 0);
      v.f_out = f_scr;
      v.cs_pairs = pairs;
      launch_by_kind(KIND_COUPLING, v);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync((R *)cpl + (size_t)q * 2 * N, pairs, sizeof(R) * 2 * N, hipMemcpyDeviceToHost, stream));
    }
    if (rv && n > 0) HIP_TRY(hipMemcpyAsync(rv, r_ring, sizeof(R) * n, hipMemcpyDeviceToHost, stream));
    // keep the unread rows at the front of the ring (in order)
    for (int q = n; q < ready; ++q) {
      if (keep_rows) HIP_TRY(hipMemcpyAsync(ring + (size_t)(q - n) * ldv, ring + (size_t)q * ldv, sizeof(R) * N, hipMemcpyDeviceToDevice, stream));
      HIP_TRY(hipMemcpyAsync(r_ring + (q - n), r_ring + q, sizeof(R), hipMemcpyDeviceToDevice, stream));
    }
    shadow.ring_count = ready - n;
    HIP_TRY(hipMemcpyAsync(&st->ring_count, &shadow.ring_count, sizeof(int), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (nout) *nout = n;
    return KMT_OK;
  }

  int read_times(int max_n, void *tv, void *hv, int *nout) override {
    if (!begun) return fail(KMT_ERR_STATE, "kmt_read_times before kmt_begin");
    KMT_TRY(fetch_state());
    const int ready = shadow.times_count;
    const int n = std::min(std::max(max_n, 0), ready);
    std::vector<R> buf(2 * std::max(ready, 1));
    if (ready > 0) {
      HIP_TRY(hipMemcpyAsync(buf.data(), times, sizeof(R) * 2 * ready, hipMemcpyDeviceToHost, stream));
      HIP_TRY(hipStreamSynchronize(stream));
    }
    for (int q = 0; q < n; ++q) {
      if (tv) ((R *)tv)[q] = buf[2 * q];
      if (hv) ((R *)hv)[q] = buf[2 * q + 1];
    }
    if (n < ready) HIP_TRY(hipMemcpyAsync(times, buf.data() + 2 * n, sizeof(R) * 2 * (ready - n), hipMemcpyHostToDevice, stream));
    shadow.times_count = ready - n;
    HIP_TRY(hipMemcpyAsync(&st->times_count, &shadow.times_count, sizeof(int), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (nout) *nout = n;
    return KMT_OK;
  }

  int step(int *acc, double *t, double *h) override {
    if (!begun) return fail(KMT_ERR_STATE, "kmt_step before kmt_begin");
    KMT_TRY(fetch_state());
    const long long before = shadow.accepted;
    if (shadow.active && shadow.times_count < shadow.times_cap) {
      if (!graphs_valid) KMT_TRY(build_graphs());
      HIP_TRY(hipGraphLaunch(g1, stream));
    }
    KMT_TRY(fetch_state());
    if (acc) *acc = shadow.accepted > before ? 1 : 0;
    if (t) *t = (double)shadow.t;
    if (h) *h = (double)shadow.h;
    if (shadow.error) return fail(KMT_ERR_NUMERIC, "error norm is not finite at t=%g", (double)shadow.t);
    return KMT_OK;
  }

  int eval(double te, void *out) override {
    if (!begun) return fail(KMT_ERR_STATE, "kmt_eval before kmt_begin");
    interp_kernel<R><<<(unsigned)((N + kBlock - 1) / kBlock), kBlock, 0, stream>>>(st, N, ylast, k[0], k[1], k[2], k[3], k[4], k[5], k[6], (R)te, vec_scr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, vec_scr, sizeof(R) * N, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return KMT_OK;
  }

  // ---------------- timing ----------------
  int time_stage(int reps, double *avg_ms) override {
    if (!coupling_ready()) return fail(KMT_ERR_STATE, "no adjacency");
    if (reps < 1) return fail(KMT_ERR_ARG, "reps < 1");
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    // stage 3 on scratch k buffers (the integrator's k_j stay untouched)
    if (!k_scr) {
      HIP_TRY(hipMalloc((void **)&k_scr, sizeof(R) * 7 * ldv));
      HIP_TRY(hipMemsetAsync(k_scr, 0, sizeof(R) * 7 * ldv, stream));
    }
    Vecs<R> v = vecs(cs[0], cs_scr, 0);
    for (int j = 0; j < 7; ++j) v.k[j] = k_scr + j * ldv;
    launch_by_kind(3, v);  // warm
    HIP_TRY(hipEventRecord(e0, stream));
    for (int r = 0; r < reps; ++r) launch_by_kind(3, v);
    HIP_TRY(hipEventRecord(e1, stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (avg_ms) *avg_ms = ms / reps;
    return KMT_OK;
  }

  int time_attempts(int attempts, double *region_ms, double *stage_avg_ms, kmt_status *s) override {
    if (!begun) return fail(KMT_ERR_STATE, "kmt_time_attempts before kmt_begin");
    if (attempts < 1) return fail(KMT_ERR_ARG, "attempts < 1");
    KMT_TRY(fetch_state());
    if (attempts > shadow.times_cap - shadow.times_count)
      return fail(KMT_ERR_ARG, "attempts exceed the accepted-step log capacity; drain kmt_read_times first");
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pairs_all(6 * attempts);
    for (auto &p : pairs_all) {
      HIP_TRY(hipEventCreate(&p.first));
      HIP_TRY(hipEventCreate(&p.second));
    }
    hipEvent_t r0, r1;
    HIP_TRY(hipEventCreate(&r0));
    HIP_TRY(hipEventCreate(&r1));
    std::vector<std::pair<hipEvent_t, hipEvent_t>> work = pairs_all;
    HIP_TRY(hipEventRecord(r0, stream));
    for (int a = 0; a < attempts; ++a) launch_attempt(&work);
    HIP_TRY(hipEventRecord(r1, stream));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventSynchronize(r1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, r0, r1));
    double tot = 0;
    for (auto &p : pairs_all) {
      float m = 0;
      HIP_TRY(hipEventElapsedTime(&m, p.first, p.second));
      tot += m;
      (void)hipEventDestroy(p.first);
      (void)hipEventDestroy(p.second);
    }
    (void)hipEventDestroy(r0);
    (void)hipEventDestroy(r1);
    if (region_ms) *region_ms = ms;
    if (stage_avg_ms) *stage_avg_ms = tot / pairs_all.size();
    KMT_TRY(fetch_state());
    fill_status(s);
    if (shadow.error) return fail(KMT_ERR_NUMERIC, "error norm is not finite");
    return KMT_OK;
  }
};


EngineBase *new_engine_f32();
EngineBase *new_engine_f64();

}  // namespace kmt
