// The C ABI declared in include/kuramoto_amd.h for single networks (the
// engine itself is kmt_engine.hpp), host-side generators, bandwidth probes
// and error reporting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/kuramoto_amd.h"
#include "kmt_common.hpp"
#include "kmt_engine.hpp"
#include "kmt_rng.h"

using namespace kmt;

namespace kmt {
// ---------------------------------------------------------------------
// Bandwidth probes: a plain 16-byte copy and a read-only stream of known
// size (the measured ceiling next to the nominal 8 TB/s, and the
// calibration kernel for the HBM byte counters).
// ---------------------------------------------------------------------
// U independent 16-byte loads in flight per lane, grid-stride.
template <int U>
__global__ __launch_bounds__(kBlock) void copy_probe_kernel(const u32x4 *src, u32x4 *dst, size_t n16) {
  const size_t stride = (size_t)gridDim.x * kBlock;
  size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  for (; i + (U - 1) * stride < n16; i += U * stride) {
    u32x4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
    for (int u = 0; u < U; ++u) dst[i + u * stride] = v[u];
  }
  for (; i < n16; i += stride) dst[i] = __builtin_nontemporal_load(src + i);
}

template <int U>
__global__ __launch_bounds__(kBlock) void read_probe_kernel(const u32x4 *src, size_t n16, unsigned int *sink) {
  const size_t stride = (size_t)gridDim.x * kBlock;
  size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  unsigned int acc = 0;
  for (; i + (U - 1) * stride < n16; i += U * stride) {
    u32x4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
    for (int u = 0; u < U; ++u) acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
  }
  for (; i < n16; i += stride) {
    const u32x4 v = __builtin_nontemporal_load(src + i);
    acc ^= v.x ^ v.y ^ v.z ^ v.w;
  }
  if (acc == 0x9E3779B9u) sink[0] = acc;  // keeps the loads alive; practically never taken
}

}  // namespace kmt

// ---------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------
static thread_local std::string g_last_error;

int kmt_fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}


struct kmt_ctx {
  int precision;
  EngineBase *eng;
};

// ---------------------------------------------------------------------
// Scratch of kmt_rhs_noalloc: the (cos, sin) vector of one RHS, 2 ldv
// values of the working precision (4 MiB: N <= 524288 in FP32, 262144 in
// FP64).  A file-scope device array, so the entry point allocates nothing;
// every GPU holds its own copy until the process ends.  In host code the
// array's name is only a host-side stand-in: its device address on each
// GPU comes from hipGetSymbolAddress, once per device with that device
// current, and only that address is passed to kernels.  Every call on a
// GPU shares that GPU's copy (calls that overlap on one GPU race); each
// call rewrites all of the values it reads, zeros past N included.
// ---------------------------------------------------------------------
static constexpr size_t kRhsScratchBytes = size_t(4) << 20;
static __device__ __align__(256) unsigned char scratch[kRhsScratchBytes];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices] = {};

static int rhs_scratch_ptr(int device, void **out) {
  if (device < 0 || device >= kMaxDevices) return kmt_fail(KMT_ERR_ARG, "device %d out of range", device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    HIP_TRY(hipGetDevice(&cur));
    HIP_TRY(hipSetDevice(device));
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    const hipError_t e2 = hipSetDevice(cur);
    if (e != hipSuccess || !ptr)
      return kmt_fail(KMT_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                      hipGetErrorString(e));
    if (e2 != hipSuccess) return kmt_fail(KMT_ERR_HIP, "hipSetDevice(%d) failed: %s", cur, hipGetErrorString(e2));
    g_scratch_ptr[device] = ptr;
  }
  *out = g_scratch_ptr[device];
  return KMT_OK;
}

// ---------------------------------------------------------------------
// host generators
// ---------------------------------------------------------------------
template <class R> static R draw_ic(R init);
template <> float draw_ic<float>(float init) { return (float)((double)(init / 2) * (2.0 / RAND_MAX * rand() - 1)); }
template <> double draw_ic<double>(double init) { return (init / 2) * (2.0 / RAND_MAX * rand() - 1); }

// kuramoto.cu:526-536 (float u, v; log(float) is logf) and
// kuramoto_64.cu:523-533 (double u, v).
static float draw_freq_f32(float freq, int normal) {
  if (normal) {
    const float u = rand() / (float)RAND_MAX;
    const float v = rand() / (float)RAND_MAX;
    const float lu = -2 * logf(u);
    return (float)((double)freq * std::pow((double)lu, 0.5) * std::cos(2 * M_PI * (double)v));
  }
  const float u = rand() / (float)RAND_MAX;
  return (float)((double)freq * std::tan(M_PI * ((double)u - 0.5)));
}
static double draw_freq_f64(double freq, int normal) {
  if (normal) {
    const double u = rand() / (double)RAND_MAX;
    const double v = rand() / (double)RAND_MAX;
    return freq * std::pow(-2 * std::log(u), 0.5) * std::cos(2 * M_PI * v);
  }
  const double u = rand() / (double)RAND_MAX;
  return freq * std::tan(M_PI * (u - 0.5));
}

extern "C" {

const char *kmt_last_error(void) { return g_last_error.c_str(); }

int kmt_device_count(int *count) {
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (count) *count = n;
  return KMT_OK;
}

int kmt_host_alloc(size_t bytes, void **ptr) {
  if (!ptr) return fail(KMT_ERR_ARG, "ptr is NULL");
  *ptr = nullptr;
  hipError_t e = hipHostMalloc(ptr, bytes > 0 ? bytes : 1, hipHostMallocDefault);
  if (e != hipSuccess) {
    *ptr = nullptr;
    return fail(KMT_ERR_MEMORY, "cannot allocate %zu bytes of pinned host memory: %s", bytes, hipGetErrorString(e));
  }
  return KMT_OK;
}
int kmt_host_free(void *ptr) {
  if (ptr) HIP_TRY(hipHostFree(ptr));
  return KMT_OK;
}
int kmt_mem_info(int device, size_t *free_bytes, size_t *total_bytes) {
  HIP_TRY(hipSetDevice(device));
  size_t f = 0, t = 0;
  HIP_TRY(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return KMT_OK;
}

int kmt_draws(int precision, long N, int seed, int normal, double freq, double init, void *theta0, void *omega) {
  if (N < 0) return fail(KMT_ERR_ARG, "N < 0");
  srand(seed);
  if (precision == KMT_F32) {
    float *th = (float *)theta0, *om = (float *)omega;
    if (th)
      for (long j = 0; j < N; ++j) th[j] = draw_ic<float>((float)init);
    if (om)
      for (long j = 0; j < N; ++j) om[j] = draw_freq_f32((float)freq, normal);
  } else if (precision == KMT_F64) {
    double *th = (double *)theta0, *om = (double *)omega;
    if (th)
      for (long j = 0; j < N; ++j) th[j] = draw_ic<double>(init);
    if (om)
      for (long j = 0; j < N; ++j) om[j] = draw_freq_f64(freq, normal);
  } else {
    return fail(KMT_ERR_ARG, "precision must be KMT_F32 or KMT_F64");
  }
  return KMT_OK;
}

int kmt_factor(long N, int K, uint32_t seed, int gaussian, double *X) {
  if (N < 0 || K < 0 || !X) return fail(KMT_ERR_ARG, "bad factor arguments");
  if (gaussian) kmt::gaussian_factor(N, K, seed, X);
  else kmt::sign_factor(N, K, seed, X);
  return KMT_OK;
}

int kmt_factor_weights(int precision, long N, int K, int gaussian, void *w) {
  if (N < 1 || K < 0 || !w) return fail(KMT_ERR_ARG, "bad weight arguments");
  for (int m = 0; m < K; ++m) {
    const double v = gaussian ? -1.0 / ((double)N * (double)K) : ((m & 1) ? -1.0 : 1.0) / (double)N;
    if (precision == KMT_F32) ((float *)w)[m] = (float)v;
    else ((double *)w)[m] = v;
  }
  return KMT_OK;
}

int kmt_create(kmt_ctx **ctx, int device, int precision, long N, double atol, double rtol, int fixed_step, double c1) {
  if (!ctx) return fail(KMT_ERR_ARG, "ctx is NULL");
  *ctx = nullptr;
  if (N < 1) return fail(KMT_ERR_ARG, "N must be >= 1 (got %ld)", N);
  if (precision != KMT_F32 && precision != KMT_F64) return fail(KMT_ERR_ARG, "precision must be KMT_F32 or KMT_F64");
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(KMT_ERR_ARG, "device %d not present (%d devices)", device, ndev);
  kmt_ctx *c = new kmt_ctx;
  c->precision = precision;
  int rc;
  if (precision == KMT_F32) {
    c->eng = new_engine_f32();
    rc = c->eng->create(device, N, atol, rtol, fixed_step, c1);
  } else {
    c->eng = new_engine_f64();
    rc = c->eng->create(device, N, atol, rtol, fixed_step, c1);
  }
  if (rc != KMT_OK) {
    std::string msg = g_last_error;
    delete c->eng;
    delete c;
    g_last_error = msg;
    return rc;
  }
  *ctx = c;
  return KMT_OK;
}

int kmt_destroy(kmt_ctx *ctx) {
  if (!ctx) return KMT_OK;
  delete ctx->eng;
  delete ctx;
  return KMT_OK;
}

#define CTX_CHECK                                                   \
  if (!ctx || !ctx->eng) return fail(KMT_ERR_ARG, "ctx is NULL");

int kmt_set_omega(kmt_ctx *ctx, const void *omega) {
  CTX_CHECK;
  if (!omega) return fail(KMT_ERR_ARG, "omega is NULL");
  return ctx->eng->set_omega(omega);
}
int kmt_set_state(kmt_ctx *ctx, const void *theta, double t, double h) {
  CTX_CHECK;
  if (!theta) return fail(KMT_ERR_ARG, "theta is NULL");
  return ctx->eng->set_state(theta, t, h);
}
int kmt_get_state(kmt_ctx *ctx, void *theta, double *t, double *h) {
  CTX_CHECK;
  if (!theta) return fail(KMT_ERR_ARG, "theta is NULL");
  return ctx->eng->get_state(theta, t, h);
}
int kmt_set_adjacency(kmt_ctx *ctx, const void *M, int src_precision, long ld, int storage) {
  CTX_CHECK;
  if (!M) return fail(KMT_ERR_ARG, "M is NULL");
  return ctx->eng->set_adjacency(M, src_precision, ld, storage);
}
int kmt_set_factor(kmt_ctx *ctx, const void *X, const void *w, int K) {
  CTX_CHECK;
  if (K > 0 && (!X || !w)) return fail(KMT_ERR_ARG, "X or w is NULL");
  return ctx->eng->set_factor(X, w, K);
}
int kmt_set_factor_sign(kmt_ctx *ctx, int K, uint32_t seed) {
  CTX_CHECK;
  return ctx->eng->set_factor_sign(K, seed);
}
int kmt_build_adjacency(kmt_ctx *ctx, int storage) {
  CTX_CHECK;
  return ctx->eng->build_adjacency(storage);
}
int kmt_use_factored(kmt_ctx *ctx, int on) {
  CTX_CHECK;
  return ctx->eng->use_factored(on);
}
int kmt_get_adjacency_rows(kmt_ctx *ctx, long i0, long i1, void *out) {
  CTX_CHECK;
  if (!out) return fail(KMT_ERR_ARG, "out is NULL");
  return ctx->eng->get_adjacency_rows(i0, i1, out);
}
int kmt_get_adjacency_scale(kmt_ctx *ctx, double *scale) {
  CTX_CHECK;
  if (scale) *scale = ctx->eng->mscale_;
  return KMT_OK;
}
int kmt_rhs(kmt_ctx *ctx, const void *theta, void *f, void *couplings) {
  CTX_CHECK;
  if (!theta || (!f && !couplings)) return fail(KMT_ERR_ARG, "theta is NULL or no output requested");
  return ctx->eng->rhs(theta, f, couplings);
}
int kmt_rhs_noalloc(kmt_ctx *ctx, const void *theta_dev, void *f_dev, void *couplings_dev, void *stream) {
  CTX_CHECK;
  if (!theta_dev || (!f_dev && !couplings_dev)) return fail(KMT_ERR_ARG, "theta_dev is NULL or no output requested");
  const size_t need = ctx->eng->rhs_scratch_bytes();
  if (need > kRhsScratchBytes)
    return fail(KMT_ERR_ARG,
                "kmt_rhs_noalloc: this context's (cos, sin) vector needs %zu bytes, more than the library's "
                "%zu-byte scratch (N <= %zu in FP32, %zu in FP64)",
                need, kRhsScratchBytes, kRhsScratchBytes / 8, kRhsScratchBytes / 16);
  void *cs = nullptr;
  KMT_TRY(rhs_scratch_ptr(ctx->eng->device_index(), &cs));
  return ctx->eng->rhs_device(theta_dev, f_dev, couplings_dev, cs, (hipStream_t)stream);
}
int kmt_order(kmt_ctx *ctx, const void *theta, double *r) {
  CTX_CHECK;
  if (!theta || !r) return fail(KMT_ERR_ARG, "NULL argument");
  return ctx->eng->order(theta, r);
}
int kmt_begin(kmt_ctx *ctx, const double *t_eval, int n_eval, double t1, int keep_rows, int ring_rows) {
  CTX_CHECK;
  if (n_eval > 0 && !t_eval) return fail(KMT_ERR_ARG, "t_eval is NULL");
  return ctx->eng->begin(t_eval, n_eval, t1, keep_rows, ring_rows);
}
int kmt_advance(kmt_ctx *ctx, int max_attempts, kmt_status *st) {
  CTX_CHECK;
  return ctx->eng->advance(max_attempts, st);
}
int kmt_read_outputs(kmt_ctx *ctx, int max_rows, void *theta_rows, void *r, int *n_rows) {
  CTX_CHECK;
  return ctx->eng->read_outputs(max_rows, theta_rows, r, nullptr, n_rows);
}
int kmt_read_outputs_couplings(kmt_ctx *ctx, int max_rows, void *theta_rows, void *r, void *couplings, int *n_rows) {
  CTX_CHECK;
  return ctx->eng->read_outputs(max_rows, theta_rows, r, couplings, n_rows);
}
int kmt_read_times(kmt_ctx *ctx, int max_n, void *t, void *h, int *n) {
  CTX_CHECK;
  return ctx->eng->read_times(max_n, t, h, n);
}
int kmt_step(kmt_ctx *ctx, int *accepted, double *t, double *h) {
  CTX_CHECK;
  return ctx->eng->step(accepted, t, h);
}
int kmt_eval(kmt_ctx *ctx, double t_eval, void *theta) {
  CTX_CHECK;
  if (!theta) return fail(KMT_ERR_ARG, "theta is NULL");
  return ctx->eng->eval(t_eval, theta);
}
int kmt_time_stage_kernel(kmt_ctx *ctx, int reps, double *avg_ms) {
  CTX_CHECK;
  return ctx->eng->time_stage(reps, avg_ms);
}
int kmt_time_attempts(kmt_ctx *ctx, int attempts, double *region_ms, double *stage_avg_ms, kmt_status *st) {
  CTX_CHECK;
  return ctx->eng->time_attempts(attempts, region_ms, stage_avg_ms, st);
}
int kmt_select_variant(kmt_ctx *ctx, int rows_per_block, int variant_flags) {
  CTX_CHECK;
  return ctx->eng->select_variant(rows_per_block, variant_flags);
}
int kmt_layout(kmt_ctx *ctx, long *lda, int *elem_bytes, int *rows_per_block, int *variant_flags) {
  CTX_CHECK;
  return ctx->eng->layout(lda, elem_bytes, rows_per_block, variant_flags);
}

int kmt_autotune(kmt_ctx *ctx, int reps, double *variant_ms) {
  CTX_CHECK;
  if (reps < 1) return fail(KMT_ERR_ARG, "reps < 1");
  return ctx->eng->tune(reps, variant_ms);
}

int kmt_bandwidth_probe(int device, size_t bytes, int reps, double *copy_gbps, double *read_gbps) {
  if (bytes < (1u << 20) || reps < 1) return fail(KMT_ERR_ARG, "probe needs >= 1 MiB and reps >= 1");
  HIP_TRY(hipSetDevice(device));
  const size_t n16 = bytes / 16;
  void *a = nullptr, *b = nullptr;
  unsigned int *sink = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&]() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (sink) (void)hipFree(sink);
    if (s) (void)hipStreamDestroy(s);
  };
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&a, n16 * 16);
  if (e == hipSuccess) e = hipMalloc(&b, n16 * 16);
  if (e == hipSuccess) e = hipMalloc((void **)&sink, 64);
  if (e == hipSuccess) e = hipMemsetAsync(a, 1, n16 * 16, s);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    cleanup();
    return fail(e == hipErrorOutOfMemory ? KMT_ERR_MEMORY : KMT_ERR_HIP, "probe setup: %s", hipGetErrorString(e));
  }
  int cus = 256;  // kept if the attribute query fails
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1) cus = 256;
  // The best of a few launch shapes (loads in flight per lane x workgroups
  // per CU), so that the figure is the device's ceiling, not one shape's.
  double best_copy = 0, best_read = 0;
  auto timed = [&](auto launch) -> double {
    float ms = 0;
    launch();
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    for (int r = 0; r < reps; ++r) launch();
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    return e == hipSuccess && ms > 0 ? 1.0 * n16 * 16 * reps / (ms * 1e-3) / 1e9 : 0.0;
  };
  for (int shape = 0; shape < 3 && e == hipSuccess; ++shape) {
    const size_t per_cu = shape == 0 ? 8 : shape == 1 ? 16 : 32;
    const unsigned grid = (unsigned)std::min<size_t>((n16 + kBlock - 1) / kBlock, (size_t)cus * per_cu);
    const double c = timed([&]() {
      if (shape == 0) copy_probe_kernel<4><<<grid, kBlock, 0, s>>>((const u32x4 *)a, (u32x4 *)b, n16);
      else copy_probe_kernel<8><<<grid, kBlock, 0, s>>>((const u32x4 *)a, (u32x4 *)b, n16);
    });
    const double r = timed([&]() {
      if (shape == 0) read_probe_kernel<4><<<grid, kBlock, 0, s>>>((const u32x4 *)a, n16, sink);
      else read_probe_kernel<8><<<grid, kBlock, 0, s>>>((const u32x4 *)a, n16, sink);
    });
    best_copy = std::max(best_copy, 2.0 * c);
    best_read = std::max(best_read, r);
  }
  cleanup();
  if (e != hipSuccess) return fail(KMT_ERR_HIP, "probe: %s", hipGetErrorString(e));
  if (copy_gbps) *copy_gbps = best_copy;
  if (read_gbps) *read_gbps = best_read;
  return KMT_OK;
}

}  // extern "C"
