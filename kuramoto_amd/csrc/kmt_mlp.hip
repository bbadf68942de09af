// Learned step-size proposer: training kernels of the network in
// kmt_mlp.hpp (8 -> 64 -> 64 -> 1, tanh, tanh, linear).  FP16 weights for the
// forward pass, FP32 master weights for the optimizer, FP32 accumulation
// everywhere.  No reference counterpart (an extension of the path).
//
// Loss: mean squared error (1/S) sum_s (out_s - y_s)^2.
//
//   mlp_fwd_bwd_kernel  one wave per sample, one lane per hidden unit: the
//                       forward pass the batched controller runs
//                       (mlp_wave_forward), output delta, back-propagated
//                       deltas; stores h1, h2, d1, d2 (S x 64) and d3 (S)
//   mlp_grad_kernel     one thread per parameter, fixed-order sum over the
//                       samples (deterministic)
//   mlp_adam_kernel     Adam update of the FP32 master weights, FP16 copy
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/kuramoto_amd.h"
#include "kmt_common.hpp"
#include "kmt_kernels.hpp"
#include "kmt_mlp.hpp"
#include "kmt_rng.h"

namespace kmt {

__global__ __launch_bounds__(kBlock) void mlp_fwd_bwd_kernel(const uint16_t *w, const float *X, const float *y, int S,
                                                             float *H1, float *H2, float *D1, float *D2, float *D3,
                                                             float *out, double *loss_part) {
  __shared__ double lsum[kBlock / kWave];
  const int wv = threadIdx.x / kWave, i = threadIdx.x % kWave;
  const int s = blockIdx.x * (kBlock / kWave) + wv;
  const bool valid = s < S;
  float xs[kIn];
#pragma unroll
  for (int j = 0; j < kIn; ++j) xs[j] = valid ? X[(size_t)s * kIn + j] : 0.f;
  float h1, h2;
  const float o = mlp_wave_forward(w, xs, h1, h2);
  const float ys = valid ? y[s] : 0.f;
  const float d3 = valid ? 2.f * (o - ys) / (float)S : 0.f;
  // back-propagation: d2_i = d3 W3_i (1 - h2_i^2), d1_i = (sum_k W2_ki d2_k) (1 - h1_i^2)
  const float d2 = d3 * hf(w[kOffW3 + i]) * (1.f - h2 * h2);
  float g = 0.f;
  for (int k = 0; k < kHid; ++k) g = fmaf(hf(w[kOffW2 + k * kHid + i]), __shfl(d2, k, kWave), g);
  const float d1 = g * (1.f - h1 * h1);
  if (valid) {
    H1[(size_t)s * kHid + i] = h1;
    H2[(size_t)s * kHid + i] = h2;
    D1[(size_t)s * kHid + i] = d1;
    D2[(size_t)s * kHid + i] = d2;
    if (i == 0) {
      D3[s] = d3;
      if (out) out[s] = o;
    }
  }
  if (i == 0) lsum[wv] = valid ? (double)(o - ys) * (double)(o - ys) : 0.0;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int q = 0; q < kBlock / kWave; ++q) t += lsum[q];
    loss_part[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kBlock) void mlp_grad_kernel(const float *X, int S, const float *H1, const float *H2,
                                                          const float *D1, const float *D2, const float *D3,
                                                          float *grad) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= kNParams) return;
  float acc = 0.f;
  if (p < kOffB1) {
    const int i = p / kIn, j = p % kIn;
    for (int s = 0; s < S; ++s) acc = fmaf(D1[(size_t)s * kHid + i], X[(size_t)s * kIn + j], acc);
  } else if (p < kOffW2) {
    const int i = p - kOffB1;
    for (int s = 0; s < S; ++s) acc += D1[(size_t)s * kHid + i];
  } else if (p < kOffB2) {
    const int q = p - kOffW2, i = q / kHid, j = q % kHid;
    for (int s = 0; s < S; ++s) acc = fmaf(D2[(size_t)s * kHid + i], H1[(size_t)s * kHid + j], acc);
  } else if (p < kOffW3) {
    const int i = p - kOffB2;
    for (int s = 0; s < S; ++s) acc += D2[(size_t)s * kHid + i];
  } else if (p < kOffB3) {
    const int i = p - kOffW3;
    for (int s = 0; s < S; ++s) acc = fmaf(D3[s], H2[(size_t)s * kHid + i], acc);
  } else {
    for (int s = 0; s < S; ++s) acc += D3[s];
  }
  grad[p] = acc;
}

__global__ __launch_bounds__(kBlock) void mlp_adam_kernel(float *w32, uint16_t *w16, float *m, float *v,
                                                          const float *grad, float lr, float b1, float b2, float eps,
                                                          float c1, float c2) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= kNParams) return;
  const float g = grad[p];
  const float mm = b1 * m[p] + (1.f - b1) * g;
  const float vv = b2 * v[p] + (1.f - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  const float wn = w32[p] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
  w32[p] = wn;
  w16[p] = __builtin_bit_cast(uint16_t, (_Float16)wn);
}

}  // namespace kmt

using namespace kmt;

struct kmt_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  float *w32 = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
  uint16_t *w16 = nullptr;
  float *X = nullptr, *y = nullptr, *H1 = nullptr, *H2 = nullptr, *D1 = nullptr, *D2 = nullptr, *D3 = nullptr,
        *out = nullptr;
  double *lpart = nullptr;
  int cap = 0;
  long step = 0;

  ~kmt_mlp() {
    if (stream) (void)hipStreamSynchronize(stream);
    void *p[] = {w32, m, v, grad, w16, X, y, H1, H2, D1, D2, D3, out, lpart};
    for (void *q : p)
      if (q) (void)hipFree(q);
    if (stream) (void)hipStreamDestroy(stream);
  }

  int init(int dev, const float *w_init, uint32_t seed) {
    device = dev;
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIP_TRY(hipMalloc((void **)&w32, sizeof(float) * kNParams));
    HIP_TRY(hipMalloc((void **)&m, sizeof(float) * kNParams));
    HIP_TRY(hipMalloc((void **)&v, sizeof(float) * kNParams));
    HIP_TRY(hipMalloc((void **)&grad, sizeof(float) * kNParams));
    HIP_TRY(hipMalloc((void **)&w16, sizeof(uint16_t) * kNParams));
    std::vector<float> w(kNParams, 0.f);
    if (w_init) {
      std::memcpy(w.data(), w_init, sizeof(float) * kNParams);
    } else {
      // Glorot-uniform weights from the library's Philox stream; zero biases
      // and a zero output layer (the untrained proposer returns the standard
      // controller's step).
      std::vector<double> u((size_t)kHid * kIn + (size_t)kHid * kHid);
      kmt::gaussian_factor((long)u.size(), 1, seed, u.data());
      const double s1 = std::sqrt(2.0 / (kIn + kHid)), s2 = std::sqrt(2.0 / (kHid + kHid));
      for (int q = 0; q < kHid * kIn; ++q) w[kOffW1 + q] = (float)(s1 * u[q]);
      for (int q = 0; q < kHid * kHid; ++q) w[kOffW2 + q] = (float)(s2 * u[kHid * kIn + q]);
    }
    std::vector<uint16_t> h(kNParams);
    for (int q = 0; q < kNParams; ++q) h[q] = __builtin_bit_cast(uint16_t, (_Float16)w[q]);
    HIP_TRY(hipMemcpyAsync(w32, w.data(), sizeof(float) * kNParams, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(w16, h.data(), sizeof(uint16_t) * kNParams, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemsetAsync(m, 0, sizeof(float) * kNParams, stream));
    HIP_TRY(hipMemsetAsync(v, 0, sizeof(float) * kNParams, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return KMT_OK;
  }

  int ensure(int n) {
    if (n <= cap) return KMT_OK;
    void *p[] = {X, y, H1, H2, D1, D2, D3, out, lpart};
    for (void *q : p)
      if (q) (void)hipFree(q);
    cap = n;
    HIP_TRY(hipMalloc((void **)&X, sizeof(float) * kIn * n));
    HIP_TRY(hipMalloc((void **)&y, sizeof(float) * n));
    HIP_TRY(hipMalloc((void **)&H1, sizeof(float) * kHid * n));
    HIP_TRY(hipMalloc((void **)&H2, sizeof(float) * kHid * n));
    HIP_TRY(hipMalloc((void **)&D1, sizeof(float) * kHid * n));
    HIP_TRY(hipMalloc((void **)&D2, sizeof(float) * kHid * n));
    HIP_TRY(hipMalloc((void **)&D3, sizeof(float) * n));
    HIP_TRY(hipMalloc((void **)&out, sizeof(float) * n));
    HIP_TRY(hipMalloc((void **)&lpart, sizeof(double) * (n / 4 + 1)));
    return KMT_OK;
  }

  // forward + backward over samples [s0, s0 + S) of the uploaded set
  void fwd_bwd(int s0, int S) {
    const int blocks = (S + 3) / 4;
    mlp_fwd_bwd_kernel<<<blocks, kBlock, 0, stream>>>(w16, X + (size_t)s0 * kIn, y + s0, S, H1, H2, D1, D2, D3,
                                                      out + s0, lpart);
    mlp_grad_kernel<<<(kNParams + kBlock - 1) / kBlock, kBlock, 0, stream>>>(X + (size_t)s0 * kIn, S, H1, H2, D1, D2,
                                                                             D3, grad);
  }

  int upload(const float *Xh, const float *yh, int n) {
    KMT_TRY(ensure(n));
    HIP_TRY(hipMemcpyAsync(X, Xh, sizeof(float) * kIn * n, hipMemcpyHostToDevice, stream));
    if (yh) HIP_TRY(hipMemcpyAsync(y, yh, sizeof(float) * n, hipMemcpyHostToDevice, stream));
    else HIP_TRY(hipMemsetAsync(y, 0, sizeof(float) * n, stream));
    return KMT_OK;
  }

  // sum of the per-block loss partials of the last fwd_bwd over S samples
  int loss_sum(int S, double *out) {
    const int blocks = (S + 3) / 4;
    std::vector<double> lp(blocks);
    HIP_TRY(hipMemcpyAsync(lp.data(), lpart, sizeof(double) * blocks, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    double t = 0;
    for (double q : lp) t += q;
    *out = t;
    return KMT_OK;
  }
};

extern "C" {

int kmt_mlp_create(kmt_mlp **out, int device, const float *w_init, uint32_t seed) {
  if (!out) return kmt_fail(KMT_ERR_ARG, "out is NULL");
  *out = nullptr;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return kmt_fail(KMT_ERR_ARG, "device %d not present", device);
  kmt_mlp *m = new kmt_mlp();
  int rc = m->init(device, w_init, seed);
  if (rc != KMT_OK) {
    delete m;
    return rc;
  }
  *out = m;
  return KMT_OK;
}

int kmt_mlp_destroy(kmt_mlp *m) {
  delete m;
  return KMT_OK;
}

int kmt_mlp_get_weights(kmt_mlp *m, float *w32, uint16_t *w16) {
  if (!m) return kmt_fail(KMT_ERR_ARG, "mlp is NULL");
  if (w32) HIP_TRY(hipMemcpyAsync(w32, m->w32, sizeof(float) * kNParams, hipMemcpyDeviceToHost, m->stream));
  if (w16) HIP_TRY(hipMemcpyAsync(w16, m->w16, sizeof(uint16_t) * kNParams, hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  return KMT_OK;
}

int kmt_mlp_forward(kmt_mlp *m, const float *Xh, int n, float *outh) {
  if (!m || !Xh || !outh || n < 1) return kmt_fail(KMT_ERR_ARG, "bad forward arguments");
  KMT_TRY(m->upload(Xh, nullptr, n));
  m->fwd_bwd(0, n);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(outh, m->out, sizeof(float) * n, hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  return KMT_OK;
}

int kmt_mlp_loss_grad(kmt_mlp *m, const float *Xh, const float *yh, int n, double *loss, float *gradh) {
  if (!m || !Xh || !yh || n < 1) return kmt_fail(KMT_ERR_ARG, "bad gradient arguments");
  KMT_TRY(m->upload(Xh, yh, n));
  m->fwd_bwd(0, n);
  HIP_TRY(hipGetLastError());
  if (gradh) HIP_TRY(hipMemcpyAsync(gradh, m->grad, sizeof(float) * kNParams, hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  if (loss) {
    double t = 0;
    KMT_TRY(m->loss_sum(n, &t));
    *loss = t / n;
  }
  return KMT_OK;
}

int kmt_mlp_train(kmt_mlp *m, const float *Xh, const float *yh, int n, int epochs, int batch, double lr,
                  double *loss_history) {
  if (!m || !Xh || !yh || n < 1 || epochs < 1 || batch < 1) return kmt_fail(KMT_ERR_ARG, "bad training arguments");
  KMT_TRY(m->upload(Xh, yh, n));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int ep = 0; ep < epochs; ++ep) {
    double tot = 0;
    for (int s0 = 0; s0 < n; s0 += batch) {
      const int S = std::min(batch, n - s0);
      m->fwd_bwd(s0, S);
      ++m->step;
      const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
      mlp_adam_kernel<<<(kNParams + kBlock - 1) / kBlock, kBlock, 0, m->stream>>>(m->w32, m->w16, m->m, m->v, m->grad,
                                                                                   (float)lr, b1, b2, eps, c1, c2);
      HIP_TRY(hipGetLastError());
      double part = 0;
      KMT_TRY(m->loss_sum(S, &part));
      tot += part;
    }
    if (loss_history) loss_history[ep] = tot / n;
  }
  HIP_TRY(hipStreamSynchronize(m->stream));
  return KMT_OK;
}

}  // extern "C"
