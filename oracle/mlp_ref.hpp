/* ORACLE — test infrastructure only.
 *
 * FP64 restatement of the step-size proposer network (8 -> 64 -> 64 -> 1,
 * tanh, tanh, linear; parameters packed W1[64][8], b1[64], W2[64][64],
 * b2[64], W3[64], b3) with its mean-squared-error loss and the gradient by
 * back-propagation.  The network has no reference counterpart; the gradient
 * is pinned by central finite differences and by the closed form of the
 * zero network (tests/test_oracle.py).
 */
#ifndef KMT_ORACLE_MLP_REF_HPP
#define KMT_ORACLE_MLP_REF_HPP
#include <cmath>
#include <vector>

namespace orc {

constexpr int MI = 8, MH = 64;
constexpr int OW1 = 0, OB1 = OW1 + MH * MI, OW2 = OB1 + MH, OB2 = OW2 + MH * MH, OW3 = OB2 + MH, OB3 = OW3 + MH,
              MP = OB3 + 1;

/* loss = (1/n) sum_s (out_s - y_s)^2; grad (MP) accumulated; out (n) */
inline double mlp_loss_grad(const double *w, const double *X, const double *y, int n, double *grad, double *out) {
  std::vector<double> h1(MH), h2(MH), d1(MH), d2(MH);
  double loss = 0;
  if (grad)
    for (int p = 0; p < MP; ++p) grad[p] = 0;
  for (int s = 0; s < n; ++s) {
    const double *x = X + (size_t)s * MI;
    for (int i = 0; i < MH; ++i) {
      double a = w[OB1 + i];
      for (int j = 0; j < MI; ++j) a += w[OW1 + i * MI + j] * x[j];
      h1[i] = std::tanh(a);
    }
    double o = w[OB3];
    for (int i = 0; i < MH; ++i) {
      double a = w[OB2 + i];
      for (int j = 0; j < MH; ++j) a += w[OW2 + i * MH + j] * h1[j];
      h2[i] = std::tanh(a);
      o += w[OW3 + i] * h2[i];
    }
    if (out) out[s] = o;
    const double r = o - y[s];
    loss += r * r;
    if (!grad) continue;
    const double d3 = 2.0 * r / n;
    grad[OB3] += d3;
    for (int i = 0; i < MH; ++i) {
      grad[OW3 + i] += d3 * h2[i];
      d2[i] = d3 * w[OW3 + i] * (1 - h2[i] * h2[i]);
      grad[OB2 + i] += d2[i];
      for (int j = 0; j < MH; ++j) grad[OW2 + i * MH + j] += d2[i] * h1[j];
    }
    for (int j = 0; j < MH; ++j) {
      double g = 0;
      for (int i = 0; i < MH; ++i) g += w[OW2 + i * MH + j] * d2[i];
      d1[j] = g * (1 - h1[j] * h1[j]);
      grad[OB1 + j] += d1[j];
      for (int k = 0; k < MI; ++k) grad[OW1 + j * MI + k] += d1[j] * x[k];
    }
  }
  return loss / n;
}

}  // namespace orc
#endif
