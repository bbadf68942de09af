/* ORACLE — test infrastructure only (see oracle.hpp header).
 *
 * C entry points onto the oracle for the Python tests (ctypes) and the
 * CPU-baseline leg of bench.py.  Nothing in the product links this library.
 */
#include "oracle.hpp"
#include "mlp_ref.hpp"

using namespace orc;

extern "C" {

void orc_philox(const uint32_t ctr[4], const uint32_t key[2], int rounds, uint32_t out[4]) {
  orc_philox4x32(ctr, key, rounds, out);
}

void orc_gaussian_factor(long N, int K, uint32_t seed, double *X) { gaussian_factor(N, K, seed, X); }
void orc_rademacher_factor(long N, int K, uint32_t seed, double *X) { rademacher_factor(N, K, seed, X); }

void orc_factor_weights_f64(long N, int K, int gauss, double *w) { factor_weights<double>(N, K, gauss, w); }
void orc_factor_weights_f32(long N, int K, int gauss, float *w) { factor_weights<float>(N, K, gauss, w); }

/* Rows [i0, i1) of M built from the factor (R-typed X, w), row-major N wide. */
void orc_adj_rows_f64(long N, int K, const double *X, const double *w, long i0, long i1, double *out) {
#pragma omp parallel for schedule(static)
  for (long i = i0; i < i1; ++i)
    for (long j = 0; j < N; ++j) out[(size_t)(i - i0) * N + j] = adj_entry<double>(X + i * K, X + j * K, w, K);
}
void orc_adj_rows_f32(long N, int K, const float *X, const float *w, long i0, long i1, float *out) {
#pragma omp parallel for schedule(static)
  for (long i = i0; i < i1; ++i)
    for (long j = 0; j < N; ++j) out[(size_t)(i - i0) * N + j] = adj_entry<float>(X + i * K, X + j * K, w, K);
}

/* The draw order of main(): srand(seed); N initial-condition draws; N
 * frequency draws (kuramoto.cu:416, :498-500, :526-536). */
void orc_draws_f32(long N, int seed, int normal, float freq, float I, float *ic, float *omega) {
  srand(seed);
  for (long j = 0; j < N; ++j) ic[j] = ic_draw<float>(I);
  for (long j = 0; j < N; ++j) omega[j] = freq_draw<float>(freq, normal);
}
void orc_draws_f64(long N, int seed, int normal, double freq, double I, double *ic, double *omega) {
  srand(seed);
  for (long j = 0; j < N; ++j) ic[j] = ic_draw<double>(I);
  for (long j = 0; j < N; ++j) omega[j] = freq_draw<double>(freq, normal);
}

/* The FP32 / FP64 cos and sin the model uses (ocos / osin). */
void orc_sincos_f32(long N, const float *x, float *c, float *s) {
  for (long i = 0; i < N; ++i) { c[i] = ocos<float>(x[i]); s[i] = osin<float>(x[i]); }
}
void orc_sincos_f64(long N, const double *x, double *c, double *s) {
  for (long i = 0; i < N; ++i) { c[i] = ocos<double>(x[i]); s[i] = osin<double>(x[i]); }
}

void orc_tableau_f64(double *a36, double *c6, double *e7, double *p28) {
  Tableau<double> T = tableau<double>();
  std::memcpy(a36, T.a, sizeof(T.a));
  std::memcpy(c6, T.c, sizeof(T.c));
  std::memcpy(e7, T.e, sizeof(T.e));
  std::memcpy(p28, T.p, sizeof(T.p));
}
void orc_tableau_f32(float *a36, float *c6, float *e7, float *p28) {
  Tableau<float> T = tableau<float>();
  std::memcpy(a36, T.a, sizeof(T.a));
  std::memcpy(c6, T.c, sizeof(T.c));
  std::memcpy(e7, T.e, sizeof(T.e));
  std::memcpy(p28, T.p, sizeof(T.p));
}

}  // extern "C"

/* adj_kind: 0 stored dense (Mf/Md with mscale), 1 dense rebuilt from factor,
 * 2 factored (-A). */
template <class R>
static void setup_model(Model<R> &m, long N, int adj_kind, const void *M, double mscale, const R *X, const R *w, int K,
                        const R *omega, R c1) {
  m.init(N);
  m.c1 = c1;
  m.omega.assign(omega, omega + N);
  m.adj.kind = adj_kind;
  m.adj.N = N;
  m.adj.K = K;
  m.adj.X = X;
  m.adj.w = w;
  m.adj.mscale = mscale;
  if (adj_kind == 0) {
    if (sizeof(R) == 4) m.adj.Mf = (const float *)M;
    else m.adj.Md = (const double *)M;
  }
}

template <class R>
static void rhs_t(long N, int adj_kind, const void *M, double mscale, const R *X, const R *w, int K, const R *omega, R c1,
                  const R *theta, R *f, R *Cs) {
  Model<R> m;
  setup_model<R>(m, N, adj_kind, M, mscale, X, w, K, omega, c1);
  m.rhs(theta, f);
  if (Cs) std::memcpy(Cs, m.Cs.data(), sizeof(R) * 2 * N);
}

/* In-memory restatement of dp45_run + step_eval's output grid walk.
 *  y      in: initial phases; out: phases at the last accepted step
 *  t, h   in/out (reference semantics: no clamp of the final step)
 *  theta_out  n_eval x N interpolated phases at t_eval (or null)
 *  r_out      order parameter at t_eval (or null)
 *  times_out  accepted t per step, up to times_cap (or null)
 *  counts     [accepted, rejected, rhs_evals, outputs_written]
 *  final_eval the final fs.dat phases (interpolation with k1 <- k7) or null */
template <class R>
static int integrate_t(long N, int adj_kind, const void *M, double mscale, const R *X, const R *w, int K, const R *omega,
                       R c1, R *y, R *t, R *h, R t1, R atol, R rtol, int fixed, const R *t_eval, int n_eval,
                       R *theta_out, R *r_out, R *times_out, int times_cap, int max_steps, long *counts, R *final_eval,
                       double *last_norms, int norms_cap) {
  Model<R> m;
  setup_model<R>(m, N, adj_kind, M, mscale, X, w, K, omega, c1);
  DP45<R> dp;
  dp.init(N, atol, rtol, fixed, y, &m);
  std::vector<R> ye(N);
  int eval_i = 0, steps = 0, attempts = 0;
  dp.f(dp.y.data(), dp.k[0].data());
  while (*t < t1) {
    int ok = dp.step(t, h);
    if (last_norms && attempts < norms_cap) last_norms[attempts] = dp.last_norm;
    ++attempts;
    if (ok) {
      if (times_out && steps < times_cap) times_out[steps] = *t;
      steps++;
      while (eval_i < n_eval && *t >= t_eval[eval_i]) {
        dp.eval(*t, t_eval[eval_i], ye.data());
        if (theta_out) std::memcpy(theta_out + (size_t)eval_i * N, ye.data(), sizeof(R) * N);
        if (r_out) r_out[eval_i] = m.order(ye.data());
        eval_i++;
      }
      dp.k[0] = dp.k[6];
      if (max_steps > 0 && steps >= max_steps) break;
    }
  }
  if (final_eval) {
    if (n_eval > 0) dp.eval(*t, t_eval[n_eval - 1], final_eval);
    else std::memcpy(final_eval, dp.y.data(), sizeof(R) * N);
  }
  std::memcpy(y, dp.y.data(), sizeof(R) * N);
  counts[0] = steps;
  counts[1] = dp.rejected;
  counts[2] = dp.rhs_evals;
  counts[3] = eval_i;
  return 0;
}

template <class R> static void interp_t(long N, const R *ylast, const R *k /*7*N*/, R t_last, R t, R t_eval, R *out) {
  DP45<R> dp;
  std::vector<R> z(N, 0);
  dp.init(N, 0, 0, 0, z.data(), nullptr);
  dp.ylast.assign(ylast, ylast + N);
  for (int j = 0; j < 7; ++j) dp.k[j].assign(k + (size_t)j * N, k + (size_t)(j + 1) * N);
  dp.t_last = t_last;
  dp.eval(t, t_eval, out);
}

extern "C" {

void orc_rhs_f64(long N, int adj_kind, const double *M, double mscale, const double *X, const double *w, int K,
                 const double *omega, double c1, const double *theta, double *f, double *Cs) {
  rhs_t<double>(N, adj_kind, M, mscale, X, w, K, omega, c1, theta, f, Cs);
}
void orc_rhs_f32(long N, int adj_kind, const float *M, double mscale, const float *X, const float *w, int K,
                 const float *omega, float c1, const float *theta, float *f, float *Cs) {
  rhs_t<float>(N, adj_kind, M, mscale, X, w, K, omega, c1, theta, f, Cs);
}

int orc_integrate_f64(long N, int adj_kind, const double *M, double mscale, const double *X, const double *w, int K,
                      const double *omega, double c1, double *y, double *t, double *h, double t1, double atol,
                      double rtol, int fixed, const double *t_eval, int n_eval, double *theta_out, double *r_out,
                      double *times_out, int times_cap, int max_steps, long *counts, double *final_eval,
                      double *norms, int norms_cap) {
  return integrate_t<double>(N, adj_kind, M, mscale, X, w, K, omega, c1, y, t, h, t1, atol, rtol, fixed, t_eval,
                             n_eval, theta_out, r_out, times_out, times_cap, max_steps, counts, final_eval, norms,
                             norms_cap);
}
int orc_integrate_f32(long N, int adj_kind, const float *M, double mscale, const float *X, const float *w, int K,
                      const float *omega, float c1, float *y, float *t, float *h, float t1, float atol, float rtol,
                      int fixed, const float *t_eval, int n_eval, float *theta_out, float *r_out, float *times_out,
                      int times_cap, int max_steps, long *counts, float *final_eval, double *norms, int norms_cap) {
  return integrate_t<float>(N, adj_kind, M, mscale, X, w, K, omega, c1, y, t, h, t1, atol, rtol, fixed, t_eval,
                            n_eval, theta_out, r_out, times_out, times_cap, max_steps, counts, final_eval, norms,
                            norms_cap);
}

void orc_interp_f64(long N, const double *ylast, const double *k, double t_last, double t, double t_eval, double *out) {
  interp_t<double>(N, ylast, k, t_last, t, t_eval, out);
}
void orc_interp_f32(long N, const float *ylast, const float *k, float t_last, float t, float t_eval, float *out) {
  interp_t<float>(N, ylast, k, t_last, t, t_eval, out);
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"

extern "C" {
double orc_mlp_loss_grad(const double *w, const double *X, const double *y, int n, double *grad, double *out) {
  return orc::mlp_loss_grad(w, X, y, n, grad, out);
}
int orc_mlp_param_count(void) { return orc::MP; }
}
