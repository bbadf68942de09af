/* ORACLE — test infrastructure only.
 *
 * A CPU restatement of the reference's hot path, templated on the working
 * precision R (float = `kuramoto`, double = `kuramoto_64`).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may build, call or
 * run anything in this directory, and only as the checker.  The product
 * (kuramoto_amd/) never links it.
 *
 * What each part follows (reference file:line) and what pins it:
 *   tableau()            dp45.cu:5-21 / dp45_64.cu:5-21; pinned against the
 *                        independently published DOPRI5 coefficients in scipy
 *                        (tests/test_oracle_tableau.py) and order conditions.
 *   stage combination    dp45.cu:34-69; FMA order (first product rounded, then
 *                        fma(k_j, a_j, acc), then fma(acc, h, y)) pinned by
 *                        the K=0 controller KAT data/5000{,_lorentz}/0/1fs.dat.
 *   error / controller   dp45.cu:79-84, :117-138; pinned by the same KAT
 *                        (t bit-exact, 12 / 43 accepted steps).
 *   interpolate          dp45.cu:70-76 (P matrix); pinned by polynomial
 *                        exactness (constant f) and scipy's P matrix.  The
 *                        build evaluates the quartic in Horner form.
 *   rhs_*                kuramoto.cu:42-55, :104-126; pinned by closed forms
 *                        (M=0, Adler N=2, symmetric-M mean-frequency
 *                        invariant) and a brute-force O(N^2) loop.  Coupling
 *                        sums accumulate in double and round once to R;
 *                        FP32 cos / sin are the float roundings of the
 *                        double functions (ocos / osin).  The reference's
 *                        cuBLAS summation order and CUDA cosf bits are not
 *                        reproducible (no CUDA here), so FP32 parity rests
 *                        on these two roundings, which the GPU follows.
 *   frequencies / ICs    kuramoto.cu:495-541; pinned bit-exactly by
 *                        data/5000{,_lorentz}/{0,2,5000}/1frequencies.dat.
 *   factor generators    the build's own Philox4x32-10 mapping (the
 *                        reference's cuRAND stream is not reproducible);
 *                        pinned by the Random123 KAT and moment checks.
 *   run / step_eval      kuramoto.cu:164-265, :543-550, :585-647.
 */
#ifndef KMT_ORACLE_HPP
#define KMT_ORACLE_HPP

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "philox_ref.h"

namespace orc {

/* ------------------------------------------------------------------ */
/* Dormand–Prince 4(5) tableau, dp45_64.cu:5-21 (dp45.cu rounds each   */
/* double expression to float).                                        */
/* ------------------------------------------------------------------ */
template <class R> struct Tableau {
  R a[6][6];   /* row s-2 holds the s-1 coefficients of stage s=2..7; row 5 = b */
  R c[6];
  R e[7];
  R p[7][4];
};

template <class R> inline Tableau<R> tableau() {
  const double A[6][6] = {
      {1.0 / 5},
      {3.0 / 40, 9.0 / 40},
      {44.0 / 45, -56.0 / 15, 32.0 / 9},
      {19372.0 / 6561, -25360.0 / 2187, 64448.0 / 6561, -212.0 / 729},
      {9017.0 / 3168, -355.0 / 33, 46732.0 / 5247, 49.0 / 176, -5103.0 / 18656},
      {35.0 / 384, 0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84}};
  const double C[6] = {0.0, 1.0 / 5, 3.0 / 10, 4.0 / 5, 8.0 / 9, 1.0};
  const double E[7] = {-71.0 / 57600, 0, 71.0 / 16695, -71.0 / 1920, 17253.0 / 339200, -22.0 / 525, 1.0 / 40};
  const double P[7][4] = {
      {1.0, -8048581381.0 / 2820520608, 8663915743.0 / 2820520608, -12715105075.0 / 11282082432},
      {0, 0, 0, 0},
      {0, 131558114200.0 / 32700410799, -68118460800.0 / 10900136933, 87487479700.0 / 32700410799},
      {0, -1754552775.0 / 470086768, 14199869525.0 / 1410260304, -10690763975.0 / 1880347072},
      {0, 127303824393.0 / 49829197408, -318862633887.0 / 49829197408, 701980252875.0 / 199316789632},
      {0, -282668133.0 / 205662961, 2019193451.0 / 616988883, -1453857185.0 / 822651844},
      {0, 40617522.0 / 29380423, -110615467.0 / 29380423, 69997945.0 / 29380423}};
  Tableau<R> T;
  std::memset(&T, 0, sizeof(T));
  for (int s = 0; s < 6; ++s)
    for (int j = 0; j <= s; ++j) T.a[s][j] = (R)A[s][j];
  for (int j = 0; j < 6; ++j) T.c[j] = (R)C[j];
  for (int j = 0; j < 7; ++j) T.e[j] = (R)E[j];
  for (int j = 0; j < 7; ++j)
    for (int q = 0; q < 4; ++q) T.p[j][q] = (R)P[j][q];
  return T;
}

inline float fmaR(float a, float b, float c) { return std::fma(a, b, c); }
inline double fmaR(double a, double b, double c) { return std::fma(a, b, c); }

/* cos / sin in the working precision (kuramoto.cu:52-53 evaluates them in
 * the working precision).  FP32: the float rounding of the double function,
 * i.e. the correctly rounded float value except within 2^-29 relative of a
 * rounding boundary, so the restatement does not depend on one float libm's
 * last bit (CUDA's cosf, glibc's cosf and the device library each differ
 * from it in some last bits).  Pinned against numpy's float64 cos / sin
 * rounded to float32 (tests/test_oracle.py). */
template <class R> inline R ocos(R x) { return (R)std::cos((double)x); }
template <class R> inline R osin(R x) { return (R)std::sin((double)x); }

/* ------------------------------------------------------------------ */
/* Host generators: frequencies and initial conditions from glibc rand */
/* (kuramoto.cu:495-541; the typing of each expression decides the     */
/* float bits and is restated exactly).                                */
/* ------------------------------------------------------------------ */
template <class R> inline R ic_draw(R I) {
  /* kuramoto.cu:499  yloc[j] = I/2*(2.0/RAND_MAX*rand()-1) */
  return (R)((double)(I / 2) * (2.0 / RAND_MAX * rand() - 1));
}

template <class R> inline R freq_draw(R freq, int normal);

template <> inline float freq_draw<float>(float freq, int normal) {
  if (normal) { /* kuramoto.cu:528-530 with float u, v */
    float u = rand() / (float)RAND_MAX;
    float v = rand() / (float)RAND_MAX;
    return (float)((double)freq * std::pow((double)(-2 * logf(u)), 0.5) * std::cos(2 * M_PI * (double)v));
  }
  float u = rand() / (float)RAND_MAX; /* kuramoto.cu:534-535 */
  return (float)((double)freq * std::tan(M_PI * ((double)u - 0.5)));
}

template <> inline double freq_draw<double>(double freq, int normal) {
  if (normal) { /* kuramoto_64.cu:525-527 */
    double u = rand() / (double)RAND_MAX;
    double v = rand() / (double)RAND_MAX;
    return freq * std::pow(-2 * std::log(u), 0.5) * std::cos(2 * M_PI * v);
  }
  double u = rand() / (double)RAND_MAX;
  return freq * std::tan(M_PI * (u - 0.5));
}

/* ------------------------------------------------------------------ */
/* Low-rank factor generators (the build's own Philox4x32-10 layout).   */
/* key = {seed, 0x4B4D5441}; counter = {idx_lo, idx_hi, stream, 0}.     */
/*  Gaussian: stream 0x58470000, one call per element, u1/u2 53-bit     */
/*  uniforms in (0,1), Box–Muller cosine branch sqrt(-2 ln u1) cos(2pi u2).*/
/*  Rademacher: stream 0x52410000, one call per 128 elements, bit b of  */
/*  the 128-bit output (word b/32, bit b%32) set => +1 else -1.         */
/* Factor layout: X[n*K + m], n < N oscillators, m < K rank.            */
/* ------------------------------------------------------------------ */
inline void gaussian_factor(long N, int K, uint32_t seed, double *X) {
  const uint64_t total = (uint64_t)N * (uint64_t)K;
  const uint32_t key[2] = {seed, 0x4B4D5441u};
  for (uint64_t e = 0; e < total; ++e) {
    uint32_t ctr[4] = {(uint32_t)e, (uint32_t)(e >> 32), 0x58470000u, 0u}, o[4];
    orc_philox4x32(ctr, key, 10, o);
    uint64_t a = ((uint64_t)o[1] << 32) | o[0];
    uint64_t b = ((uint64_t)o[3] << 32) | o[2];
    double u1 = ((double)(a >> 11) + 0.5) * 0x1.0p-53;
    double u2 = ((double)(b >> 11) + 0.5) * 0x1.0p-53;
    X[e] = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
}

inline void rademacher_factor(long N, int K, uint32_t seed, double *X) {
  const uint64_t total = (uint64_t)N * (uint64_t)K;
  const uint32_t key[2] = {seed, 0x4B4D5441u};
  for (uint64_t blk = 0; 128 * blk < total; ++blk) {
    uint32_t ctr[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), 0x52410000u, 0u}, o[4];
    orc_philox4x32(ctr, key, 10, o);
    for (int b = 0; b < 128 && 128 * blk + b < total; ++b)
      X[128 * blk + b] = ((o[b >> 5] >> (b & 31)) & 1u) ? 1.0 : -1.0;
  }
}

/* Per-rank weights w[m] so that M = X diag(w) X^T.
 *  volcano (reference, kuramoto.cu:137-157): w_m = (-1)^m / N, rounded to R
 *  as makeu2 does.
 *  gauss   (A = X X^T / K with the standard attractive sign and a c/N
 *  prefactor): w_m = -1/(N K), rounded to R. */
template <class R> inline void factor_weights(long N, int K, int gauss, R *w) {
  for (int m = 0; m < K; ++m) {
    if (gauss) w[m] = (R)(-1.0 / ((double)N * (double)K));
    else w[m] = (R)((m % 2 ? -1.0 : 1.0) / (double)N);
  }
}

/* One adjacency entry, exactly symmetric: acc = fma(w_m, X_im*X_jm, acc). */
template <class R> inline R adj_entry(const R *Xi, const R *Xj, const R *w, int K) {
  R acc = 0;
  for (int m = 0; m < K; ++m) acc = fmaR(w[m], Xi[m] * Xj[m], acc);
  return acc;
}

/* ------------------------------------------------------------------ */
/* RHS, kuramoto.cu:42-55 (tot_kuramoto, makey2), :104-126 (makecoupling, */
/* dydt).  Sign and scale follow the reference (no 1/N in the RHS):       */
/*   f_i = w_i + c1 * (sin(th_i) * C_i - cos(th_i) * S_i),                */
/*   C_i = sum_j M_ij cos(th_j),  S_i = sum_j M_ij sin(th_j).             */
/* Coupling sums accumulate in double, then round to R.                   */
/* ------------------------------------------------------------------ */
template <class R> inline R rhs_epilogue(R om, R c1, R cs, R sn, R C, R S) {
  R d = fmaR(sn, C, -(cs * S));
  return fmaR(c1, d, om);
}

struct Adjacency {
  /* exactly one of the representations is used */
  int kind = 0;          /* 0 = stored dense, 1 = dense rebuilt from factor, 2 = factored (-A) */
  long N = 0;
  int K = 0;
  const float *Mf = nullptr;   /* stored dense (float) */
  const double *Md = nullptr;  /* stored dense (double) */
  double mscale = 1.0;         /* M = mscale * stored */
  const void *X = nullptr;     /* factor, R-typed, N x K */
  const void *w = nullptr;     /* weights, R-typed, K */
};

template <class R> struct Model {
  long N = 0;
  R c1 = 0;
  std::vector<R> omega;
  Adjacency adj;
  std::vector<R> cosv, sinv;   /* makey2 scratch */
  std::vector<R> Cs;           /* interleaved couplings (C_i, S_i) of the last evaluation */

  void init(long n) {
    N = n;
    omega.assign(N, 0);
    cosv.assign(N, 0);
    sinv.assign(N, 0);
    Cs.assign(2 * N, 0);
  }

  /* makey2 (kuramoto.cu:49-55) */
  void sincos(const R *y) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
      cosv[i] = ocos<R>(y[i]);
      sinv[i] = osin<R>(y[i]);
    }
  }

  /* makecoupling (kuramoto.cu:104-116) on the current cosv/sinv */
  void coupling() {
    if (adj.kind == 2) {
      const R *X = (const R *)adj.X, *w = (const R *)adj.w;
      int K = adj.K;
      /* G_m = w_m sum_n X_nm (cos_n, sin_n), each sum in n order; a thread
       * owns a strip of m and walks the rows of X (contiguous reads) */
      std::vector<double> gc(K, 0.0), gs(K, 0.0);
#pragma omp parallel
      {
        int nt = 1, tid = 0;
#ifdef _OPENMP
        nt = omp_get_num_threads();
        tid = omp_get_thread_num();
#endif
        const int m0 = (int)((long)K * tid / nt), m1 = (int)((long)K * (tid + 1) / nt);
        for (long n = 0; n < N; ++n) {
          const double c = (double)cosv[n], sn = (double)sinv[n];
          const R *Xn = X + n * K;
          for (int m = m0; m < m1; ++m) {
            gc[m] += (double)Xn[m] * c;
            gs[m] += (double)Xn[m] * sn;
          }
        }
        for (int m = m0; m < m1; ++m) {
          gc[m] *= (double)w[m];
          gs[m] *= (double)w[m];
        }
      }
#pragma omp parallel for schedule(static)
      for (long n = 0; n < N; ++n) {
        double ac = 0, as = 0;
        for (int m = 0; m < K; ++m) {
          ac += (double)X[n * K + m] * gc[m];
          as += (double)X[n * K + m] * gs[m];
        }
        Cs[2 * n] = (R)ac;
        Cs[2 * n + 1] = (R)as;
      }
      return;
    }
#pragma omp parallel
    {
      std::vector<R> row;
      if (adj.kind == 1) row.resize(N);
#pragma omp for schedule(dynamic, 16)
      for (long i = 0; i < N; ++i) {
        double ac = 0, as = 0;
        if (adj.kind == 1) {
          const R *X = (const R *)adj.X, *w = (const R *)adj.w;
          int K = adj.K;
          for (long j = 0; j < N; ++j) row[j] = adj_entry<R>(X + i * K, X + j * K, w, K);
          row_sums(row.data(), ac, as);
        } else if (adj.Mf) {
          row_sums(adj.Mf + (size_t)i * N, ac, as);
        } else {
          row_sums(adj.Md + (size_t)i * N, ac, as);
        }
        Cs[2 * i] = (R)(adj.mscale * ac);
        Cs[2 * i + 1] = (R)(adj.mscale * as);
      }
    }
  }

  /* C = sum_j m_j cos_j, S = sum_j m_j sin_j in double: four interleaved
   * partial sums (j mod 4), combined as (p0 + p1) + (p2 + p3).  Any double
   * order rounds to the same float sums except at a rounding boundary; four
   * independent chains let the loop run at memory speed instead of the
   * latency of one chain of additions. */
  template <class T> void row_sums(const T *m, double &C, double &S) const {
    double pc[4] = {0, 0, 0, 0}, ps[4] = {0, 0, 0, 0};
    long j = 0;
    for (; j + 4 <= N; j += 4)
      for (int q = 0; q < 4; ++q) {
        pc[q] += (double)m[j + q] * (double)cosv[j + q];
        ps[q] += (double)m[j + q] * (double)sinv[j + q];
      }
    for (int q = 0; j < N; ++j, ++q) {
      pc[q] += (double)m[j] * (double)cosv[j];
      ps[q] += (double)m[j] * (double)sinv[j];
    }
    C = (pc[0] + pc[1]) + (pc[2] + pc[3]);
    S = (ps[0] + ps[1]) + (ps[2] + ps[3]);
  }

  /* dydt (kuramoto.cu:118-126) */
  void rhs(const R *y, R *f) {
    sincos(y);
    coupling();
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) f[i] = rhs_epilogue<R>(omega[i], c1, cosv[i], sinv[i], Cs[2 * i], Cs[2 * i + 1]);
  }

  /* order parameter (kuramoto.cu:216-220): r = sqrt((X/N)^2 + (Y/N)^2) with
   * X, Y the sums of cos / sin (accumulated in double, rounded to R) */
  R order(const R *y) {
    double X = 0, Y = 0;
    for (long i = 0; i < N; ++i) {
      X += (double)ocos<R>(y[i]);
      Y += (double)osin<R>(y[i]);
    }
    R XR = (R)X, YR = (R)Y, NR = (R)N;
    return (R)std::pow((double)(XR / NR * XR / NR + YR / NR * YR / NR), 0.5);
  }
};

/* ------------------------------------------------------------------ */
/* DP45 integrator, dp45.cu:34-164.                                     */
/* ------------------------------------------------------------------ */
template <class R> struct DP45 {
  long N = 0;
  R atl = 0, rtl = 0;
  int fixed = 0;
  R t_last = 0;
  Tableau<R> T = tableau<R>();
  std::vector<R> y, ylast, ytemp, yerr, k[7];
  Model<R> *model = nullptr;
  long rhs_evals = 0, rejected = 0;
  double last_norm = 0;

  void init(long n, R atol, R rtol, int fixedstep, const R *y0, Model<R> *m) {
    N = n; atl = atol; rtl = rtol; fixed = fixedstep; model = m;
    y.assign(y0, y0 + N);
    ylast.assign(N, 0); ytemp.assign(N, 0); yerr.assign(N, 0);
    for (auto &v : k) v.assign(N, 0);
  }

  void f(const R *yy, R *out) { model->rhs(yy, out); ++rhs_evals; }

  /* stage combination (step2..step7): acc = k1*a0; acc = fma(k_j, a_j, acc);
   * ytemp = fma(acc, h, y) */
  void combine(int s /*2..7*/, R h) {
    const R *a = T.a[s - 2];
    int nk = s - 1;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
      R acc = k[0][i] * a[0];
      for (int j = 1; j < nk; ++j) acc = fmaR(k[j][i], a[j], acc);
      ytemp[i] = fmaR(acc, h, y[i]);
    }
  }

  /* dp45_step (dp45.cu:87-141) */
  int step(R *t, R *h) {
    for (int s = 2; s <= 7; ++s) {
      combine(s, *h);
      f(ytemp.data(), k[s - 1].data());
    }
    if (fixed) {
      t_last = *t;
      ylast = y;
      y = ytemp;
      *t = *t + *h;
      return 1;
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
      R acc = k[0][i] * T.e[0];
      for (int j = 1; j < 7; ++j) acc = fmaR(k[j][i], T.e[j], acc);
      R num = *h * acc;
      R den = fmaR(rtl, std::fmax(std::fabs(y[i]), std::fabs(ytemp[i])), atl);
      yerr[i] = num / den;
    }
    double sum = 0; /* sequential, so the norm does not depend on the thread count */
    for (long i = 0; i < N; ++i) sum += (double)yerr[i] * (double)yerr[i];
    R nrm = (R)std::sqrt(sum);                                     /* {S,D}nrm2 */
    R norm = (R)((double)nrm / std::pow((double)N, 0.5));          /* dp45.cu:119 */
    R factor = (R)(0.9 * std::pow((double)norm, -0.2));            /* dp45.cu:120 */
    last_norm = (double)norm;
    if (norm < 1) {
      t_last = *t;
      ylast = y;
      y = ytemp;
      *t = *t + *h;
      if (factor > 10) factor = 10;
      *h *= factor;
      return 1;
    } else if (factor < 1) {
      if ((double)factor < 0.2) factor = (R)0.2;
      *h *= factor;
    }
    ++rejected;
    if (std::isnan((double)norm)) {
      std::fprintf(stderr, "oracle: error norm is NaN\n");
      std::exit(3);
    }
    return 0;
  }

  /* dp45_eval / interpolate (dp45.cu:70-76, :143-146), Horner form:
   * Q_q = sum_j p_j[q] k_j;  y = ylast + (h x) * (Q0 + x (Q1 + x (Q2 + x Q3))) */
  void eval(R t, R t_eval, R *out) {
    R hh = t - t_last, h2 = t_eval - t_last;
    R x = h2 / hh;
    R hx = hh * x;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
      R Q[4];
      for (int q = 0; q < 4; ++q) {
        R acc = k[0][i] * T.p[0][q];
        for (int j = 1; j < 7; ++j) acc = fmaR(k[j][i], T.p[j][q], acc);
        Q[q] = acc;
      }
      R in = fmaR(x, Q[3], Q[2]);
      in = fmaR(x, in, Q[1]);
      in = fmaR(x, in, Q[0]);
      out[i] = fmaR(hx, in, ylast[i]);
    }
  }
};

}  // namespace orc
#endif
