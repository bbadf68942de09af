/* ORACLE — test infrastructure only (see oracle.hpp header).
 *
 * Command-line restatement of the reference programs `kuramoto` /
 * `kuramoto_64`: main() kuramoto.cu:267-678 (kuramoto_64.cu:264-675) and
 * step_eval() kuramoto.cu:164-265 (kuramoto_64.cu:164-261), same getopt
 * string, defaults, reload rules, output grid and FILEBASE files.  Built once
 * per precision (-DKMT_REAL=float -> kuramoto_oracle, double ->
 * kuramoto_oracle_64).  Extensions (long options, shared with the product
 * CLI):
 *   --gauss        adjacency M = -X X^T/(N K) with Gaussian X (standard
 *                  attractive sign, c/N prefactor) instead of the volcano
 *                  M = sum_m (-1)^m u_m u_m^T / N with Rademacher u.
 *   --rebuild      oracle only: rebuild dense rows of M from the factor at
 *                  every RHS instead of storing N^2 values.
 *   --max-steps n  stop after n accepted steps and leave the files as a
 *                  process killed at that point would (no final block).
 *   --exact-final  final fs.dat holds the integrator's state (y at its t,
 *                  t, h, couplings of y) instead of the reference's
 *                  interpolation with k1 already replaced by k7.
 * A summary line "oracle: ..." with loop time and osc*steps/s goes to stdout.
 */
#include <getopt.h>
#include <sys/time.h>
#include <unistd.h>

#include <string>

#include "oracle.hpp"

#ifndef KMT_REAL
#define KMT_REAL double
#endif
typedef KMT_REAL R;

using namespace orc;

static double now_s() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + 1e-6 * tv.tv_usec;
}

struct Ctx {
  long N, K;
  R t0, t1;
  int steps = 0, verbose = 0, dense = 3;
  std::vector<R> t_eval;
  int n_eval = 0, eval_i = 0;
  std::vector<R> yloc, floc;
  std::string filebase;
  double start;
  Model<R> *model;
  DP45<R> *dp;
  std::vector<R> yeval;
  int max_steps = -1;
};

static std::string fname(const Ctx &p, const char *suffix) { return p.filebase + suffix; }

/* step_eval, kuramoto.cu:164-265 */
static void step_eval(R t, R h, Ctx &p) {
  p.steps++;
  if (p.verbose) {
    double el = now_s() - p.start;
    double frac = (double)((t - p.t0) / (p.t1 - p.t0));
    double eta = el / ((double)((t - p.t0 + h) / (p.t1 - p.t0))) * (1 - frac);
    std::printf("%.3f\t%1.3e\t%1.3e\t%f\t%i\t\r", frac, el, eta, (double)h, p.steps);
    std::fflush(stdout);
    FILE *out = std::fopen(fname(p, ".out").c_str(), "ab");
    std::fprintf(out, "%.3f\t%1.3e\t%1.3e\t%f\t%i\t\n", frac, el, eta, (double)h, p.steps);
    std::fclose(out);
  }
  if (p.dense >= 1) {
    FILE *ft = std::fopen(fname(p, "times.dat").c_str(), "ab");
    std::fwrite(&t, sizeof(R), 1, ft);
    std::fclose(ft);
  }
  FILE *fth = nullptr, *fcp = nullptr;
  if (p.dense >= 1) fth = std::fopen(fname(p, "thetas.dat").c_str(), "ab");
  if (p.dense >= 2) fcp = std::fopen(fname(p, "couplings.dat").c_str(), "ab");
  std::vector<R> r;
  while (p.eval_i < p.n_eval && t >= p.t_eval[p.eval_i]) {
    p.dp->eval(t, p.t_eval[p.eval_i], p.yeval.data());
    r.push_back(p.model->order(p.yeval.data()));
    if (p.dense >= 1) std::fwrite(p.yeval.data(), sizeof(R), p.N, fth);
    if (p.dense >= 2) {
      p.model->sincos(p.yeval.data());
      p.model->coupling();
      p.floc = p.model->Cs;
      std::fwrite(p.floc.data(), sizeof(R), 2 * p.N, fcp);
    }
    p.eval_i++;
  }
  if (fth) std::fclose(fth);
  if (fcp) std::fclose(fcp);
  if (!r.empty()) {
    FILE *fo = std::fopen(fname(p, "order.dat").c_str(), "ab");
    std::fwrite(r.data(), sizeof(R), r.size(), fo);
    std::fclose(fo);
  }
  /* kuramoto.cu:251 guards the per-step fs.dat with dense>=0;
   * kuramoto_64.cu:250 writes it unconditionally */
  if (sizeof(R) == 8 || p.dense >= 0) {
    FILE *fl = std::fopen(fname(p, "fs.dat").c_str(), "wb");
    std::fwrite(p.dp->y.data(), sizeof(R), p.N, fl);
    std::fwrite(&t, sizeof(R), 1, fl);
    std::fwrite(&h, sizeof(R), 1, fl);
    std::fwrite(p.floc.data(), sizeof(R), 2 * p.N, fl);
    std::fclose(fl);
  }
}

static void usage() {
  std::printf("usage:\tkuramoto [-hvnRFA] [-N N] [-K K] [-D D]\n");
  std::printf("\t[-c c] [-t t] [-d dt] [-f f] [-s seed] \n");
  std::printf("\t[-I init] [-r rtol] [-a atol] [-g gpu] filebase \n\n");
  std::printf("-h for help \n-v for verbose \n-n for normal random frequencies (default is cauchy) \n");
  std::printf("-R to reload adjacency, frequencies, and initial conditions from files if possible\n");
  std::printf("-F for fixed timestep \n-A to regenerate volcano adjacency each step to save memory \n");
  std::printf("D is the output density level. 0 for minimal, 1 for phases, 2 for phases and couplings, 3 for phases, couplings, and adjacency\n");
  std::printf("N is number of oscillators. Default 5000. \nK is rank of volcano adjacency. Default 5. \n");
  std::printf("c is the coupling coefficient. Default 3.0. \nt is total integration time. Default 1e2. \n");
  std::printf("dt is the time between outputs. Default 1e0. \nf is the scale of the frequencies. Default 1e0. \n");
  std::printf("seed is random seed. Default 1. \ninit is uniform random initial condition scale. Default 0. \n");
  std::printf("rtol is relative error tolerance. Default 0.\natol is absolute error tolerance. Default 1e-6.\n");
  std::printf("gpu is index of the gpu. Default 0.\nfilebase is base file name for output. \n");
  std::printf("oracle extensions: --gauss --rebuild --max-steps n --exact-final\n");
}

int main(int argc, char *argv[]) {
  double start = now_s();
  R c1 = 3.0, t1 = 1e2, dt = 1e0, freq = 1.0, atl = 1e-6, rtl = 0, I = 0;
  int seed = 1, fixed = 0, A = 0, verbose = 0, help = 1, dense = 3, normal = 0, reload = 0;
  int gauss = 0, rebuild = 0, max_steps = -1, exact_final = 0;
  long N = 5000, K = 6;
  const char *filebase = nullptr;
  static struct option longopts[] = {{"gauss", no_argument, nullptr, 1001},
                                     {"rebuild", no_argument, nullptr, 1002},
                                     {"max-steps", required_argument, nullptr, 1003},
                                     {"exact-final", no_argument, nullptr, 1004},
                                     {nullptr, 0, nullptr, 0}};
  while (optind < argc) {
    int c = getopt_long(argc, argv, "N:K:I:D:c:g:t:d:f:s:r:a:hvFnRA", longopts, nullptr);
    if (c != -1) {
      switch (c) {
        case 'N': N = (int)atoi(optarg); break;
        case 'K': K = (int)atoi(optarg); break;
        case 'A': A = 1; break;
        case 'I': I = (R)atof(optarg); break;
        case 'c': c1 = (R)atof(optarg); break;
        case 'g': break;
        case 't': t1 = (R)atof(optarg); break;
        case 'd': dt = (R)atof(optarg); break;
        case 'f': freq = (R)atof(optarg); break;
        case 's': seed = (int)atoi(optarg); break;
        case 'r': rtl = (R)atof(optarg); break;
        case 'a': atl = (R)atof(optarg); break;
        case 'D': dense = (int)atoi(optarg); break;
        case 'R': reload = 1; break;
        case 'F': fixed = 1; break;
        case 'h': help = 1; break;
        case 'v': verbose = 1; break;
        case 'n': normal = 1; break;
        case 1001: gauss = 1; break;
        case 1002: rebuild = 1; break;
        case 1003: max_steps = atoi(optarg); break;
        case 1004: exact_final = 1; break;
      }
    } else {
      filebase = argv[optind];
      optind++;
      help = 0;
    }
  }
  if (help || !filebase) {
    usage();
    return 0;
  }

  R t = 0, h;
  Ctx p;
  p.filebase = filebase;
  FILE *out = std::fopen(fname(p, ".out").c_str(), "ab");
  std::vector<R> yloc(N, 0), omegas(N, 0);
  p.floc.assign(2 * N, 0);
  srand(seed);
  std::fprintf(out, "%lu %lu %f %f %f %i\n", (unsigned long)N, (unsigned long)K, (double)t1, (double)dt, (double)c1, seed);
  for (int i = 0; i < argc; i++) std::fprintf(out, "%s ", argv[i]);
  std::fprintf(out, "\n");
  size_t req = A ? 100 * N * sizeof(R) : (100 + N + 2 * K) * N * sizeof(R);
  std::printf("Host Memory: req %zu\n", req);
  std::fprintf(out, "Host Memory: req %zu\n", req);

  h = fixed ? dt : dt / 100;
  int reloaded = 0;
  FILE *in;
  if (reload && (in = std::fopen(fname(p, "fs.dat").c_str(), "r"))) {
    reloaded = 1;
    std::printf("Using initial conditions from file\n");
    std::fprintf(out, "Using initial conditions from file\n");
    size_t rd = std::fread(yloc.data(), sizeof(R), N, in);
    if (rd != (size_t)N) {
      std::printf("initial conditions file not compatible with N!\n");
      std::fprintf(out, "initial conditions file not compatible with N!\n");
      reloaded = 0;
    }
    if (reloaded) {
      rd = std::fread(&t, sizeof(R), 1, in);
      rd = std::fread(&h, sizeof(R), 1, in);
      if (rd != 1) {
        std::printf("Couldn't read start time and step!\n");
        std::fprintf(out, "Couldn't read start time and step!\n");
        reloaded = 0;
      }
    }
    std::fclose(in);
    std::printf("Restarting at t=%f with h=%f\n", (double)t, (double)h);
    std::fprintf(out, "Restarting at t=%f with h=%f\n", (double)t, (double)h);
  }
  if (!reloaded) {
    std::printf("Using random initial conditions\n");
    std::fprintf(out, "Using random initial conditions\n");
    for (long j = 0; j < N; j++) yloc[j] = ic_draw<R>(I);
    in = std::fopen(fname(p, "fs.dat").c_str(), "wb");
    std::fwrite(yloc.data(), sizeof(R), N, in);
    std::fclose(in);
  }
  if (reload && (in = std::fopen(fname(p, "frequencies.dat").c_str(), "r"))) {
    std::printf("Using frequencies from file\n");
    std::fprintf(out, "Using frequencies from file\n");
    size_t rd = std::fread(omegas.data(), sizeof(R), N, in);
    std::fclose(in);
    if (rd != (size_t)N) {
      std::printf("frequencies not compatible with N!");
      std::fprintf(out, "frequencies not compatible with N!");
      return 0;
    }
  } else {
    std::printf("Using random frequencies\n");
    std::fprintf(out, "Using random frequencies\n");
    for (long j = 0; j < N; j++) omegas[j] = freq_draw<R>(freq, normal);
    in = std::fopen(fname(p, "frequencies.dat").c_str(), "wb");
    std::fwrite(omegas.data(), sizeof(R), N, in);
    std::fclose(in);
  }
  int n0 = int(t / dt) + 1;
  int n1 = int(t1 / dt) + 1;
  int n_eval = n1 - n0;
  if (n_eval < 0) n_eval = 0;
  p.t_eval.resize(n_eval);
  for (int n = n0, ind = 0; n < n1; n++) p.t_eval[ind++] = dt * n;

  Model<R> model;
  model.init(N);
  model.c1 = c1;
  model.omega = omegas;
  /* factor and weights (the adjacency generator) */
  std::vector<double> Xd(N * K);
  if (gauss) gaussian_factor(N, (int)K, (uint32_t)seed, Xd.data());
  else rademacher_factor(N, (int)K, (uint32_t)seed, Xd.data());
  std::vector<R> X(Xd.begin(), Xd.end()), w(K);
  factor_weights<R>(N, (int)K, gauss, w.data());
  std::vector<R> adjv;
  model.adj.N = N;
  model.adj.K = (int)K;
  model.adj.X = X.data();
  model.adj.w = w.data();
  if (!A && reload && (in = std::fopen(fname(p, "adj.dat").c_str(), "r"))) {
    std::printf("Using adjacency from file\n");
    std::fprintf(out, "Using adjacency from file\n");
    adjv.assign((size_t)N * N, 0);
    size_t rd = std::fread(adjv.data(), sizeof(R), (size_t)N * N, in);
    (void)rd;
    std::fclose(in);
    model.adj.kind = 0;
  } else {
    std::printf("Using random adjacency matrix\n");
    std::fprintf(out, "Using random adjacency matrix\n");
    if (A) {
      model.adj.kind = 2;
    } else if (rebuild) {
      model.adj.kind = 1;
    } else {
      model.adj.kind = 0;
      adjv.assign((size_t)N * N, 0);
#pragma omp parallel for schedule(static)
      for (long i = 0; i < N; ++i)
        for (long j = 0; j < N; ++j) adjv[(size_t)i * N + j] = adj_entry<R>(&X[i * K], &X[j * K], w.data(), (int)K);
      if (dense >= 3) {
        in = std::fopen(fname(p, "adj.dat").c_str(), "wb");
        std::fwrite(adjv.data(), sizeof(R), (size_t)N * N, in);
        std::fclose(in);
      }
    }
  }
  if (model.adj.kind == 0) {
    if (sizeof(R) == 4) model.adj.Mf = (const float *)adjv.data();
    else model.adj.Md = (const double *)adjv.data();
  }
  std::fclose(out);

  DP45<R> dp;
  dp.init(N, atl, rtl, fixed, yloc.data(), &model);
  p.N = N; p.K = K; p.t0 = t; p.t1 = t1; p.verbose = verbose; p.dense = dense;
  p.n_eval = n_eval; p.start = start; p.model = &model; p.dp = &dp;
  p.yeval.assign(N, 0);
  p.max_steps = max_steps;

  if (!reloaded) {
    R r = model.order(dp.y.data());
    FILE *fo = std::fopen(fname(p, "order.dat").c_str(), "wb");
    std::fwrite(&r, sizeof(R), 1, fo);
    std::fclose(fo);
    if (dense >= 1) {
      FILE *fa = std::fopen(fname(p, "thetas.dat").c_str(), "wb");
      std::fwrite(yloc.data(), sizeof(R), N, fa);
      std::fclose(fa);
      FILE *ft = std::fopen(fname(p, "times.dat").c_str(), "wb");
      std::fclose(ft);
    }
    if (dense >= 2) {
      model.sincos(dp.y.data());
      model.coupling();
      p.floc = model.Cs;
      FILE *fc = std::fopen(fname(p, "couplings.dat").c_str(), "wb");
      std::fwrite(p.floc.data(), sizeof(R), 2 * N, fc);
      std::fclose(fc);
    }
  }

  /* dp45_run, dp45.cu:148-164 */
  double loop0 = now_s();
  bool stopped = false;
  dp.f(dp.y.data(), dp.k[0].data());
  while (t < t1) {
    int ok = dp.step(&t, &h);
    if (ok) {
      step_eval(t, h, p);
      dp.k[0] = dp.k[6];
      if (max_steps > 0 && p.steps >= max_steps) {
        stopped = true;
        break;
      }
    }
  }
  double loop_s = now_s() - loop0;
  int threads = 1;
#ifdef _OPENMP
  threads = omp_get_max_threads();
#endif
  if (stopped) {  // leave the files as a process killed after this step would
    std::printf("oracle: stopped accepted=%d rejected=%ld rhs=%ld loop_seconds=%.6f osc_steps_per_s=%.6e threads=%d t=%.9g h=%.9g\n",
                p.steps, dp.rejected, dp.rhs_evals, loop_s, (double)N * p.steps / loop_s, threads, (double)t, (double)h);
    return 0;
  }

  /* final block, kuramoto.cu:630-647: interpolation uses k1 already replaced
   * by k7 and the integrator t */
  if (n_eval > 0 && !exact_final) dp.eval(t, p.t_eval[n_eval - 1], p.yeval.data());
  else p.yeval = dp.y;
  model.sincos(p.yeval.data());
  model.coupling();
  p.floc = model.Cs;
  FILE *fl = std::fopen(fname(p, "fs.dat").c_str(), "wb");
  std::fwrite(p.yeval.data(), sizeof(R), N, fl);
  std::fwrite(&t, sizeof(R), 1, fl);
  std::fwrite(&h, sizeof(R), 1, fl);
  std::fwrite(p.floc.data(), sizeof(R), 2 * N, fl);
  std::fclose(fl);

  out = std::fopen(fname(p, ".out").c_str(), "ab");
  double rt = now_s() - start;
  std::printf("\nruntime: %f\n", rt);
  std::fprintf(out, "\nsteps: %i\n", p.steps);
  std::fprintf(out, "runtime: %f\n", rt);
  std::fclose(out);
  std::printf("oracle: accepted=%d rejected=%ld rhs=%ld loop_seconds=%.6f osc_steps_per_s=%.6e threads=%d t=%.9g h=%.9g\n",
              p.steps, dp.rejected, dp.rhs_evals, loop_s, (double)N * p.steps / loop_s, threads, (double)t, (double)h);
  return 0;
}
