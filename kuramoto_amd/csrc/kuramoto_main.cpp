// `kuramoto` / `kuramoto_64`: the drop-in command-line programs.
//
// Same getopt string, defaults, reload rules, output grid, messages and
// FILEBASE files as the reference main() (kuramoto.cu:267-678,
// kuramoto_64.cu:264-675) and step_eval() (kuramoto.cu:164-265).  All GPU
// work goes through the C ABI of include/kuramoto_amd.h.  Built once per
// working precision: -DKMT_REAL=float -> kuramoto, double -> kuramoto_64.
//
// Behaviour that differs from the reference, on purpose:
//   * the step controller runs on the device; the host drains accepted-step
//     records and dense-output rows in chunks, so FILEBASEfs.dat (the restart
//     file) is rewritten after every chunk instead of after every accepted
//     step — it always holds a consistent accepted state (phases, t, h);
//   * every HIP failure is reported and the program exits with status 1
//     (the reference ignores CUDA/cuBLAS status codes);
//   * a non-finite error norm stops the run with status 2 (the reference
//     loops forever, dp45.cu:122-138).
// Extensions (long options only, so every reference command line keeps its
// meaning):
//   --gauss             adjacency M = -X X^T/(N K) with Gaussian X (the
//                       standard attractive Kuramoto coupling with a c/N
//                       prefactor) instead of the volcano matrix
//   --adj-dtype=T       HBM storage of the dense adjacency: native (default),
//                       bf16 or f16 (FP32 accumulation of the products;
//                       kuramoto_64 keeps the state, stage combinations and
//                       error norm in FP64), or sym, sym-bf16, sym-f16 (only
//                       the upper block triangle, in the working precision
//                       or in 16 bits: half the bytes per RHS; a reloaded
//                       adj.dat must be symmetric)
//   --rebuild           with -A: rebuild the adjacency tiles from the factor in
//                       registers at every stage (dense semantics, O(N^2 K)
//                       work, no N^2 storage) instead of the O(N K) factored
//                       contraction
//   --max-steps=n       stop after n accepted steps, leaving the files as a
//                       process killed at that point would (no final block)
//   --chunk=n           maximum step attempts per host synchronisation
//   --exact-final       final fs.dat holds the integrator's state (theta at
//                       its t, that t and h, the couplings of that theta),
//                       as the per-step restart file does, instead of the
//                       reference's interpolation with k1 already replaced
//                       by k7 beside the integrator's t (kuramoto.cu:630-647,
//                       dp45.cu:158-161): -R then continues the run exactly
#include <getopt.h>
#include <sys/time.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "kuramoto_amd.h"

#ifndef KMT_REAL
#define KMT_REAL float
#endif
typedef KMT_REAL R;
static const int PREC = sizeof(R) == 4 ? KMT_F32 : KMT_F64;

static double now_s() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + 1e-6 * tv.tv_usec;
}

static std::string base;
static std::string fname(const char *suffix) { return base + suffix; }

static kmt_ctx *g_ctx = nullptr;

static void die(const char *what, int rc) {
  std::fprintf(stderr, "kuramoto: %s failed (%d): %s\n", what, rc, kmt_last_error());
  if (g_ctx) kmt_destroy(g_ctx);
  std::exit(rc == KMT_ERR_NUMERIC ? 2 : 1);
}
#define CALL(x)                       \
  do {                                \
    int rc_ = (x);                    \
    if (rc_ != KMT_OK) die(#x, rc_);  \
  } while (0)

static void usage() {
  std::printf("usage:\tkuramoto [-hvnRFA] [-N N] [-K K] [-D D]\n");
  std::printf("\t[-c c] [-t t] [-d dt] [-f f] [-s seed] \n");
  std::printf("\t[-I init] [-r rtol] [-a atol] [-g gpu] filebase \n\n");
  std::printf("-h for help \n");
  std::printf("-v for verbose \n");
  std::printf("-n for normal random frequencies (default is cauchy) \n");
  std::printf("-R to reload adjacency, frequencies, and initial conditions from files if possible\n");
  std::printf("-F for fixed timestep \n");
  std::printf("-A to regenerate volcano adjacency each step to save memory \n");
  std::printf("D is the output density level. 0 for minimal, 1 for phases, 2 for phases and couplings, 3 for phases, couplings, and adjacency\n");
  std::printf("N is number of oscillators. Default 5000. \n");
  std::printf("K is rank of volcano adjacency. Default 5. \n");
  std::printf("c is the coupling coefficient. Default 3.0. \n");
  std::printf("t is total integration time. Default 1e2. \n");
  std::printf("dt is the time between outputs. Default 1e0. \n");
  std::printf("f is the scale of the frequencies. Default 1e0. \n");
  std::printf("seed is random seed. Default 1. \n");
  std::printf("init is uniform random initial condition scale. Default 0. \n");
  std::printf("rtol is relative error tolerance. Default 0.\n");
  std::printf("atol is absolute error tolerance. Default 1e-6.\n");
  std::printf("gpu is index of the gpu. Default 0.\n");
  std::printf("filebase is base file name for output. \n");
  std::printf("extensions: --gauss --adj-dtype=native|bf16|f16|sym|sym-bf16|sym-f16 --rebuild --max-steps=n --chunk=n --exact-final\n");
}

static void append(const char *suffix, const void *data, size_t elem, size_t n) {
  FILE *f = std::fopen(fname(suffix).c_str(), "ab");
  if (!f) {
    std::fprintf(stderr, "kuramoto: cannot open %s\n", fname(suffix).c_str());
    std::exit(1);
  }
  if (n) std::fwrite(data, elem, n, f);
  std::fclose(f);
}

static void write_fs(const std::vector<R> &th, R t, R h, const std::vector<R> &floc) {
  FILE *f = std::fopen(fname("fs.dat").c_str(), "wb");
  if (!f) return;
  std::fwrite(th.data(), sizeof(R), th.size(), f);
  std::fwrite(&t, sizeof(R), 1, f);
  std::fwrite(&h, sizeof(R), 1, f);
  std::fwrite(floc.data(), sizeof(R), floc.size(), f);
  std::fclose(f);
}

// Output files are appended by a writer thread, so the file I/O of one
// drained chunk overlaps the device's next chunk.  Batches are written in
// the order they were drained (the same bytes and order as appending in the
// loop).  Phase rows are staged in page-locked buffers (two, reused) that
// kmt_read_outputs copies into at full PCIe rate.
struct OutBatch {
  std::vector<R> times;        // times.dat
  R *rows = nullptr;           // thetas.dat rows (a pooled pinned buffer), nrows x n_osc
  size_t nrows = 0, n_osc = 0;
  std::vector<R> cpl;          // couplings.dat rows
  std::vector<R> r;            // order.dat
  bool fs = false;             // rewrite fs.dat with the state below
  std::vector<R> fs_th, fs_f;
  R fs_t = 0, fs_h = 0;
};

class Writer {
 public:
  Writer(size_t row_bytes, int dense) : dense_(dense) {
    for (int q = 0; q < 2 && row_bytes > 0; ++q) {
      void *p = nullptr;
      if (kmt_host_alloc(row_bytes, &p) != KMT_OK) {
        std::fprintf(stderr, "kuramoto: %s\n", kmt_last_error());
        std::exit(1);
      }
      free_.push_back((R *)p);
      all_.push_back((R *)p);
    }
    th_ = std::thread([this] { run(); });
  }
  ~Writer() {
    finish();
    for (R *p : all_) kmt_host_free(p);
  }
  // a free pinned row buffer (waits for the writer to release one)
  R *take_rows() {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [this] { return !free_.empty(); });
    R *p = free_.front();
    free_.pop_front();
    return p;
  }
  void give_back(R *p) {
    std::lock_guard<std::mutex> lk(mu_);
    free_.push_back(p);
    cv_.notify_all();
  }
  void push(OutBatch &&b) {
    std::lock_guard<std::mutex> lk(mu_);
    q_.push_back(std::move(b));
    cv_.notify_all();
  }
  // write everything queued, stop the thread
  void finish() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (done_) return;
      done_ = true;
      cv_.notify_all();
    }
    th_.join();
  }

 private:
  void run() {
    for (;;) {
      OutBatch b;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [this] { return done_ || !q_.empty(); });
        if (q_.empty()) return;
        b = std::move(q_.front());
        q_.pop_front();
      }
      if (dense_ >= 1) append("times.dat", b.times.data(), sizeof(R), b.times.size());
      if (b.nrows > 0) {
        if (dense_ >= 1) append("thetas.dat", b.rows, sizeof(R), b.nrows * b.n_osc);
        if (dense_ >= 2) append("couplings.dat", b.cpl.data(), sizeof(R), b.cpl.size());
        append("order.dat", b.r.data(), sizeof(R), b.r.size());
      }
      if (b.rows) give_back(b.rows);
      if (b.fs) write_fs(b.fs_th, b.fs_t, b.fs_h, b.fs_f);
    }
  }
  int dense_;
  std::thread th_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<OutBatch> q_;
  std::deque<R *> free_;
  std::vector<R *> all_;
  bool done_ = false;
};

int main(int argc, char *argv[]) {
  const double start = now_s();
  R c1 = 3.0, t1 = 1e2, dt = 1e0, freq = 1.0, atl = 1e-6, rtl = 0, I = 0;
  int gpu = 0, seed = 1, fixed = 0, A = 0, verbose = 0, help = 1, dense = 3, normal = 0, reload = 0;
  int gauss = 0, storage = KMT_ADJ_NATIVE, max_steps = -1, chunk_max = 64, rebuild = 0, exact_final = 0;
  unsigned long N = 5000, K = 6;
  const char *filebase = nullptr;
  static struct option longopts[] = {{"gauss", no_argument, nullptr, 1001},
                                     {"adj-dtype", required_argument, nullptr, 1002},
                                     {"max-steps", required_argument, nullptr, 1003},
                                     {"chunk", required_argument, nullptr, 1004},
                                     {"rebuild", no_argument, nullptr, 1005},
                                     {"exact-final", no_argument, nullptr, 1006},
                                     {nullptr, 0, nullptr, 0}};
  // kuramoto.cu:298-362 (options may follow the positional filebase)
  while (optind < argc) {
    int c = getopt_long(argc, argv, "N:K:I:D:c:g:t:d:f:s:r:a:hvFnRA", longopts, nullptr);
    if (c != -1) {
      switch (c) {
        case 'N': N = (int)atoi(optarg); break;
        case 'K': K = (int)atoi(optarg); break;
        case 'A': A = 1; break;
        case 'I': I = (R)atof(optarg); break;
        case 'c': c1 = (R)atof(optarg); break;
        case 'g': gpu = (int)atof(optarg); break;
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
        case 1002:
          if (!std::strcmp(optarg, "native")) storage = KMT_ADJ_NATIVE;
          else if (!std::strcmp(optarg, "bf16")) storage = KMT_ADJ_BF16;
          else if (!std::strcmp(optarg, "f16")) storage = KMT_ADJ_F16;
          else if (!std::strcmp(optarg, "sym")) storage = KMT_ADJ_SYM;
          else if (!std::strcmp(optarg, "sym-bf16")) storage = KMT_ADJ_SYM | KMT_ADJ_BF16;
          else if (!std::strcmp(optarg, "sym-f16")) storage = KMT_ADJ_SYM | KMT_ADJ_F16;
          else { std::fprintf(stderr, "kuramoto: unknown --adj-dtype %s\n", optarg); return 1; }
          break;
        case 1003: max_steps = atoi(optarg); break;
        case 1004: chunk_max = std::max(1, atoi(optarg)); break;
        case 1005: rebuild = 1; break;
        case 1006: exact_final = 1; break;
        default: break;
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
  base = filebase;

  R t = 0, h;
  FILE *out = std::fopen(fname(".out").c_str(), "ab");
  if (!out) {
    std::fprintf(stderr, "kuramoto: cannot open %s\n", fname(".out").c_str());
    return 1;
  }
  std::vector<R> yloc(N, 0), floc(2 * N, 0), omegas(N, 0);

  std::fprintf(out, "%lu %lu %f %f %f %i\n", N, K, (double)t1, (double)dt, (double)c1, seed);
  for (int i = 0; i < argc; i++) std::fprintf(out, "%s ", argv[i]);
  std::fprintf(out, "\n");

  // memory check, kuramoto.cu:424-438 (stored adjacency bytes per element
  // follow the chosen storage)
  size_t fr = 0, total = 0;
  CALL(kmt_mem_info(gpu, &fr, &total));
  const size_t adj_elem = (storage == KMT_ADJ_NATIVE || storage == KMT_ADJ_SYM) ? sizeof(R) : 2;
  const bool sym = (storage & KMT_ADJ_SYM) != 0;
  // symmetric storage: 256 x 256 blocks of the upper block triangle, plus
  // two partial sums (C, S) per row of every block
  const size_t nbk = (N + 255) / 256, sym_blocks = nbk * (nbk + 1) / 2;
  const size_t adj_bytes = sym ? sym_blocks * 256 * (256 * adj_elem + 4 * sizeof(R)) : N * N * adj_elem;
  const size_t req = A ? 100 * N * sizeof(R) : (100 + 2 * K) * N * sizeof(R) + adj_bytes;
  std::printf("GPU Memory: %lu %lu %lu\n", (unsigned long)fr, (unsigned long)total, (unsigned long)req);
  std::fprintf(out, "GPU Memory: %lu %lu %lu\n", (unsigned long)fr, (unsigned long)total, (unsigned long)req);
  if (fr < req) {
    std::printf("GPU Memory low!\n");
    std::fprintf(out, "GPU Memory low!\n");
    std::fclose(out);
    return 0;
  }
  std::fflush(out);

  CALL(kmt_create(&g_ctx, gpu, PREC, (long)N, (double)atl, (double)rtl, fixed, (double)c1));

  h = fixed ? dt : dt / 100;

  // initial conditions, kuramoto.cu:466-506
  int reloaded = 0;
  FILE *in;
  if (reload && (in = std::fopen(fname("fs.dat").c_str(), "r"))) {
    reloaded = 1;
    std::printf("Using initial conditions from file\n");
    std::fprintf(out, "Using initial conditions from file\n");
    size_t rd = std::fread(yloc.data(), sizeof(R), N, in);
    if (rd != N) {
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
  bool freq_file = false;
  if (reload) {
    FILE *fq = std::fopen(fname("frequencies.dat").c_str(), "r");
    if (fq) { freq_file = true; std::fclose(fq); }
  }
  // the rand() stream: IC draws (unless reloaded), then frequency draws
  // (unless read from file), kuramoto.cu:416, :498, :526
  CALL(kmt_draws(PREC, (long)N, seed, normal, (double)freq, (double)I, reloaded ? nullptr : yloc.data(),
                 freq_file ? nullptr : omegas.data()));
  if (!reloaded) {
    std::printf("Using random initial conditions\n");
    std::fprintf(out, "Using random initial conditions\n");
    FILE *f0 = std::fopen(fname("fs.dat").c_str(), "wb");
    std::fwrite(yloc.data(), sizeof(R), N, f0);
    std::fclose(f0);
  }
  if (freq_file) {
    std::printf("Using frequencies from file\n");
    std::fprintf(out, "Using frequencies from file\n");
    in = std::fopen(fname("frequencies.dat").c_str(), "r");
    size_t rd = std::fread(omegas.data(), sizeof(R), N, in);
    std::fclose(in);
    if (rd != N) {
      std::printf("frequencies not compatible with N!");
      std::fprintf(out, "frequencies not compatible with N!");
      std::fclose(out);
      kmt_destroy(g_ctx);
      return 0;
    }
  } else {
    std::printf("Using random frequencies\n");
    std::fprintf(out, "Using random frequencies\n");
    FILE *fo = std::fopen(fname("frequencies.dat").c_str(), "wb");
    std::fwrite(omegas.data(), sizeof(R), N, fo);
    std::fclose(fo);
  }

  // output grid, kuramoto.cu:543-550
  const int n0 = int(t / dt) + 1;
  const int n1 = int(t1 / dt) + 1;
  const int n_eval = std::max(n1 - n0, 0);
  std::vector<R> t_eval(n_eval);
  std::vector<double> t_eval_d(n_eval);
  for (int n = n0, ind = 0; n < n1; n++, ind++) {
    t_eval[ind] = dt * n;
    t_eval_d[ind] = (double)t_eval[ind];
  }
  CALL(kmt_set_omega(g_ctx, omegas.data()));

  // adjacency, kuramoto.cu:557-581.  The volcano factor of rank K > 16 is
  // drawn on the device and kept as packed signs (the same signs as the host
  // draw); small ranks and Gaussian factors go through a real N x K factor.
  if (!gauss && K > 16) {
    CALL(kmt_set_factor_sign(g_ctx, (int)K, (uint32_t)seed));
  } else {
    std::vector<double> Xd((size_t)N * K);
    CALL(kmt_factor((long)N, (int)K, (uint32_t)seed, gauss, Xd.data()));
    std::vector<R> X(Xd.begin(), Xd.end()), w(K);
    Xd.clear();
    Xd.shrink_to_fit();
    CALL(kmt_factor_weights(PREC, (long)N, (int)K, gauss, w.data()));
    CALL(kmt_set_factor(g_ctx, X.data(), w.data(), (int)K));
  }
  if (!A && reload && (in = std::fopen(fname("adj.dat").c_str(), "r"))) {
    std::printf("Using adjacency from file\n");
    std::fprintf(out, "Using adjacency from file\n");
    std::vector<R> adj((size_t)N * N, 0);
    size_t rd = std::fread(adj.data(), sizeof(R), (size_t)N * N, in);
    (void)rd;  // the reference does not check the size either
    std::fclose(in);
    CALL(kmt_set_adjacency(g_ctx, adj.data(), PREC, (long)N, storage));
  } else {
    std::printf("Using random adjacency matrix\n");
    std::fprintf(out, "Using random adjacency matrix\n");
    if (!A) {
      CALL(kmt_build_adjacency(g_ctx, storage));
      if (dense >= 3) {
        FILE *fa = std::fopen(fname("adj.dat").c_str(), "wb");
        const long rows = std::max<long>(1, (long)((256ul << 20) / (sizeof(R) * N)));
        std::vector<R> buf((size_t)std::min<long>(rows, (long)N) * N);
        for (long r0 = 0; r0 < (long)N; r0 += rows) {
          const long nr = std::min<long>(rows, (long)N - r0);
          CALL(kmt_get_adjacency_rows(g_ctx, r0, r0 + nr, buf.data()));
          std::fwrite(buf.data(), sizeof(R), (size_t)nr * N, fa);
        }
        std::fclose(fa);
      }
    }
  }
  if (A) CALL(kmt_use_factored(g_ctx, rebuild ? KMT_LOWRANK_REBUILD : KMT_LOWRANK_CONTRACT));
  std::fflush(out);
  std::fclose(out);

  CALL(kmt_set_state(g_ctx, yloc.data(), (double)t, (double)h));
  const R t_start = t;

  // t = t0 outputs, kuramoto.cu:587-626
  if (!reloaded) {
    double r0 = 0;
    CALL(kmt_order(g_ctx, yloc.data(), &r0));
    R r = (R)r0;
    FILE *fo = std::fopen(fname("order.dat").c_str(), "wb");
    std::fwrite(&r, sizeof(R), 1, fo);
    std::fclose(fo);
    if (dense >= 1) {
      FILE *fa = std::fopen(fname("thetas.dat").c_str(), "wb");
      std::fwrite(yloc.data(), sizeof(R), N, fa);
      std::fclose(fa);
      FILE *ft = std::fopen(fname("times.dat").c_str(), "wb");
      std::fclose(ft);
    }
    if (dense >= 2) {
      CALL(kmt_rhs(g_ctx, yloc.data(), nullptr, floc.data()));
      FILE *fc = std::fopen(fname("couplings.dat").c_str(), "wb");
      std::fwrite(floc.data(), sizeof(R), 2 * N, fc);
      std::fclose(fc);
    }
  }

  // dp45_run (dp45.cu:148-164) on the device; the host drains outputs
  const int keep_rows = dense >= 1 ? 1 : 0;
  int ring_rows = 4096;
  if (keep_rows) ring_rows = (int)std::max<long>(4, std::min<long>(1024, (long)((256ul << 20) / (sizeof(R) * N))));
  CALL(kmt_begin(g_ctx, t_eval_d.data(), n_eval, (double)t1, keep_rows, ring_rows));
  const double loop0 = now_s();
  kmt_status st;
  std::memset(&st, 0, sizeof(st));
  std::vector<R> ts(4096), hs(4096), rv(ring_rows);
  Writer writer(keep_rows ? sizeof(R) * (size_t)ring_rows * N : 0, dense);
  long steps = 0;
  int chunk = 1;
  bool started = false, stopped = false;
  for (;;) {
    int n = chunk;
    if (max_steps > 0) n = (int)std::max<long>(1, std::min<long>(n, max_steps - steps));
    CALL(kmt_advance(g_ctx, n, &st));
    started = true;
    chunk = std::min(chunk * 2, chunk_max);
    OutBatch b;
    // per accepted step: verbose line and times.dat (kuramoto.cu:169-188)
    int nt = 0;
    CALL(kmt_read_times(g_ctx, (int)ts.size(), ts.data(), hs.data(), &nt));
    for (int q = 0; q < nt; ++q) {
      ++steps;
      if (verbose) {
        const double el = now_s() - start;
        const double frac = (double)((ts[q] - t_start) / (t1 - t_start));
        const double eta = el / ((double)((ts[q] - t_start + hs[q]) / (t1 - t_start))) * (1 - frac);
        std::printf("%.3f\t%1.3e\t%1.3e\t%f\t%li\t\r", frac, el, eta, (double)hs[q], steps);
        std::fflush(stdout);
        FILE *fo = std::fopen(fname(".out").c_str(), "ab");
        std::fprintf(fo, "%.3f\t%1.3e\t%1.3e\t%f\t%li\t\n", frac, el, eta, (double)hs[q], steps);
        std::fclose(fo);
      }
    }
    b.times.assign(ts.begin(), ts.begin() + nt);
    // dense-output rows and order parameters (kuramoto.cu:212-249)
    int nr = 0;
    if (keep_rows) b.rows = writer.take_rows();
    if (dense >= 2) {  // couplings of each row from the row in device memory (kuramoto.cu:226-231)
      const int want = std::min(st.rows_ready, ring_rows);  // rows the last advance left in the ring
      b.cpl.resize((size_t)want * 2 * N);
      CALL(kmt_read_outputs_couplings(g_ctx, want, b.rows, rv.data(), b.cpl.data(), &nr));
      b.cpl.resize((size_t)nr * 2 * N);
    } else {
      CALL(kmt_read_outputs(g_ctx, ring_rows, b.rows, rv.data(), &nr));
    }
    b.nrows = (size_t)nr;
    b.n_osc = N;
    if (nr > 0) {
      if (dense >= 2) floc.assign(b.cpl.end() - 2 * N, b.cpl.end());
      b.r.assign(rv.begin(), rv.begin() + nr);
    }
    // restart file (kuramoto.cu:251-264; kuramoto_64.cu writes it always)
    if (nt > 0 && (sizeof(R) == 8 || dense >= 0)) {
      double tt = 0, hh = 0;
      CALL(kmt_get_state(g_ctx, yloc.data(), &tt, &hh));
      b.fs = true;
      b.fs_t = (R)tt;
      b.fs_h = (R)hh;
      b.fs_f = floc;
      b.fs_th = yloc;
    }
    writer.push(std::move(b));
    if (!st.active) break;
    if (max_steps > 0 && steps >= max_steps && st.pending == 0) {
      stopped = true;
      break;
    }
  }
  writer.finish();
  const double loop_s = now_s() - loop0;
  if (stopped) {  // leave the files as a process killed after this step would
    kmt_destroy(g_ctx);
    return 0;
  }

  // final state with couplings, kuramoto.cu:630-647: interpolation over the
  // last step with k1 already replaced by k7, integrator t and h; with
  // --exact-final the integrator's own state at that t
  std::vector<R> yev(N);
  double tt = 0, hh = 0;
  CALL(kmt_get_state(g_ctx, yev.data(), &tt, &hh));
  if (!exact_final && started && st.accepted > 0 && n_eval > 0) CALL(kmt_eval(g_ctx, t_eval_d[n_eval - 1], yev.data()));
  CALL(kmt_rhs(g_ctx, yev.data(), nullptr, floc.data()));
  write_fs(yev, (R)tt, (R)hh, floc);

  out = std::fopen(fname(".out").c_str(), "ab");
  const double rt = now_s() - start;
  std::printf("\nruntime: %f\n", rt);
  std::fprintf(out, "\nsteps: %li\n", steps);
  std::fprintf(out, "runtime: %f\n", rt);
  std::fprintf(out, "rejected: %li\nrhs_evals: %li\nloop_seconds: %f\nosc_steps_per_s: %e\n", st.rejected, st.rhs_evals,
               loop_s, loop_s > 0 ? (double)N * steps / loop_s : 0.0);
  std::fclose(out);
  kmt_destroy(g_ctx);
  return 0;
}
