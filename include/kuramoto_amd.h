/*
 * kuramoto_amd — C ABI of the MI355X Kuramoto / Dormand–Prince 4(5) integrator.
 *
 * This is the in-process operator boundary that replaces the reference's
 * integrator API (dp45.h:7-11, dp45_64.h:7-11) together with the model
 * callbacks it drives (dydt kuramoto.cu:118, makecoupling kuramoto.cu:104,
 * step_eval kuramoto.cu:164) and the `parameters` context
 * (kuramoto.cu:14-40).  Differences from the reference API, by design:
 *   - one opaque context per integrator instead of file-static globals
 *     (dp45.cu:24-31), so several integrators can live in one process;
 *   - the RHS is built in (fused HIP kernels) instead of a dydt callback;
 *   - the step controller runs on the device, so stepping is a pull API
 *     (kmt_advance + kmt_read_*) instead of a step_eval callback;
 *   - every entry point returns a status code (KMT_OK = 0); the message of
 *     the last failure on the calling thread is kmt_last_error().
 * All array arguments are HOST pointers (kmt_rhs_noalloc's alone are device
 * pointers); sizes are element counts.  Real
 * arrays use the context precision: float for KMT_F32, double for KMT_F64.
 * Adjacency layout everywhere: row-major, element [i*ld + j] multiplies
 * oscillator j in the equation of oscillator i (README.md:50, kuramoto.cu:113).
 */
#ifndef KURAMOTO_AMD_H
#define KURAMOTO_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KMT_OK 0
#define KMT_ERR_ARG (-1)     /* invalid argument or size */
#define KMT_ERR_HIP (-2)     /* a HIP runtime call failed */
#define KMT_ERR_STATE (-3)   /* call out of order (e.g. advance before begin) */
#define KMT_ERR_NUMERIC (-4) /* non-finite error norm: integration stopped */
#define KMT_ERR_MEMORY (-5)  /* not enough device memory */

/* working precision: `kuramoto` (float) / `kuramoto_64` (double) */
#define KMT_F32 0
#define KMT_F64 1

/* adjacency storage in HBM */
#define KMT_ADJ_NATIVE 0 /* same type as the working precision */
#define KMT_ADJ_BF16 2   /* bfloat16 with FP32 accumulation of the sums (with
                            KMT_F64 the state, stage combinations and error
                            norm stay FP64) */
#define KMT_ADJ_F16 3    /* IEEE half, the same way */
#define KMT_ADJ_SYM 4    /* flag: only the upper block triangle of a symmetric
                            matrix (half the bytes per RHS), in the working
                            precision (KMT_ADJ_SYM) or in 16 bits
                            (KMT_ADJ_SYM | KMT_ADJ_BF16 / _F16); no reference
                            counterpart: its gemvs read the full matrix */

typedef struct kmt_ctx kmt_ctx;

/* Snapshot of the device-resident integrator state, filled by kmt_advance. */
typedef struct kmt_status {
  double t;          /* integrator time (reference: *t in dp45_run) */
  double h;          /* next step size */
  double t_last;     /* start time of the last accepted step */
  long accepted;     /* accepted steps (reference: parameters.steps) */
  long rejected;     /* rejected attempts */
  long attempts;     /* all attempts */
  long rhs_evals;    /* RHS evaluations (6 per attempt + 1 per begin) */
  int active;        /* 1 while t < t1 or outputs are pending */
  int error;         /* nonzero after a non-finite error norm */
  int rows_ready;    /* output rows waiting to be read (kmt_read_outputs) */
  int times_ready;   /* accepted-step records waiting (kmt_read_times) */
  int eval_i;        /* outputs emitted so far (index into t_eval) */
  int pending;       /* outputs of the last accepted step not yet emitted */
} kmt_status;

/* Message of the last failed call on this thread ("" if none). */
const char *kmt_last_error(void);

/* Device query, used for the reference's memory check (kuramoto.cu:424-438). */
int kmt_device_count(int *count);
int kmt_mem_info(int device, size_t *free_bytes, size_t *total_bytes);
/* Page-locked host buffers for output staging (device-to-host copies into
 * them run at full PCIe rate and may overlap host work).  No reference
 * counterpart (the reference copies into pageable vectors, kuramoto.cu:
 * 221-231). */
int kmt_host_alloc(size_t bytes, void **ptr);
int kmt_host_free(void *ptr);

/* ---- host-side generators (no GPU work) ------------------------------- */
/* main()'s draw order: srand(seed); N initial-condition draws
 * I/2*(2u-1); N frequency draws, normal or Cauchy (kuramoto.cu:416,
 * :495-541; kuramoto_64.cu:492-538).  Bit-identical to the reference.
 * A NULL output skips its draws, as the reference skips the
 * initial-condition draws when -R reloads them. */
int kmt_draws(int precision, long N, int seed, int normal, double freq, double init, void *theta0, void *omega);
/* Low-rank factor X (N x K, row-major, double) from the library's
 * Philox4x32-10 stream: Gaussian (Box–Muller) or Rademacher (+-1).
 * Replaces makeu1 (kuramoto.cu:128-135) with a host generator. */
int kmt_factor(long N, int K, uint32_t seed, int gaussian, double *X);
/* Weights w so that M = X diag(w) X^T: volcano w_m = (-1)^m/N
 * (makeu2, kuramoto.cu:137-143) or Gaussian w_m = -1/(N K). */
int kmt_factor_weights(int precision, long N, int K, int gaussian, void *w);

/* ---- context (parameters + dp45_init / dp45_destroy) ------------------ */
/* dp45_init (dp45.cu:166-220) + the parameters struct (kuramoto.cu:14-40,
 * :555).  c1 is the coupling coefficient of the RHS
 * f_i = w_i + c1 (sin th_i C_i - cos th_i S_i) (kuramoto.cu:45). */
int kmt_create(kmt_ctx **ctx, int device, int precision, long N, double atol, double rtol, int fixed_step, double c1);
/* dp45_destroy (dp45.cu:222-251) plus the frees of main (kuramoto.cu:659-675). */
int kmt_destroy(kmt_ctx *ctx);

/* omegas upload (kuramoto.cu:552). */
int kmt_set_omega(kmt_ctx *ctx, const void *omega);
/* Phases, time and step size (dp45_init's y copy, dp45.cu:202; the -R
 * restart of kuramoto.cu:475-483). */
int kmt_set_state(kmt_ctx *ctx, const void *theta, double t, double h);
/* Latest accepted state (phases at the last accepted step, t, h): what the
 * reference writes to FILEBASEfs.dat after every step (kuramoto.cu:251-264). */
int kmt_get_state(kmt_ctx *ctx, void *theta, double *t, double *h);

/* Dense adjacency from host memory (the -R adj.dat path, kuramoto.cu:559-566).
 * M: N x N values of type `src_precision` with row stride ld (>= N).
 * storage: KMT_ADJ_NATIVE / _BF16 / _F16 (reduced storage uses a power-of-two
 * scale chosen from max|M| so the stored values are O(1)), or KMT_ADJ_SYM
 * (M must equal its transpose exactly; KMT_ERR_ARG otherwise). */
int kmt_set_adjacency(kmt_ctx *ctx, const void *M, int src_precision, long ld, int storage);
/* Low-rank factor for the adjacency: X (N x K, row-major, context
 * precision) and weights w (K).  Used by kmt_build_adjacency and -A. */
int kmt_set_factor(kmt_ctx *ctx, const void *X, const void *w, int K);
/* The reference's -A volcano factor for any rank K, drawn on the device:
 * X_nm = +-1 from the library's Philox stream (the same signs as
 * kmt_factor(N, K, seed, 0)) with w_m = (-1)^m / N (makeu1 / makeu2 / m1 /
 * m2, kuramoto.cu:62-143, which regenerate the signs from Philox counters
 * and store none), kept packed at one bit per entry plus its transpose
 * (N K / 4 bytes in all, 2.5 GB at N = K = 10^5) instead of an N x K real
 * array from the host.  Afterwards kmt_use_factored(ctx, KMT_LOWRANK_CONTRACT)
 * couples through it (two passes per RHS on the int8 matrix cores: the
 * vectors as 56-bit fixed point, exact integer sums, one rounding to
 * double), and kmt_build_adjacency builds the dense matrix from it (native,
 * BF16 or FP16 storage) with the same fma chain as from a real factor.
 * N and K up to 2^20 (KMT_ERR_ARG beyond).  Replaces a previous
 * kmt_set_factor; no register-rebuild form. */
int kmt_set_factor_sign(kmt_ctx *ctx, int K, uint32_t seed);
/* Build the dense adjacency on the device from the factor (getadj,
 * kuramoto.cu:145-161): M_ij = sum_m fma(w_m, X_im X_jm, .) in the working
 * precision, stored as `storage`.  Exactly symmetric. */
int kmt_build_adjacency(kmt_ctx *ctx, int storage);
/* -A: couple through the factor instead of a dense matrix (makecoupling's
 * m1/m2 branch, kuramoto.cu:62-108).  mode KMT_LOWRANK_CONTRACT: X diag(w)
 * (X^T v), O(N K) per RHS (the reference's algorithm); KMT_LOWRANK_REBUILD:
 * adjacency tiles M_ij rebuilt in registers from X for every stage, O(N^2 K)
 * VALU work, no N x N storage, results equal to the dense path on the
 * generated matrix (K <= 16); 0 returns to the dense matrix. */
#define KMT_LOWRANK_CONTRACT 1
#define KMT_LOWRANK_REBUILD 2
int kmt_use_factored(kmt_ctx *ctx, int mode);
/* Read back rows [i0, i1) of the dense adjacency (dequantized, context
 * precision, N values per row) — the adj.dat writer (kuramoto.cu:570-578). */
int kmt_get_adjacency_rows(kmt_ctx *ctx, long i0, long i1, void *out);
/* Storage scale: M = scale * stored (1 for native storage). */
int kmt_get_adjacency_scale(kmt_ctx *ctx, double *scale);

/* ---- model evaluations ------------------------------------------------- */
/* dydt (kuramoto.cu:118-126): f = RHS(theta); optionally the couplings
 * (C_i, S_i) interleaved, 2N values (makecoupling, kuramoto.cu:104-116). */
int kmt_rhs(kmt_ctx *ctx, const void *theta, void *f, void *couplings);
/* kmt_rhs without any allocation, on DEVICE vectors: theta_dev (N values)
 * in, f_dev (N values) and/or couplings_dev (2N values, interleaved C_i,
 * S_i) out, launched on `stream` (a hipStream_t; NULL: the context's own
 * stream) with no synchronisation; the caller synchronises before reading
 * the outputs.  No counterpart in the reference: this is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The (cos, sin) vector lives in
 * a fixed 4 MiB device array of the library, one per GPU, so N <= 524288
 * (FP32) / 262144 (FP64); larger contexts get KMT_ERR_ARG.  All calls on one
 * GPU share that array, and the factored (-A) path also uses the context's
 * partial-sum buffers: calls that overlap on one GPU (other streams, other
 * host threads) race.  Results equal kmt_rhs's bit for bit (same kernels). */
int kmt_rhs_noalloc(kmt_ctx *ctx, const void *theta_dev, void *f_dev, void *couplings_dev, void *stream);
/* Order parameter r = |mean exp(i theta)| (kuramoto.cu:216-220, :592-597). */
int kmt_order(kmt_ctx *ctx, const void *theta, double *r);

/* ---- integration (dp45_run / dp45_step / dp45_eval / step_eval) ------- */
/* Start a run (dp45_run's first dydt, dp45.cu:150-151): output grid
 * t_eval[n_eval] (kuramoto.cu:543-550), end time t1, whether interpolated
 * phase rows are kept for kmt_read_outputs (dense >= 1), and the number of
 * output rows the device ring holds before the integrator pauses. */
int kmt_begin(kmt_ctx *ctx, const double *t_eval, int n_eval, double t1, int keep_rows, int ring_rows);
/* Run up to max_attempts step attempts on the device (the while loop of
 * dp45_run, dp45.cu:153-162), then synchronise and report the state. */
int kmt_advance(kmt_ctx *ctx, int max_attempts, kmt_status *st);
/* Drain the output ring: up to max_rows rows of interpolated phases
 * (N values each, if keep_rows) and order parameters (step_eval,
 * kuramoto.cu:212-249).  Either pointer may be NULL. */
int kmt_read_outputs(kmt_ctx *ctx, int max_rows, void *theta_rows, void *r, int *n_rows);
/* kmt_read_outputs plus the couplings (C_i, S_i) of every drained row, 2N
 * values a row, interleaved as kmt_rhs writes them (step_eval's
 * makecoupling for couplings.dat, kuramoto.cu:226-231), computed on the
 * device from the row still in the ring: no host round trip of the phases.
 * Needs keep_rows; equal bit for bit to kmt_rhs on the returned rows. */
int kmt_read_outputs_couplings(kmt_ctx *ctx, int max_rows, void *theta_rows, void *r, void *couplings, int *n_rows);
/* Drain the accepted-step log: t and h after each accepted step
 * (times.dat, kuramoto.cu:181-188, and the verbose line's h). */
int kmt_read_times(kmt_ctx *ctx, int max_n, void *t, void *h, int *n);
/* One attempt with a host synchronisation (dp45_step, dp45.cu:87-141). */
int kmt_step(kmt_ctx *ctx, int *accepted, double *t, double *h);
/* Dense output over the last accepted step at t_eval (dp45_eval,
 * dp45.cu:143-146).  After the run this is the final-state interpolation of
 * kuramoto.cu:632 (with k1 already replaced by k7, as in the reference). */
int kmt_eval(kmt_ctx *ctx, double t_eval, void *theta);
/* Time the dominant kernel: `reps` launches of the fused stage kernel on
 * the library's stream, bracketed by HIP events; average ms per launch. */
int kmt_time_stage_kernel(kmt_ctx *ctx, int reps, double *avg_ms);
/* Time `attempts` step attempts on the library's stream with HIP events;
 * reports wall ms for the whole region and the average ms of the stage
 * kernel launches inside it (events around each stage launch). */
int kmt_time_attempts(kmt_ctx *ctx, int attempts, double *region_ms, double *stage_avg_ms, kmt_status *st);

/* Measured bandwidth ceiling of the device: a plain 16-byte copy kernel
 * (reads + writes `bytes` each way) and a read-only stream of `bytes`,
 * `reps` launches each in three launch shapes (4 or 8 loads in flight per
 * lane; 8, 16 or 32 workgroups per CU), timed with HIP events; the best
 * shape's GB/s (1e9 B/s) is reported.  No reference counterpart: the
 * denominator of the roofline next to the nominal peak, and the
 * calibration kernel of the HBM byte counters. */
int kmt_bandwidth_probe(int device, size_t bytes, int reps, double *copy_gbps, double *read_gbps);

/* Autotune the dense stage kernel on this context's adjacency: rows per
 * workgroup {4, 8, 16} x non-temporal loads {off, on} x grid {one workgroup
 * per row group (timed with and without the chunk ring), balanced (one
 * resident round, equal contiguous row ranges)}, `reps` timed launches
 * each; keeps the fastest.  variant_ms (12 values, may be NULL) receives the
 * average ms of each variant (the faster ring form for one workgroup per
 * group), index 4*r + 2*nt + balanced with r = 0, 1, 2 for 4, 8, 16 rows.
 * Every row is summed in the same order in every variant; the error norm's
 * partial sums follow the variant's workgroups (the ring groups them as the
 * loop without it does), so its last bits (and h's) can differ between
 * variants.  Without this call kmt_begin uses a fixed choice (4 rows, plain
 * loads when N >= 4096 and the matrix is <= 128 MB; else 8 rows, 16 for
 * 16-bit entries, non-temporal; the chunk ring unless the entries are 16
 * bits), so that every context and process gives the same bits.
 * Calling it (mid-run, or in only one of two contexts) therefore breaks
 * bit-exact restart against contexts that use the fixed choice.  KMT_ADJ_SYM has one
 * form and is not tuned.  No reference counterpart (cuBLAS chooses its own
 * gemv kernel). */
int kmt_autotune(kmt_ctx *ctx, int reps, double *variant_ms);
/* Pin the dense stage-kernel variant instead of the autotuner's choice:
 * rows per group 4, 8 or 16; variant_flags as kmt_layout reports them. */
int kmt_select_variant(kmt_ctx *ctx, int rows_per_block, int variant_flags);

/* Sizes the library chose (for roofline bookkeeping): padded row stride of
 * the adjacency in elements (KMT_ADJ_SYM: the block edge, 256), bytes per
 * stored element, rows per group of the
 * stage kernel, and variant_flags (bit 0: non-temporal adjacency loads, bit
 * 1: balanced grid, bit 2: chunk ring — each row's next chunk loads while
 * the current one is reduced; not with bit 1).  Any pointer may be NULL. */
int kmt_layout(kmt_ctx *ctx, long *lda, int *elem_bytes, int *rows_per_block, int *variant_flags);

/* ---- batched integration (an extension; no reference function:
 * the reference runs one instance per process, SLURM arrays of
 * kuramoto calls, job1.sh:13-25) ------------------------------------------
 * B independent FP32 runs advanced in lockstep stages on one GPU, each with
 * its own step size, accept/reject, output cursor and order-parameter ring.
 * shared = 1: all instances have the same N and share one adjacency stored
 * in BF16/FP16, coupled as one MFMA product A.[C S] per stage (B <= 64).
 * shared = 0: instance b has its own N_b and its own adjacency (any
 * storage), coupled by a grouped gemv-pair launch. */
typedef struct kmt_batch kmt_batch;
typedef struct kmt_batch_status {
  long accepted;          /* summed over instances */
  long rejected;
  long attempts;          /* max over instances */
  long rhs_evals;         /* summed */
  int active;             /* instances still running */
  int errors;             /* instances stopped by a non-finite error norm */
  int rows_ready;         /* max order-parameter values waiting per instance */
  int pad_;
  double t_min, t_max;    /* instance times */
  double osc_steps;       /* sum_b N_b * accepted_b */
  long proposer_accepts;  /* attempts whose h came from the learned proposer */
  long proposer_rejects;
} kmt_batch_status;

int kmt_batch_create(kmt_batch **batch, int device, int B, const long *N, double atol, double rtol, int shared,
                     int storage, int fixed_step);
int kmt_batch_destroy(kmt_batch *batch);
/* frequencies, phases, coupling c1, start time and first step of instance b */
int kmt_batch_set_instance(kmt_batch *batch, int b, const float *omega, const float *theta0, double c1, double t0,
                           double h0);
/* dense adjacency of instance b (b = -1: the shared one) built on the device
 * from a factor X (N x K, FP32) and weights w: M = X diag(w) X^T */
int kmt_batch_set_factor(kmt_batch *batch, int b, const float *X, const float *w, int K);
/* dequantized rows [i0, i1) of instance b's adjacency (b = -1: shared) */
int kmt_batch_get_adjacency_rows(kmt_batch *batch, int b, long i0, long i1, float *out);
/* shared output grid and end time for all instances; ring_rows = r values
 * per instance kept on the device between reads */
int kmt_batch_begin(kmt_batch *batch, const double *t_eval, int n_eval, double t1, int ring_rows);
int kmt_batch_advance(kmt_batch *batch, int max_attempts, kmt_batch_status *st);
/* drain instance b's order parameters r(t_eval[j]) in output order */
int kmt_batch_read_order(kmt_batch *batch, int b, int max_n, float *r, int *n);
/* phases at the last accepted step; state8 = {t, h, accepted, rejected,
 * active, error, proposer accepts, proposer rejects} (either may be NULL) */
int kmt_batch_get_instance(kmt_batch *batch, int b, float *theta, double *state8);
/* tile autotuner of the shared MFMA product: workgroup shape `variant`
 * (0: 4 waves x 32 rows of A, 1: 8 waves x 32 rows, 2: 8 waves x 64 rows,
 * 3 / 4: 8 waves x 32 rows with two / three 64-wide chunks of A in flight,
 * 5: 4 waves x 32 rows with two chunks in flight) x split-K 1..16, `reps`
 * launches each; keeps the fastest (time plus the split-K partial
 * traffic).  kmt_batch_tune_table returns the 96 average times (ms),
 * variant major: ms96[16 variant + split_k - 1]. */
int kmt_batch_autotune(kmt_batch *batch, int reps, int *variant, int *split_k, double *best_ms);
int kmt_batch_tune_table(kmt_batch *batch, double *ms96);
/* pin the shared product's workgroup variant and split-K instead of the
 * autotuner's choice */
int kmt_batch_select_tile(kmt_batch *batch, int variant, int split_k);
/* shared mode: once finished instances leave the active ones fitting in
 * fewer 32-row MFMA tiles, kmt_batch_advance renumbers the active instances'
 * rows and runs the product on those rows only (default on; results are
 * bit-identical either way).  kmt_batch_product_rows reports the rows the
 * product currently computes (2B in grouped mode) and how many times the
 * rows were compacted since kmt_batch_create. */
int kmt_batch_set_compaction(kmt_batch *batch, int enabled);
int kmt_batch_product_rows(kmt_batch *batch, int *rows, int *compactions);
/* HIP-event timing of `attempts` attempts; coupling_ms = average of the
 * coupling launches (MFMA product or grouped gemv) inside the region */
int kmt_batch_time_attempts(kmt_batch *batch, int attempts, double *region_ms, double *coupling_ms,
                            kmt_batch_status *st);
/* shared mode: 1 (default) = cos/sin as hi + lo 16-bit parts (two MFMAs),
 * 0 = hi part only */
int kmt_batch_set_vector_split(kmt_batch *batch, int hi_lo);
/* learned step-size proposer (MLP 8-64-64-1, FP16 weights, FP32
 * accumulation): after an accepted step the next h is the standard
 * controller's proposal times exp(mlp(features)); the embedded error still
 * accepts or rejects every step and a rejection falls back to the standard
 * controller.  weights: kmt_proposer_param_count() FP16 values (NULL keeps
 * the loaded ones). */
int kmt_batch_set_proposer(kmt_batch *batch, const uint16_t *weights_f16, int n_params, int enabled);
int kmt_proposer_param_count(void);
/* attempt log for training the proposer: capacity records of 14 floats
 * {8 features, h used, the standard controller's proposal behind it, error
 *  norm, accepted, instance, 1 if h came from the proposal made after an
 *  accepted step}, returned sorted by (instance, attempt) so that a run
 * reads back the same log every time; reading a log that overflowed its
 * capacity fails (the kept subset would depend on the order the
 * instances' controllers ran in). */
int kmt_batch_set_attempt_log(kmt_batch *batch, int capacity);
int kmt_batch_read_attempt_log(kmt_batch *batch, int max_n, float *records, int *n);

/* ---- training the learned step-size proposer on the GPU ----------------
 * MLP 8 -> 64 -> 64 -> 1 (tanh, tanh, linear), parameters packed as
 * W1[64][8], b1[64], W2[64][64], b2[64], W3[64], b3 (kmt_proposer_param_count()
 * values).  FP16 copy for inference (the one kmt_batch_set_proposer takes),
 * FP32 master weights for Adam; FP32 accumulation.  Loss: mean squared error. */
typedef struct kmt_mlp kmt_mlp;
/* w_init: FP32 parameters, or NULL for Glorot-uniform hidden layers from
 * `seed` with a zero output layer (the untrained proposer returns the
 * standard controller's step). */
int kmt_mlp_create(kmt_mlp **mlp, int device, const float *w_init, uint32_t seed);
int kmt_mlp_destroy(kmt_mlp *mlp);
int kmt_mlp_get_weights(kmt_mlp *mlp, float *w32, uint16_t *w16);
/* outputs for n samples X[n][8] with the FP16 weights */
int kmt_mlp_forward(kmt_mlp *mlp, const float *X, int n, float *out);
/* mean squared error and its gradient with respect to the (FP16-rounded)
 * parameters over n samples */
int kmt_mlp_loss_grad(kmt_mlp *mlp, const float *X, const float *y, int n, double *loss, float *grad);
/* Adam (beta 0.9 / 0.999) over minibatches of `batch` samples in the given
 * order, `epochs` passes; loss_history[e] = mean loss of epoch e */
int kmt_mlp_train(kmt_mlp *mlp, const float *X, const float *y, int n, int epochs, int batch, double lr,
                  double *loss_history);

#ifdef __cplusplus
}
#endif
#endif
