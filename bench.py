"""Benchmark of the fused DP45 Kuramoto integrator on one MI355X.

A "step" is one DP45 step attempt over the whole network: six fused
gemv-pair stage kernels (each streams the adjacency from HBM once), the
device controller, dense output / commit and bookkeeping.

Default workload (`--config fp64-n50000`): N = 50000 oscillators, FP64,
dense rank-5 Gaussian adjacency A = X X^T / K (X ~ N(0,1), Philox4x32-10
seed 2) stored in FP64 (20 GB), normal frequencies, theta0 = pi(2u - 1),
c = 3, atol = 1e-9, output every dt = 1 up to t = 100.  Inputs are synthetic
and resident in HBM before the timed region starts.

Printed JSON line (one):
  value       = N * accepted steps in the timed region / timed seconds
                (osc*steps/s, the reference's own unit)
  roofline    = the stage kernel: algorithmic bytes per launch
                (s_A N^2 + N (2 s_y + s_w)) / average launch time (HIP events
                around every stage launch of the timed region, on the
                library's stream), against 8 TB/s
  cpu_baseline = the oracle (FP64 C++ transcription, OpenMP) on the same
                workload, first accepted steps only (bounded sample)
  secondary   = (default workload only) the FP32 N=50000 dense run and the
                64-instance shared BF16 batch, measured in the same process
                after the headline figure (per-RHS roofline; batched vs 64
                plain attempts)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from kuramoto_amd.workloads import BATCH as BATCH_CONFIGS  # noqa: E402
from kuramoto_amd.workloads import PLAIN as CONFIGS  # noqa: E402
from kuramoto_amd.workloads import make_batch  # noqa: E402

HBM_PEAK = 8.0e12


def profiled_traffic(config, kernel="stage_kernel"):
    """HBM bytes per launch of the dominant kernel measured by the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs recorded in profiles/<kernel>_<config>.txt
    (None if that workload has not been profiled)."""
    path = os.path.join(ROOT, "profiles", kernel + "_" + config.replace("-", "_") + ".txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1].split()[0])
    return None


def timed_advance(run, steps):
    """The timed region: exactly `steps` attempts through the production path
    (replayed hipGraphs of 8 and 1 attempts), host clock from a synchronized
    start to the status read-back that waits for the last attempt."""
    t0 = time.perf_counter()
    st = run.advance(steps)
    return (time.perf_counter() - t0) * 1e3, st


def gpu_run(cfg, steps, warmup):
    import kuramoto_amd as km

    p = km.F32 if cfg["prec"] == "f32" else km.F64
    dtype = np.float32 if p == km.F32 else np.float64
    N, K = cfg["N"], cfg["K"]
    theta0, omega = km.draws(p, N, cfg["seed"], normal=cfg["normal"], freq=1.0, init=cfg["init"])
    dt = dtype(cfg["dt"])
    te = (dt * np.arange(1, int(dtype(cfg["t1"]) / dt) + 1)).astype(np.float64)
    it = km.Integrator(N, p, atol=cfg["atol"], c1=cfg["c1"])
    it.set_omega(omega)
    t0 = time.time()
    if cfg["storage"] == "signs":
        it.set_factor_sign(K, cfg["seed"])
    else:
        X = km.factor(N, K, cfg["seed"], gaussian=True).astype(dtype)
        it.set_factor(X, km.factor_weights(p, N, K, gaussian=True))
        t0 = time.time()
    if cfg["storage"] in ("factored", "signs"):
        it.use_factored("contract")
    elif cfg["storage"] == "rebuild":
        it.use_factored("rebuild")
    else:
        it.build_adjacency(cfg["storage"])
    gen_s = time.time() - t0
    it.set_state(theta0, 0.0, float(dt / dtype(100)))
    it.begin(te, cfg["t1"], keep_rows=False, ring_rows=4096)
    if warmup:
        it.advance(warmup)
    it.read_times()
    it.read_outputs()
    st0 = it.advance(0)
    region_ms, st = timed_advance(it, steps)
    # per-launch kernel time for the roofline: HIP events around every stage
    # launch of further attempts, after the timed region (the event markers
    # add a few microseconds between launches, so they stay out of it)
    _, stage_ms, _ = it.time_attempts(min(steps, 10))
    lda, eb, rows, nt, bal = it.layout()
    ring = bool(it.variant_flags() & 4)
    it.close()
    copy_gbps, read_gbps = km.bandwidth_probe(8 << 30, 5)
    return dict(dtype=dtype, N=N, region_ms=region_ms, stage_ms=stage_ms, st0=st0, st=st, lda=lda, elem_bytes=eb,
                gen_s=gen_s, rows=rows, nt=nt, balanced=bal, ring=ring, copy_gbps=copy_gbps, read_gbps=read_gbps)


def cpu_baseline(cfg, max_steps):
    """Oracle (FP64 transcription, OpenMP) on the same inputs, first
    `max_steps` accepted steps."""
    import kuramoto_amd as km
    from oracle import oracle_py as orc

    N, K = cfg["N"], cfg["K"]
    p = km.F32 if cfg["prec"] == "f32" else km.F64
    dtype = np.float32 if p == km.F32 else np.float64
    theta0, omega = orc.draws(N, cfg["seed"], cfg["normal"], 1.0, cfg["init"], dtype=dtype)
    signs = cfg["storage"] == "signs"
    if signs and N * K > 2e8:
        return dict(value=None, unit="osc*steps/s", cores=orc.num_threads(), kind="port",
                    sample=f"not run: the oracle's -A path holds X as an N x K array ({4 * N * K / 1e9:.0f} GB in FP32)")
    X = (orc.rademacher_factor(N, K, cfg["seed"]) if signs else orc.gaussian_factor(N, K, cfg["seed"])).astype(dtype)
    w = orc.factor_weights(N, K, not signs, dtype=dtype)
    # N^2 does not fit the host for the N=200000 cases: the oracle runs the
    # reference's own -A path (coupling through the factor) there
    lowrank = cfg["storage"] in ("factored", "f16", "rebuild", "signs")
    adj = orc.Adj(2, X=X, w=w) if lowrank else orc.Adj(0, M=orc.adj_rows(X, w, 0, N))
    dt = dtype(cfg["dt"])
    te = (dt * np.arange(1, int(dtype(cfg["t1"]) / dt) + 1)).astype(dtype)
    t0 = time.time()
    res = orc.integrate(adj, omega, cfg["c1"], theta0, 0.0, float(dt / dtype(100)), cfg["t1"],
                        cfg["atol"], t_eval=te, keep_rows=False, max_steps=max_steps, dtype=dtype)
    secs = time.time() - t0
    del adj
    return dict(value=N * res["accepted"] / secs, unit="osc*steps/s", cores=orc.num_threads(), kind="port",
                sample=f"oracle {np.dtype(dtype).name} {'low-rank (-A)' if lowrank else 'dense'}, N={N}, "
                       f"first {res['accepted']} accepted steps "
                       f"({res['rejected']} rejected, {res['rhs_evals']} RHS evals) in {secs:.1f} s, "
                       f"OpenMP threads={orc.num_threads()}")


def batch_to_time(name, proposer, t_end, tile=None):
    """Whole integration of every instance from t=0 to t_end (host wall
    clock around the advance loop, adjacency already resident)."""
    bt, Ns, _ = make_batch(name, proposer, tile)
    dt = np.float32(BATCH_CONFIGS[name]["dt"])
    te = (dt * np.arange(1, int(np.float32(t_end) / dt) + 1)).astype(np.float64)
    t0 = time.perf_counter()
    res = bt.run(te, t_end, ring_rows=1024, chunk=64)
    secs = time.perf_counter() - t0
    bt.close()
    st = res["status"]
    return dict(seconds=secs, r=res["r"], rhs_evals=st.rhs_evals, accepted=st.accepted, rejected=st.rejected,
                osc_steps=st.osc_steps, proposer_accepts=st.proposer_accepts, proposer_rejects=st.proposer_rejects)


def sequential_plain_to_time(name, t_end):
    """The shared batch's instances run one after another as plain FP32
    runs with the dense FP32 adjacency (the reference's own path), each from
    t=0 to t_end: host wall clock around each run's integration loop (the
    adjacency is generated before the clock starts, as for the batch)."""
    import kuramoto_amd as km

    cfg = BATCH_CONFIGS[name]
    N = cfg["Ns"][0]
    X = km.factor(N, 5, 1, gaussian=True).astype(np.float32)
    w = km.factor_weights(km.F32, N, 5, gaussian=True)
    dt = np.float32(cfg["dt"])
    te = (dt * np.arange(1, int(np.float32(t_end) / dt) + 1)).astype(np.float64)
    secs, rs, tot = 0.0, [], dict(rhs_evals=0, accepted=0, rejected=0)
    for b in range(cfg["B"]):
        th, om = km.draws(km.F32, N, cfg["seeds"][b], normal=cfg["normal"][b], init=6.283185307)
        with km.Integrator(N, km.F32, atol=1e-6, c1=float(cfg["cs"][b])) as it:
            it.set_omega(om)
            it.set_factor(X, w)
            it.build_adjacency()
            it.set_state(th, 0.0, 0.01)
            t0 = time.perf_counter()
            res = it.run(te, t_end, keep_rows=False, chunk=64)
            secs += time.perf_counter() - t0
        st = res["status"]
        for k in tot:
            tot[k] += getattr(st, k)
        rs.append(res["r"])
    return dict(seconds=secs, r=rs, **tot)


def batch_run(name, steps, warmup, proposer, tile=None):
    cfg = BATCH_CONFIGS[name]
    bt, Ns, tune = make_batch(name, proposer, tile)
    dt = np.float32(cfg["dt"])
    te = (dt * np.arange(1, int(np.float32(cfg["t1"]) / dt) + 1)).astype(np.float64)
    bt.begin(te, cfg["t1"], ring_rows=1024)
    if warmup:
        bt.advance(warmup)
    st0 = bt.advance(0)
    region_ms, st = timed_advance(bt, steps)
    _, coup_ms, _ = bt.time_attempts(min(steps, 10))
    bt.close()
    return dict(Ns=Ns, region_ms=region_ms, coupling_ms=coup_ms, st0=st0, st=st, tune=tune)


def plain_attempt_ms(N, steps=10):
    """One plain FP32 dense run at size N (FP32 adjacency): ms per attempt."""
    import kuramoto_amd as km

    X = km.factor(N, 5, 1, gaussian=True).astype(np.float32)
    th, om = km.draws(km.F32, N, 1, normal=False, init=6.283185307)
    with km.Integrator(N, km.F32, atol=1e-6, c1=3.0) as it:
        it.set_omega(om)
        it.set_factor(X, km.factor_weights(km.F32, N, 5, gaussian=True))
        it.build_adjacency()
        it.set_state(th, 0.0, 0.01)
        it.begin(np.arange(1, 51, dtype=np.float64), 50.0, keep_rows=False, ring_rows=64)
        it.advance(2)
        region_ms, _ = timed_advance(it, steps)
    return region_ms / steps


def batch_main(args):
    import kuramoto_amd as km

    cfg = BATCH_CONFIGS[args.config]
    tile = tuple(int(v) for v in args.pin_tile.split(",")) if args.pin_tile else None
    r = batch_run(args.config, args.steps, args.warmup, proposer=not cfg["shared"], tile=tile)
    st, st0 = r["st"], r["st0"]
    secs = r["region_ms"] / 1e3
    osc_steps = st.osc_steps - st0.osc_steps
    Ns = r["Ns"]
    hilo = cfg.get("hilo", True)
    storage = cfg.get("storage", "bf16")
    if cfg["shared"]:
        N = Ns[0]
        mpad = 128
        alg = 2 * N * N + (2 if hilo else 1) * 2 * mpad * N  # 16-bit A once + the hi (and lo) V^T once
        seq_ms = plain_attempt_ms(N)
        inst_steps = st.accepted - st0.accepted
        per_attempt_batched = r["region_ms"] / args.steps
        extra = {"rhs_evals_in_region": st.rhs_evals - st0.rhs_evals,
                 "rejected_in_region": st.rejected - st0.rejected,
                 "sequential_plain_fp32_ms_per_attempt_x64": 64 * seq_ms,
                 "batched_ms_per_attempt": per_attempt_batched,
                 "speedup_vs_64_sequential_plain": 64 * seq_ms / per_attempt_batched,
                 "tile": {"workgroup_variant": r["tune"][0], "split_k": r["tune"][1]}}
        # time to solution: every instance integrated from t=0 to t_end by the
        # batch (standard controller and learned proposer) and by 64 plain
        # runs one after another (t_end = 0: skipped)
        t_end = args.batch_t_end
        if t_end > 0:
            runs = {m: batch_to_time(args.config, m == "batched_learned_proposer", t_end, tile)
                    for m in ("batched_standard_controller", "batched_learned_proposer")}
            runs["sequential_plain_fp32"] = sequential_plain_to_time(args.config, t_end)
            plain_r = runs["sequential_plain_fp32"]["r"]
            dr = {m: max(float(np.max(np.abs(a - b))) for a, b in zip(runs[m]["r"], plain_r))
                  for m in ("batched_standard_controller", "batched_learned_proposer")}
            for m in runs:
                runs[m].pop("r")
            seq_s = runs["sequential_plain_fp32"]["seconds"]
            extra.update({f"whole_run_to_t{t_end:g}": runs,
                          "max_abs_dr_vs_sequential_plain": dr,
                          "time_to_solution_speedup_vs_sequential_plain": {
                              m: seq_s / runs[m]["seconds"]
                              for m in ("batched_standard_controller", "batched_learned_proposer")}})
    else:
        alg = sum(2 * n * n + 3 * 4 * n for n in Ns)
        # the learned proposer changes how many attempts reach a given t, so
        # the two controllers are compared on whole runs to the same end time
        t_end = args.batch_t_end
        extra = {"sum_N": int(sum(Ns))}
        if t_end > 0:
            runs = {m: batch_to_time(args.config, m == "learned_proposer", t_end)
                    for m in ("standard_controller", "learned_proposer")}
            dr = max(float(np.max(np.abs(a - b))) for a, b in zip(runs["standard_controller"]["r"],
                                                                   runs["learned_proposer"]["r"]))
            for m in runs:
                runs[m].pop("r")
                runs[m]["osc_steps_per_s"] = runs[m]["osc_steps"] / runs[m]["seconds"]
            extra.update({f"whole_run_to_t{t_end:g}": runs, "max_abs_dr_learned_vs_standard": dr,
                          "speedup_learned_vs_standard": runs["standard_controller"]["seconds"]
                          / runs["learned_proposer"]["seconds"]})
    achieved = alg / (r["coupling_ms"] / 1e3)
    copy_gbps, read_gbps = km.bandwidth_probe(8 << 30, 5)
    kname = "bgemm_kernel" if cfg["shared"] else "bgemv_kernel"
    out = {
        "metric": "osc*steps/s (sum over instances of N_b x accepted steps per second)",
        "value": osc_steps / secs,
        "unit": "osc*steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": r["region_ms"] / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": f"f32 ({storage} adjacency" + ("" if hilo else f", {storage} cos/sin in the product") + ")",
        "data": "synthetic (Philox Gaussian factors, glibc-rand frequencies/phases)",
        "config": {"workload": args.config, "B": cfg["B"], "instance_steps_in_region": st.accepted - st0.accepted,
                   **extra},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": profiled_traffic(args.config, kname),
                     "kernel": kname + ((" (MFMA, hi+lo vectors)" if hilo else " (MFMA, hi vectors only)")
                                        if cfg["shared"] else " (grouped gemv-pairs)"),
                     "bytes_per_launch": alg, "avg_launch_ms": r["coupling_ms"],
                     "measured_read_stream_gbps": read_gbps, "measured_copy_gbps": copy_gbps},
    }
    print(json.dumps(out), flush=True)


# The reference's own published runs of its programs (volcano.ipynb): the
# command line, the notebook cell and the "runtime:" and final step count
# it printed (an unnamed CUDA GPU with 25 GB; whole-process time: adjacency
# generation, integration, every output file).
PUBLISHED = [
    dict(prog="kuramoto", args="-N 10000 -K 10000 -c 1.75 -t 100 -d 0.01 -s 1 -nv", runtime=26.206418, steps=846,
         source="volcano.ipynb:47-52 (dense FP32)"),
    dict(prog="kuramoto", args="-N 10000 -K 10000 -c 1.75 -t 100 -d 0.01 -s 1 -nvA", runtime=93.887063, steps=846,
         source="volcano.ipynb:55-60 (-A FP32)"),
    dict(prog="kuramoto_64", args="-N 10000 -K 10000 -c 1.75 -t 100 -d 0.01 -s 1 -nv", runtime=60.521239,
         steps=845, source="volcano.ipynb:63-68 (dense FP64)"),
    dict(prog="kuramoto", args="-N 100000 -K 100000 -c 3 -t 20 -d 0.02 -s 1 -D1 -nvA", runtime=422.154478,
         steps=230, source="volcano.ipynb:319-346 (-A FP32, N = K = 10^5)"),
    dict(prog="kuramoto", args="-N 100000 -K 6 -c 3 -t 20 -d 0.02 -s 1 -D1 -nvA", runtime=13.55, steps=120,
         source="volcano.ipynb:547-573 (-A FP32, K = 6)"),
]


def published_main(args):
    """Each published command through this build's drop-in programs, in a
    scratch directory (files deleted after each run): the programs' own
    "runtime:" line and accepted steps against the reference's."""
    import shutil
    import subprocess
    import tempfile

    rows = []
    for p in PUBLISHED:
        work = tempfile.mkdtemp(prefix="kmt_pub_", dir=args.scratch)
        try:
            cmd = [os.path.join(ROOT, "kuramoto_amd", "bin", p["prog"])] + p["args"].split() + [
                os.path.join(work, "run")]
            t0 = time.perf_counter()
            out = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
            wall = time.perf_counter() - t0
            text = out.stdout + open(os.path.join(work, "run.out")).read()
            rt = [float(l.split()[1]) for l in text.splitlines() if l.startswith("runtime:")]
            st = [int(l.split()[1]) for l in text.splitlines() if l.startswith("steps:")]
            sizes = sum(os.path.getsize(os.path.join(work, f)) for f in os.listdir(work))
        finally:
            shutil.rmtree(work, ignore_errors=True)
        rows.append({"command": f"{p['prog']} {p['args']}", "source": p["source"], "rc": out.returncode,
                     "runtime_s": rt[-1] if rt else None, "process_wall_s": wall,
                     "accepted_steps": st[-1] if st else None, "reference_runtime_s": p["runtime"],
                     "reference_steps": p["steps"], "output_bytes": sizes,
                     "speedup_vs_reference": (p["runtime"] / rt[-1]) if rt else None})
        print(f"published: {rows[-1]}", file=sys.stderr, flush=True)
    print(json.dumps({"metric": "whole-process runtime of the reference's published commands (s)",
                      "unit": "s", "higher_is_better": False, "n_gpus": 1, "data": "synthetic, as the programs draw",
                      "runs": rows}), flush=True)


def rhs_bytes(cfg, r):
    """Algorithmic bytes of one RHS of a single-context workload."""
    N = r["N"]
    s_y = np.dtype(r["dtype"]).itemsize
    if cfg["storage"] == "factored":
        # per RHS: X read twice (N K s_y each) + theta-derived vectors in, k out, omega
        return 2 * N * cfg["K"] * s_y + N * (2 * s_y + s_y)
    if cfg["storage"] == "signs":
        # per RHS: the packed signs read twice (N K / 8 bytes each pass)
        return 2 * N * cfg["K"] // 8 + N * (2 * s_y + s_y)
    if cfg["storage"].startswith("sym"):
        # a symmetric matrix carries N (N + 1) / 2 distinct entries
        return r["elem_bytes"] * N * (N + 1) // 2 + N * (2 * s_y + s_y)
    return r["elem_bytes"] * N * N + N * (2 * s_y + s_y)


def secondary_measurements(steps, warmup):
    """More workloads measured in the same process after the headline one,
    so that one bench run also covers them: the FP32 N=50000 dense run (the
    plain integrator's target configuration, >= 85% of 8 TB/s per RHS), the
    FP64 N=50000 run on symmetric storage (upper block triangle), the
    reference's high-rank -A network at N = K = 10^5 (packed signs, int8
    passes), and the 64-instance shared-adjacency batch (BF16 A, MFMA
    product) per attempt against 64 plain FP32 attempts of the same size.
    No whole runs to a fixed time here (those take minutes;
    --config batch64-n20000-bf16 reports them)."""
    out = {}
    for name in ("fp32-n50000", "fp64-n50000-sym", "fp32-n100000-k100000-signs"):
        cfg = CONFIGS[name]
        r = gpu_run(cfg, steps, warmup)
        N = r["N"]
        out[name] = {
            "value_osc_steps_per_s": N * (r["st"].accepted - r["st0"].accepted) / (r["region_ms"] / 1e3),
            "ms_per_step": r["region_ms"] / steps, "stage_ms": r["stage_ms"],
            "roofline_frac": rhs_bytes(cfg, r) / (r["stage_ms"] / 1e3) / HBM_PEAK}
        if cfg["storage"] not in ("sym", "signs"):
            out[name]["variant"] = {"rows_per_group": r["rows"], "nontemporal_loads": r["nt"], "chunk_ring": r["ring"]}
    name = "batch64-n20000-bf16"
    b = batch_run(name, steps, warmup, proposer=False)
    Nb = b["Ns"][0]
    per_attempt = b["region_ms"] / steps
    plain_ms = plain_attempt_ms(Nb)
    alg_b = 2 * Nb * Nb + 2 * 2 * 128 * Nb
    out[name] = {"batched_ms_per_attempt": per_attempt, "plain_fp32_ms_per_attempt": plain_ms,
                 "instance_steps_speedup_vs_64_sequential_plain": 64 * plain_ms / per_attempt,
                 "product_ms": b["coupling_ms"], "product_roofline_frac": alg_b / (b["coupling_ms"] / 1e3) / HBM_PEAK,
                 "tile": {"workgroup_variant": b["tune"][0], "split_k": b["tune"][1]}}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="fp64-n50000", choices=sorted(CONFIGS) + sorted(BATCH_CONFIGS))
    ap.add_argument("--cpu-steps", type=int, default=30,
                    help="accepted steps of the CPU baseline sample: a fixed window from t=0 (0: skip); "
                         "30 steps of the FP64 N=50000 default take 15-25 s on 16 host threads")
    ap.add_argument("--batch-t-end", type=float, default=10.0,
                    help="batched workloads: end time of the whole runs that compare the step controllers "
                         "(and, for the shared batch, 64 sequential plain runs); 0 skips them")
    ap.add_argument("--pin-tile", default=None,
                    help="shared batch: 'variant,split_k' of the MFMA product instead of the autotuner's choice")
    ap.add_argument("--published", action="store_true",
                    help="run the reference's published command lines through the drop-in programs instead")
    ap.add_argument("--scratch", default=None, help="--published: directory for the programs' output files")
    ap.add_argument("--no-secondary", action="store_true",
                    help="default workload: skip the FP32 N=50000, symmetric, packed-sign and batched "
                         "measurements that follow it")
    args = ap.parse_args()
    if args.published:
        published_main(args)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported (the row-sharded "
              "multi-GPU layout is a design only)", file=sys.stderr)
        sys.exit(2)
    if args.config in BATCH_CONFIGS:
        batch_main(args)
        return
    cfg = CONFIGS[args.config]
    r = gpu_run(cfg, args.steps, args.warmup)
    N = r["N"]
    accepted = r["st"].accepted - r["st0"].accepted
    secs = r["region_ms"] / 1e3
    alg_bytes = rhs_bytes(cfg, r)
    achieved = alg_bytes / (r["stage_ms"] / 1e3)
    roof = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK}
    if cfg["storage"] == "signs":
        # two passes over the N x K signs per RHS on the int8 matrix cores:
        # the bits (N K / 8 bytes per pass) streamed from HBM is the binding
        # roof (the MFMA work, 16 digit columns x 2 ops per entry and pass,
        # 64 N K int8 ops per RHS, takes 40% of the streaming time at the
        # 5.0 POP/s dense int8 peak)
        ops = 64 * N * cfg["K"]
        roof.update({"int8_ops_per_launch": ops,
                     "int8_pops": ops / (r["stage_ms"] / 1e3) / 1e15, "int8_peak_pops": 5.0})
    if cfg["storage"] == "rebuild":
        # (K products + K fmas to form M_ij, 2 fmas into C and S) per pair:
        # 3K + 4 flops in 2K + 3 VALU instructions per (i, j) (M_ij converted
        # to double once; the sums accumulate in double, as the dense path's),
        # against the published FP32 vector peak (157.3 TFLOP/s, all-FMA mix)
        flops = (3 * cfg["K"] + 4) * N * N
        tf = flops / (r["stage_ms"] / 1e3) / 1e12
        roof = {"bound": "valu", "achieved": tf, "peak": 157.3, "unit": "TFLOP/s", "frac": tf / 157.3,
                "flops_per_launch": flops,
                "valu_instructions_per_launch": (2 * cfg["K"] + 3) * N * N}
    out = {
        "metric": "osc*steps/s (N x accepted DP45 steps per second)",
        "value": N * accepted / secs,
        "unit": "osc*steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["region_ms"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if np.dtype(r["dtype"]).itemsize == 8 else "f32",
        "data": "synthetic (Philox Gaussian factor, glibc-rand frequencies/phases, as the CLI generates)",
        "config": {"workload": f"{args.config}: N={N}, K={cfg['K']}, "
                               + ("low-rank factor only (-A)" if cfg["storage"] == "factored" else
                                  "-A, tiles rebuilt in registers" if cfg["storage"] == "rebuild" else
                                  "-A, volcano signs packed 1 bit per entry" if cfg["storage"] == "signs" else
                                  f"symmetric adjacency, upper block triangle ({r['elem_bytes']} B/elem)"
                                  if cfg["storage"].startswith("sym") else
                                  f"dense adjacency {cfg['storage']} ({r['elem_bytes']} B/elem)")
                               + f", atol={cfg['atol']}, c={cfg['c1']}",
                   "N": N, "accepted_in_region": accepted,
                   "rejected_in_region": r["st"].rejected - r["st0"].rejected,
                   "rhs_evals_in_region": r["st"].rhs_evals - r["st0"].rhs_evals,
                   "t_after": r["st"].t, "adjacency_build_s": round(r["gen_s"], 3)},
        "roofline": {**roof,
                     "traffic": profiled_traffic(args.config, "lowrank_stage_kernel" if cfg["storage"] == "factored"
                                                 else "sym_tile_kernel" if cfg["storage"].startswith("sym")
                                                 else "sign_mfma_kernel" if cfg["storage"] == "signs"
                                                 else "stage_ring_kernel" if r["ring"] else "stage_kernel"),
                     "kernel": ("sign_mfma_kernel x 2 (int8 MFMA passes over the bits) + sign_limbs / sign_gram / "
                                "sign_stage kernels (one RHS)"
                                if cfg["storage"] == "signs" else
                                "lowrank_stage_kernel (+ gram at stage 2; one RHS)" if cfg["storage"] == "factored"
                                else "rebuild_stage_kernel (tiles rebuilt in registers)" if cfg["storage"] == "rebuild"
                                else "sym_tile_kernel + sym_reduce_kernel (one RHS from the upper block triangle)"
                                if cfg["storage"].startswith("sym")
                                else "stage_ring_kernel (fused gemv-pair + DP45 stage epilogue, chunk ring)"
                                if r["ring"] else "stage_kernel (fused gemv-pair + DP45 stage epilogue)"),
                     "bytes_per_launch": alg_bytes, "avg_launch_ms": r["stage_ms"],
                     "variant": ({"block_edge": r["lda"]} if cfg["storage"].startswith("sym") else
                                 {"rows_per_group": r["rows"], "nontemporal_loads": r["nt"],
                                  "balanced_grid": r["balanced"], "chunk_ring": r["ring"]}),
                     "measured_read_stream_gbps": r["read_gbps"], "measured_copy_gbps": r["copy_gbps"],
                     "frac_of_measured_read_stream": achieved / 1e9 / r["read_gbps"]},
    }
    if args.cpu_steps > 0:
        try:
            out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_steps)
        except Exception as e:  # reported, never fatal for the GPU figure
            out["cpu_baseline"] = {"value": None, "error": repr(e)}
    # after the CPU baseline: the GPU has idled for its duration, as before
    # the headline run (the stage kernels stream 2-5% slower right after
    # minutes of load)
    if args.config == "fp64-n50000" and not args.no_secondary:
        try:
            out["secondary"] = secondary_measurements(args.steps, args.warmup)
        except Exception as e:  # reported, never fatal for the headline figure
            out["secondary"] = {"error": repr(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
