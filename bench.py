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
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

CONFIGS = {
    # the FP64 plain path, the memory-bound gemv-roofline case
    "fp64-n50000": dict(N=50000, K=5, seed=2, prec="f64", normal=True, init=6.283185307, c1=3.0, atol=1e-9,
                        t1=100.0, dt=1.0, storage="native"),
    # the north-star roofline target: N=50000 FP32 dense
    "fp32-n50000": dict(N=50000, K=5, seed=2, prec="f32", normal=True, init=6.283185307, c1=3.0, atol=1e-6,
                        t1=100.0, dt=1.0, storage="native"),
    # the reference-defaults FP32 case (MALL-resident A)
    "fp32-n5000": dict(N=5000, K=5, seed=1, prec="f32", normal=False, init=0.0, c1=3.0, atol=1e-6, t1=100.0,
                       dt=1.0, storage="native"),
}

HBM_PEAK = 8.0e12


def profiled_traffic(config):
    """HBM bytes per stage launch measured by the rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs recorded in profiles/stage_kernel_<config>.txt (None if
    that workload has not been profiled)."""
    path = os.path.join(ROOT, "profiles", "stage_kernel_" + config.replace("-", "_") + ".txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def gpu_run(cfg, steps, warmup):
    import kuramoto_amd as km

    p = km.F32 if cfg["prec"] == "f32" else km.F64
    dtype = np.float32 if p == km.F32 else np.float64
    N, K = cfg["N"], cfg["K"]
    theta0, omega = km.draws(p, N, cfg["seed"], normal=cfg["normal"], freq=1.0, init=cfg["init"])
    X = km.factor(N, K, cfg["seed"], gaussian=True).astype(dtype)
    w = km.factor_weights(p, N, K, gaussian=True)
    dt = dtype(cfg["dt"])
    te = (dt * np.arange(1, int(dtype(cfg["t1"]) / dt) + 1)).astype(np.float64)
    it = km.Integrator(N, p, atol=cfg["atol"], c1=cfg["c1"])
    it.set_omega(omega)
    it.set_factor(X, w)
    t0 = time.time()
    it.build_adjacency(cfg["storage"])
    gen_s = time.time() - t0
    it.set_state(theta0, 0.0, float(dt / dtype(100)))
    it.begin(te, cfg["t1"], keep_rows=False, ring_rows=4096)
    if warmup:
        it.advance(warmup)
    it.read_times()
    it.read_outputs()
    st0 = it.advance(0)
    region_ms, stage_ms, st = it.time_attempts(steps)
    lda, eb, rows = it.layout()
    it.close()
    copy_gbps, read_gbps = km.bandwidth_probe(8 << 30, 5)
    return dict(dtype=dtype, N=N, region_ms=region_ms, stage_ms=stage_ms, st0=st0, st=st, lda=lda, elem_bytes=eb,
                gen_s=gen_s, rows=abs(rows), nt=rows > 0, copy_gbps=copy_gbps, read_gbps=read_gbps)


def cpu_baseline(cfg, max_steps):
    """Oracle (FP64 transcription, OpenMP) on the same inputs, first
    `max_steps` accepted steps."""
    import kuramoto_amd as km
    from oracle import oracle_py as orc

    N, K = cfg["N"], cfg["K"]
    p = km.F32 if cfg["prec"] == "f32" else km.F64
    dtype = np.float32 if p == km.F32 else np.float64
    theta0, omega = orc.draws(N, cfg["seed"], cfg["normal"], 1.0, cfg["init"], dtype=dtype)
    X = orc.gaussian_factor(N, K, cfg["seed"]).astype(dtype)
    w = orc.factor_weights(N, K, True, dtype=dtype)
    M = orc.adj_rows(X, w, 0, N)
    dt = dtype(cfg["dt"])
    te = (dt * np.arange(1, int(dtype(cfg["t1"]) / dt) + 1)).astype(dtype)
    t0 = time.time()
    res = orc.integrate(orc.Adj(0, M=M), omega, cfg["c1"], theta0, 0.0, float(dt / dtype(100)), cfg["t1"],
                        cfg["atol"], t_eval=te, keep_rows=False, max_steps=max_steps, dtype=dtype)
    secs = time.time() - t0
    del M
    return dict(value=N * res["accepted"] / secs, unit="osc*steps/s", cores=orc.num_threads(), kind="port",
                sample=f"oracle {np.dtype(dtype).name} dense, N={N}, first {res['accepted']} accepted steps "
                       f"({res['rejected']} rejected, {res['rhs_evals']} RHS evals) in {secs:.1f} s, "
                       f"OpenMP threads={orc.num_threads()}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="fp64-n50000", choices=sorted(CONFIGS))
    ap.add_argument("--cpu-steps", type=int, default=20, help="accepted steps of the CPU baseline sample (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported (see DESIGN.md, "
              "'How the path would shard')", file=sys.stderr)
        sys.exit(2)
    cfg = CONFIGS[args.config]
    r = gpu_run(cfg, args.steps, args.warmup)
    N = r["N"]
    s_y = np.dtype(r["dtype"]).itemsize
    accepted = r["st"].accepted - r["st0"].accepted
    secs = r["region_ms"] / 1e3
    alg_bytes = r["elem_bytes"] * N * N + N * (2 * s_y + s_y)
    achieved = alg_bytes / (r["stage_ms"] / 1e3)
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
        "dtype": "f64" if s_y == 8 else "f32",
        "data": "synthetic (Philox Gaussian factor, glibc-rand frequencies/phases, as the CLI generates)",
        "config": {"workload": f"{args.config}: N={N}, K={cfg['K']}, dense adjacency {cfg['storage']} "
                               f"({r['elem_bytes']} B/elem), atol={cfg['atol']}, c={cfg['c1']}",
                   "N": N, "accepted_in_region": accepted,
                   "rejected_in_region": r["st"].rejected - r["st0"].rejected,
                   "t_after": r["st"].t, "adjacency_build_s": round(r["gen_s"], 3)},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": profiled_traffic(args.config),
                     "kernel": "stage_kernel (fused gemv-pair + DP45 stage epilogue)",
                     "bytes_per_launch": alg_bytes, "avg_launch_ms": r["stage_ms"],
                     "variant": {"rows_per_workgroup": r["rows"], "nontemporal_loads": r["nt"]},
                     "measured_read_stream_gbps": r["read_gbps"], "measured_copy_gbps": r["copy_gbps"],
                     "frac_of_measured_read_stream": achieved / 1e9 / r["read_gbps"]},
    }
    if args.cpu_steps > 0:
        try:
            out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_steps)
        except Exception as e:  # reported, never fatal for the GPU figure
            out["cpu_baseline"] = {"value": None, "error": repr(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
