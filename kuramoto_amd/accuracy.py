"""Error of a reduced-precision or batched run against an FP64 run of the
same problem, at full size.

The FP64 reference is the library's own FP64 dense path (the kuramoto_64
arithmetic) at atol 1e-9, which the GPU tests hold to the CPU FP64 oracle
within 1e-10 in the phases; at these sizes (N up to 200000, t up to 100) the
CPU oracle would take hours, so the FP64 GPU path stands in for it here.

For every checked run: max |r(t) - r_64(t)| over the output grid, the max
relative phase error at t1 (|dtheta| / max(1, |theta_64|)), and the accepted
steps of both.  Storage-reduced runs (BF16/FP16 A) are compared with FP64 on
the same dequantized matrix (arithmetic error) and with FP64 on the exact
generated matrix (arithmetic + storage error).

    python -m kuramoto_amd.accuracy plain fp32-n50000 [--storage bf16] [--t1 100]
    python -m kuramoto_amd.accuracy storage64 [--t1 10]
    python -m kuramoto_amd.accuracy batch64 [--instances 0,9,18,...] [--t1 50]
    python -m kuramoto_amd.accuracy batch48 [--instances ...] [--t1 100]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import kuramoto_amd as km  # noqa: E402
from kuramoto_amd import workloads  # noqa: E402


def _grid(t1, dt):
    dt = np.float32(dt)
    return (dt * np.arange(1, int(np.float32(t1) / dt) + 1)).astype(np.float64)


# above this many bytes of FP64 matrix the reference couples through the
# factor instead (the same matrix, contracted as X diag(w) (X^T v))
DENSE64_LIMIT = 100e9


def fp64_run(M64, X, w, omega, theta0, c1, t1, dt, h0, atol=1e-9):
    """FP64 run; M64 = explicit matrix (N x N float64) or None to use the
    factor (dense matrix generated on the device, or the factored
    contraction when the FP64 matrix would exceed DENSE64_LIMIT)."""
    N = omega.size
    with km.Integrator(N, km.F64, atol=atol, c1=c1) as it:
        it.set_omega(omega.astype(np.float64))
        if M64 is None:
            it.set_factor(X.astype(np.float64), w.astype(np.float64))
            if 8.0 * N * N > DENSE64_LIMIT:
                it.use_factored("contract")
            else:
                it.build_adjacency()
        else:
            it.set_adjacency(M64)
        it.set_state(theta0.astype(np.float64), 0.0, h0)
        res = it.run(_grid(t1, dt), t1, keep_rows=False, ring_rows=256, chunk=64)
        yf = it.eval(float(_grid(t1, dt)[-1]))
    return dict(r=res["r"], final=yf, accepted=res["status"].accepted, rejected=res["status"].rejected)


def compare(r, final, ref):
    n = min(len(r), len(ref["r"]))
    scale = np.maximum(1.0, np.abs(ref["final"]))
    return dict(max_dr=float(np.max(np.abs(r[:n].astype(np.float64) - ref["r"][:n]))),
                max_rel_dtheta_t1=float(np.max(np.abs(final.astype(np.float64) - ref["final"]) / scale)),
                accepted_fp64=int(ref["accepted"]))


def plain(name, t1=None, storage=None):
    cfg = dict(workloads.PLAIN[name])
    if cfg["prec"] != "f32":
        raise ValueError(f"{name}: the check covers FP32 workloads")
    if storage:
        cfg["storage"] = storage
    N, K = cfg["N"], cfg["K"]
    t1 = t1 or cfg["t1"]
    th, om = km.draws(km.F32, N, cfg["seed"], normal=cfg["normal"], init=cfg["init"])
    X = km.factor(N, K, cfg["seed"], gaussian=True)
    w32 = km.factor_weights(km.F32, N, K, gaussian=True)
    h0 = float(np.float32(cfg["dt"]) / np.float32(100))
    t0 = time.time()
    with km.Integrator(N, km.F32, atol=cfg["atol"], c1=cfg["c1"]) as it:
        it.set_omega(om)
        it.set_factor(X.astype(np.float32), w32)
        if cfg["storage"] == "factored":
            it.use_factored("contract")
        elif cfg["storage"] == "rebuild":
            it.use_factored("rebuild")
        else:
            it.build_adjacency(cfg["storage"])
        same = cfg["storage"] in ("bf16", "f16") and 8.0 * N * N <= DENSE64_LIMIT
        Mq = it.adjacency_rows(0, N).astype(np.float64) if same else None
        it.set_state(th, 0.0, h0)
        res = it.run(_grid(t1, cfg["dt"]), t1, keep_rows=False, ring_rows=256, chunk=64)
        fin = it.eval(float(_grid(t1, cfg["dt"])[-1]))
    out = dict(config=name, storage=cfg["storage"], t1=t1, accepted=res["status"].accepted,
               rejected=res["status"].rejected)
    w64 = km.factor_weights(km.F64, N, K, gaussian=True)
    print(f"{name}: FP32 run done ({time.time() - t0:.1f} s); FP64 run on the exact matrix", flush=True)
    ref = fp64_run(None, X, w64, om, th, cfg["c1"], t1, cfg["dt"], h0)
    out["vs_fp64_exact_matrix"] = compare(res["r"], fin, ref)
    if Mq is not None:
        refq = fp64_run(Mq, None, None, om, th, cfg["c1"], t1, cfg["dt"], h0)
        out["vs_fp64_same_stored_matrix"] = compare(res["r"], fin, refq)
    out["seconds"] = time.time() - t0
    return out


def storage64(t1=None):
    """The FP64 N=50000 workload (kuramoto_64 arithmetic) with each adjacency
    storage against the dense FP64 storage of the same factor: symmetric FP64
    (the same sums in another order), dense and symmetric BF16 (FP32 sums of
    BF16 entries; FP64 state, stage combinations and norm)."""
    cfg = workloads.PLAIN["fp64-n50000"]
    N, K, c1 = cfg["N"], cfg["K"], cfg["c1"]
    t1 = t1 or 10.0
    th, om = km.draws(km.F64, N, cfg["seed"], normal=cfg["normal"], init=cfg["init"])
    X = km.factor(N, K, cfg["seed"], gaussian=True)
    w = km.factor_weights(km.F64, N, K, gaussian=True)
    te = _grid(t1, cfg["dt"]).astype(np.float64)
    runs = {}
    for storage in ("native", "sym", "bf16", "sym-bf16"):
        with km.Integrator(N, km.F64, atol=cfg["atol"], c1=c1) as it:
            it.set_omega(om)
            it.set_factor(X, w)
            it.build_adjacency(storage)
            it.set_state(th, 0.0, 0.01)
            res = it.run(te, t1, keep_rows=False, ring_rows=64)
            y, _, _ = it.get_state()
        runs[storage] = (res["r"], y, res["status"].accepted)
        print(f"storage64 {storage}: {res['status'].accepted} accepted steps", flush=True)
    r0, y0, a0 = runs["native"]
    out = dict(config="fp64-n50000", t1=t1, accepted_dense_fp64=a0)
    for k in ("sym", "bf16", "sym-bf16"):
        r, y, acc = runs[k]
        out[k] = dict(max_dr=float(np.max(np.abs(r - r0))),
                      max_rel_phase=float(np.max(np.abs(y - y0) / np.maximum(1.0, np.abs(y0)))), accepted=acc)
    out["sym-bf16_vs_bf16_max_dr"] = float(np.max(np.abs(runs["sym-bf16"][0] - runs["bf16"][0])))
    return out


def batch(which, instances, t1, ref_atol=1e-9, exact_only=False):
    name = "batch64-n20000-bf16" if which == "batch64" else "batch48-uneven-bf16"
    cfg = workloads.BATCH[name]
    t1 = t1 or cfg["t1"]
    bt, Ns, _ = workloads.make_batch(name, proposer=not cfg["shared"])
    te = _grid(t1, cfg["dt"])
    t0 = time.time()
    res = bt.run(te, t1, ring_rows=1024, chunk=64)
    finals = [bt.get_instance(b) for b in range(cfg["B"])]
    print(f"{name}: batched run to t={t1:g} done in {time.time() - t0:.1f} s", flush=True)
    Mq_shared = bt.adjacency_rows(-1, 0, Ns[0]).astype(np.float64) if cfg["shared"] else None
    rows = []
    for b in instances:
        N = Ns[b]
        seed_x = 1 if cfg["shared"] else cfg["seeds"][b]
        th, om = km.draws(km.F32, N, cfg["seeds"][b], normal=cfg["normal"][b], init=6.283185307)
        X = km.factor(N, 5, seed_x, gaussian=True)
        w64 = km.factor_weights(km.F64, N, 5, gaussian=True)
        Mq = Mq_shared if cfg["shared"] else bt.adjacency_rows(b, 0, N).astype(np.float64)
        c1 = float(cfg["cs"][b])
        st_b = finals[b][1]
        t0 = time.time()
        print(f"instance {b} (N={N}): FP64 run on the exact matrix", flush=True)
        ref = fp64_run(None, X, w64, om, th, c1, t1, cfg["dt"], 0.01, atol=ref_atol)
        refq = None
        if not exact_only:
            print(f"instance {b}: FP64 run on the stored matrix ({time.time() - t0:.1f} s so far)", flush=True)
            refq = fp64_run(Mq, None, None, om, th, c1, t1, cfg["dt"], 0.01, atol=ref_atol)
        row = dict(instance=b, N=N, c=c1, accepted=int(st_b["accepted"]), rejected=int(st_b["rejected"]))
        rb = res["r"][b]
        for key, rr in (("vs_fp64_exact_matrix", ref), ("vs_fp64_same_stored_matrix", refq)):
            if rr is None:
                continue
            n = min(len(rb), len(rr["r"]))
            row[key] = dict(max_dr=float(np.max(np.abs(rb[:n].astype(np.float64) - rr["r"][:n]))),
                            accepted_fp64=int(rr["accepted"]))
        rows.append(row)
        print(json.dumps(row), flush=True)
    bt.close()
    out = dict(config=name, t1=t1, reference_atol=ref_atol, instances=rows,
               max_dr_exact=max(r["vs_fp64_exact_matrix"]["max_dr"] for r in rows))
    if not exact_only:
        out["max_dr_same_matrix"] = max(r["vs_fp64_same_stored_matrix"]["max_dr"] for r in rows)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["plain", "storage64", "batch64", "batch48"])
    ap.add_argument("config", nargs="?", default="fp32-n50000")
    ap.add_argument("--instances", default=None, help="comma-separated instance indices (batch modes)")
    ap.add_argument("--t1", type=float, default=None)
    ap.add_argument("--storage", default=None, help="plain mode: native, bf16 or f16 instead of the workload's")
    ap.add_argument("--ref-atol", type=float, default=1e-9, help="atol of the FP64 reference runs")
    ap.add_argument("--exact-only", action="store_true", help="batch modes: skip the FP64 run on the stored matrix")
    args = ap.parse_args()
    if args.mode == "plain":
        print(json.dumps(plain(args.config, args.t1, args.storage)), flush=True)
    elif args.mode == "storage64":
        print(json.dumps(storage64(args.t1)), flush=True)
    else:
        B = 64 if args.mode == "batch64" else 48
        if args.instances == "all":
            inst = list(range(B))
        else:
            inst = [int(x) for x in args.instances.split(",")] if args.instances else list(range(0, B, max(1, B // 8)))
        print(json.dumps(batch(args.mode, inst, args.t1, args.ref_atol, args.exact_only)), flush=True)


if __name__ == "__main__":
    main()
