"""Workload definitions shared by bench.py and kuramoto_amd.accuracy.

Plain runs (one network per context) and batched runs (B networks advanced
in lockstep stages).  Every input is synthetic: the rank-5 Gaussian factor
X ~ N(0, 1) comes from the library's Philox4x32-10 generator (kmt_factor),
the adjacency is A = X X^T / K with the attractive sign (M = -A / N), and
frequencies / initial phases come from the reference's glibc-rand draws
(kmt_draws).
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

PLAIN = {
    # FP64 plain path (the kuramoto_64 arithmetic), the memory-bound gemv case
    "fp64-n50000": dict(N=50000, K=5, seed=2, prec="f64", normal=True, init=6.283185307, c1=3.0, atol=1e-9,
                        t1=100.0, dt=1.0, storage="native"),
    # the same FP64 run with the adjacency stored in BF16 (5 GB): FP32
    # accumulation of the products, FP64 state, stage combinations and norm
    "fp64-n50000-bf16": dict(N=50000, K=5, seed=2, prec="f64", normal=True, init=6.283185307, c1=3.0, atol=1e-9,
                             t1=100.0, dt=1.0, storage="bf16"),
    # FP32 dense at the same size (10 GB adjacency)
    "fp32-n50000": dict(N=50000, K=5, seed=2, prec="f32", normal=True, init=6.283185307, c1=3.0, atol=1e-6,
                        t1=100.0, dt=1.0, storage="native"),
    # the same two runs with the adjacency stored as its upper block triangle
    # (symmetric storage: 10 GB / 5 GB, half the bytes per RHS)
    "fp64-n50000-sym": dict(N=50000, K=5, seed=2, prec="f64", normal=True, init=6.283185307, c1=3.0, atol=1e-9,
                            t1=100.0, dt=1.0, storage="sym"),
    "fp32-n50000-sym": dict(N=50000, K=5, seed=2, prec="f32", normal=True, init=6.283185307, c1=3.0, atol=1e-6,
                            t1=100.0, dt=1.0, storage="sym"),
    # the FP64 run with BF16 symmetric storage (2.5 GB): FP32 coupling sums,
    # FP64 state, stage combinations and norm
    "fp64-n50000-sym-bf16": dict(N=50000, K=5, seed=2, prec="f64", normal=True, init=6.283185307, c1=3.0,
                                 atol=1e-9, t1=100.0, dt=1.0, storage="sym-bf16"),
    # the reference's default FP32 run: N=5000, Cauchy, theta0=0 (A fits the 256 MB cache)
    "fp32-n5000": dict(N=5000, K=5, seed=1, prec="f32", normal=False, init=0.0, c1=3.0, atol=1e-6, t1=100.0,
                       dt=1.0, storage="native"),
    # -A at N=200000: only the N x K factor is stored, coupled through it (O(N K) per RHS)
    "fp32-n200000-lowrank": dict(N=200000, K=5, seed=3, prec="f32", normal=False, init=0.0, c1=3.0, atol=1e-6,
                                 t1=20.0, dt=0.5, storage="factored"),
    # -A at N=200000 with adjacency tiles rebuilt in registers from X at every
    # stage (dense semantics, VALU-bound)
    "fp32-n200000-rebuild": dict(N=200000, K=5, seed=3, prec="f32", normal=False, init=0.0, c1=3.0, atol=1e-6,
                                 t1=20.0, dt=0.5, storage="rebuild"),
    # the reference's published high-rank -A runs (volcano.ipynb:55-60 and
    # :319-324): the volcano factor X = +-1 with K = N, drawn on the device
    # and kept as packed signs (1 bit per entry), w_m = (-1)^m / N
    "fp32-n10000-k10000-signs": dict(N=10000, K=10000, seed=1, prec="f32", normal=True, init=0.0, c1=1.75,
                                     atol=1e-6, t1=100.0, dt=0.01, storage="signs"),
    "fp32-n100000-k100000-signs": dict(N=100000, K=100000, seed=1, prec="f32", normal=True, init=0.0, c1=3.0,
                                       atol=1e-6, t1=20.0, dt=0.02, storage="signs"),
    # the same N=200000 network with A stored explicitly in FP16 (80 GB)
    "fp32-n200000-f16": dict(N=200000, K=5, seed=3, prec="f32", normal=False, init=0.0, c1=3.0, atol=1e-6,
                             t1=20.0, dt=0.5, storage="f16"),
}

# batched configurations (FP32 state, 16-bit adjacency)
BATCH = {
    # B=64 sharing one N=20000 BF16 adjacency, c_b = linspace(0.5, 4, 64),
    # Cauchy frequencies, theta0 = pi(2u-1), seeds 1..64, t=50
    "batch64-n20000-bf16": dict(B=64, Ns=[20000] * 64, shared=True, seeds=list(range(1, 65)),
                                cs=list(np.linspace(0.5, 4.0, 64)), normal=[False] * 64, t1=50.0, dt=1.0),
    # the same sweep with the adjacency in FP16 and the cos/sin vectors of the
    # product in FP16 only (11 significant bits, one MFMA per tile instead of
    # the hi + lo pair that keeps them FP32-class)
    "batch64-n20000-f16-hi": dict(B=64, Ns=[20000] * 64, shared=True, seeds=list(range(1, 65)),
                                  cs=list(np.linspace(0.5, 4.0, 64)), normal=[False] * 64, t1=50.0, dt=1.0,
                                  storage="f16", hilo=False),
    # B=48, N uniform in {5000..40000}, own adjacency each (seeds
    # 100..147), normal/Cauchy alternating, c=3, t=100, with the learned
    # step proposer against the same batch without it
    "batch48-uneven-bf16": dict(B=48, Ns=None, shared=False, seeds=list(range(100, 148)), cs=[3.0] * 48,
                                normal=[b % 2 == 0 for b in range(48)], t1=100.0, dt=1.0),
}

def uneven_sizes(B, seed=147):
    """N_b drawn uniformly from {5000, 10000, ..., 40000} (Philox-free, fixed)."""
    rng = np.random.default_rng(seed)
    return [int(5000 * rng.integers(1, 9)) for _ in range(B)]


def make_batch(name, proposer, tile=None):
    """The batch of workload `name`; the shared product's tile is autotuned,
    or pinned to tile = (workgroup variant, split-K)."""
    import kuramoto_amd as km

    cfg = BATCH[name]
    B = cfg["B"]
    Ns = cfg["Ns"] or uneven_sizes(B)
    bt = km.Batch(Ns, atol=1e-6, shared=cfg["shared"], storage=cfg.get("storage", "bf16"))
    bt.set_vector_split(cfg.get("hilo", True))
    if cfg["shared"]:
        X = km.factor(Ns[0], 5, 1, gaussian=True).astype(np.float32)
        bt.set_factor(-1, X, km.factor_weights(km.F32, Ns[0], 5, gaussian=True))
    for b in range(B):
        th, om = km.draws(km.F32, Ns[b], cfg["seeds"][b], normal=cfg["normal"][b], init=6.283185307)
        bt.set_instance(b, om, th, float(cfg["cs"][b]), 0.0, 0.01)
        if not cfg["shared"]:
            X = km.factor(Ns[b], 5, cfg["seeds"][b], gaussian=True).astype(np.float32)
            bt.set_factor(b, X, km.factor_weights(km.F32, Ns[b], 5, gaussian=True))
    if cfg["shared"] and tile is not None:
        bt.select_tile(*tile)
        tune = (tile[0], tile[1], None)
    else:
        tune = bt.autotune(3) if cfg["shared"] else None
    if proposer:
        w = np.load(os.path.join(HERE, "proposer_weights.npz"))["w16"]
        bt.set_proposer(w, True)
    return bt, Ns, tune


