"""The reference's -A volcano factor at high rank (kuramoto.cu:62-143), drawn on
the device and kept as packed signs (kmt_set_factor_sign), through the C ABI.

* The coupling through the packed signs against the oracle's factored path
  (kuramoto.cu:62-101 restated: G_m = w_m sum_n X_nm v_n, C_n = sum_m X_nm G_m)
  on the host generator's signs (kmt_factor(N, K, seed, 0)), K = 17 .. 5000,
  K > N, K = 0, a single row, and sizes on the 64-row / 1024-column tile
  boundaries of the passes' bit layout: FP32 bit-identical in >= 99% of the values and within 2.5e-7 of
  the row scale |w_i| + c K everywhere (both sides sum in double and round
  once); FP64 within 1e-13 of that scale.
* The dense matrix built from the packed signs equals the oracle generator's
  rows bit for bit (the same fma chain), in the working precision and, in
  BF16, the matrix built from the host factor.
* The published high-rank run (volcano.ipynb:55-60: kuramoto -N 10000
  -K 10000 -c 1.75 -d 0.01 -s 1 -n -A, theta0 = 0) to t = 1 against the FP32
  oracle's -A run on the same signs: phases at t=1 within 1e-5 relative to
  max(1, |theta|), accepted steps within 2% (or 1), r(t) within 1e-3; the
  dense form of the same network (the matrix built from the signs) against
  the same oracle run: steps within 2% (or 1), r(t) within 1e-3.
* At the largest published size, N = K = 10^5 (no oracle can hold that
  factor), the symmetry invariants sum_i (sin t_i C_i - cos t_i S_i) = 0 and
  sum_i f_i = sum_i w_i within 1e-8 (FP32) / 1e-14 (FP64) of the scale.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _phase_err(a, b):
    b64 = np.asarray(b, np.float64)
    return float(np.max(np.abs(np.asarray(a, np.float64) - b64) / np.maximum(1.0, np.abs(b64))))


@pytest.mark.parametrize("N,K", [(1000, 17), (3001, 300), (257, 1000), (5000, 5000), (1, 2049), (64, 1024), (130, 0)])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_sign_factor_rhs_matches_oracle(km, orc, N, K, prec):
    dtype = np.float32 if prec == "f32" else np.float64
    p = km.F32 if prec == "f32" else km.F64
    seed, c1 = 7, 1.75
    rng = np.random.default_rng(N + K)
    omega = rng.standard_normal(N).astype(dtype)
    theta = rng.uniform(-30, 30, N).astype(dtype)
    print(f"case sign factor rhs N={N} K={K} {prec}", flush=True)
    with km.Integrator(N, p, c1=c1) as it:
        it.set_omega(omega)
        it.set_factor_sign(K, seed)
        it.use_factored(True)
        f, cp = it.rhs(theta, couplings=True)
    X = orc.rademacher_factor(N, K, seed).astype(dtype)
    w = orc.factor_weights(N, K, False, dtype)
    fo, cso = orc.rhs(orc.Adj(2, X=X, w=w), omega, c1, theta, dtype=dtype)
    scale = np.abs(omega.astype(np.float64)) + c1 * K
    tol = (2.5e-7 if prec == "f32" else 1e-13) * scale
    assert np.all(np.abs(f.astype(np.float64) - fo) <= tol)
    assert np.all(np.abs(cp.astype(np.float64) - cso) <= np.repeat(tol, 2))
    if prec == "f32":
        assert np.count_nonzero(f == fo) >= 0.99 * N and np.count_nonzero(cp == cso) >= 0.99 * 2 * N


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_sign_factor_dense_build_bit_exact(km, orc, prec):
    N, K, seed = 700, 100, 5
    dtype = np.float32 if prec == "f32" else np.float64
    p = km.F32 if prec == "f32" else km.F64
    X = orc.rademacher_factor(N, K, seed).astype(dtype)
    w = orc.factor_weights(N, K, False, dtype)
    print(f"case sign factor dense build {prec}", flush=True)
    with km.Integrator(N, p) as it:
        it.set_factor_sign(K, seed)
        it.build_adjacency()
        M = it.adjacency_rows(0, N)
        it.build_adjacency("bf16")
        Mb, sb = it.adjacency_rows(0, N), it.adjacency_scale()
    assert np.array_equal(M, orc.adj_rows(X, w, 0, N))
    with km.Integrator(N, p) as it:
        it.set_factor(X, km.factor_weights(p, N, K, gaussian=False))
        it.build_adjacency("bf16")
        assert np.array_equal(Mb, it.adjacency_rows(0, N)) and sb == it.adjacency_scale()


def test_sign_factor_published_run_n10000_k10000(km, orc):
    N, K, seed, c1, dt, t1, atol = 10000, 10000, 1, 1.75, 0.01, 1.0, 1e-6
    theta0, omega = km.draws(km.F32, N, seed, normal=True, freq=1.0, init=0.0)
    te = (np.float32(dt) * np.arange(1, int(np.float32(t1) / np.float32(dt)) + 1)).astype(np.float32)
    h0 = float(np.float32(dt) / np.float32(100))
    res = {}
    for mode in ("factored", "dense"):
        print(f"case N=K=10000 -A published run, GPU {mode}", flush=True)
        with km.Integrator(N, km.F32, atol=atol, c1=c1) as it:
            it.set_omega(omega)
            it.set_factor_sign(K, seed)
            if mode == "factored":
                it.use_factored(True)
            else:
                it.build_adjacency()
            it.set_state(theta0, 0.0, h0)
            res[mode] = it.run(te.astype(np.float64), t1, keep_rows=True, ring_rows=128, chunk=64)
    print("  oracle -A on the same signs", flush=True)
    X = orc.rademacher_factor(N, K, seed).astype(np.float32)
    w = orc.factor_weights(N, K, False, np.float32)
    ref = orc.integrate(orc.Adj(2, X=X, w=w), omega, c1, theta0, 0.0, h0, t1, atol, t_eval=te, dtype=np.float32)
    del X
    g, d = res["factored"], res["dense"]
    e1 = _phase_err(g["rows"][-1], ref["rows"][-1])
    print(f"  accepted gpu -A={g['status'].accepted} dense={d['status'].accepted} oracle={ref['accepted']}; "
          f"phase err t=1 {e1:.3e}; max|dr| -A {np.max(np.abs(g['r'] - ref['r'])):.3e} "
          f"dense {np.max(np.abs(d['r'] - ref['r'])):.3e}", flush=True)
    for run in (g, d):
        assert run["status"].error == 0 and run["status"].active == 0 and len(run["r"]) == len(te)
        assert abs(run["status"].accepted - ref["accepted"]) <= max(1, int(0.02 * ref["accepted"]))
        assert np.max(np.abs(run["r"] - ref["r"])) < 1e-3
    assert e1 < 1e-5


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_sign_factor_full_size_symmetry(km, prec):
    """The reference's largest published -A network, N = K = 10^5
    (volcano.ipynb:319-324), where the oracle cannot hold the factor (10^10
    signs): M = X diag(w) X^T is symmetric, so sum_i (sin t_i C_i - cos t_i S_i)
    = sum_ij M_ij sin(t_i - t_j) = 0 and sum_i f_i = sum_i w_i.  The packed
    path sums exactly in integers and rounds C_i, S_i once, so the residual is
    that rounding (and the device's cos / sin against the host's); a pass
    that used the signs other than transposed to the first (an effective M
    that is not symmetric) leaves a residual of order sum |C| / sqrt(N)."""
    N = K = 100000
    p = km.F32 if prec == "f32" else km.F64
    dtype = np.float32 if prec == "f32" else np.float64
    c1 = 3.0
    rng = np.random.default_rng(11)
    omega = rng.standard_normal(N).astype(dtype)
    theta = rng.uniform(-30, 30, N).astype(dtype)
    print(f"case sign factor full size N=K={N} {prec}", flush=True)
    with km.Integrator(N, p, c1=c1) as it:
        it.set_omega(omega)
        it.set_factor_sign(K, 1)
        it.use_factored(True)
        f, cp = it.rhs(theta, couplings=True)
    C, S = cp[0::2].astype(np.float64), cp[1::2].astype(np.float64)
    t = theta.astype(np.float64)
    scale = float(np.sum(np.abs(C)) + np.sum(np.abs(S)))
    inv = float(np.sum(np.sin(t) * C - np.cos(t) * S))
    # FP32: the float rounding of every C_i, S_i summed with random signs,
    # ~6e-8 sqrt(N) |C| ~ 1e-10 of the scale (measured 1e-11); FP64 1e-17
    tol = 1e-8 if prec == "f32" else 1e-14
    print(f"  sum|C|+|S| = {scale:.4e}, antisymmetric sum = {inv:.3e}", flush=True)
    assert scale > 0 and abs(inv) <= tol * scale
    fsum = float(np.sum(f.astype(np.float64) - omega.astype(np.float64)))
    assert abs(fsum) <= tol * (float(np.sum(np.abs(omega.astype(np.float64)))) + c1 * scale)
