"""GPU parity of the symmetric-storage path (KMT_ADJ_SYM): the upper block
triangle of a symmetric adjacency in 256 x 256 blocks, one block kernel
(row sums, and column sums standing in for the lower triangle) and one row
reduction + stage epilogue per RHS.  The reference reads the full matrix in
its two gemvs (kuramoto.cu:104-126); only the summation order differs.

* The stored triangle, read back as full rows, equals the dense adjacency
  of the same factor bit for bit (the generator's formula on both sides).
  N on both sides of the block edge.
* A symmetric host matrix packed by kmt_set_adjacency reads back unchanged;
  a non-symmetric one is refused (KMT_ERR_ARG).
* RHS and couplings against the oracle, with the dense path's tolerances
  (FP64 1e-12 of the row scale sum_j |M_ij|; FP32 sums accumulate in double
  and round once, so within one float ulp, 1.2e-7 of the row scale).
* Whole integrations against the oracle, with the dense path's tolerances.
  Two runs agree bit for bit.
* At N = 50000 in FP64 (10 GB stored instead of 20 GB), the RHS equals the
  dense path's to 1e-12 of the row scale.  The symmetric-M invariant
  sum_i f_i = sum_i w_i holds, and so does the drift invariant through
  whole steps.
"""
import numpy as np
import pytest

from kuramoto_amd.workloads import PLAIN

pytestmark = pytest.mark.gpu


def _grid(t1, dt, dtype):
    dt = dtype(dt)
    return (dt * np.arange(1, int(dtype(t1) / dt) + 1)).astype(dtype)


def _factor_case(km, prec, N, seed, K=5):
    dtype = np.float32 if prec == km.F32 else np.float64
    theta0, omega = km.draws(prec, N, seed, normal=True, init=6.283185307)
    X = km.factor(N, K, seed, gaussian=True).astype(dtype)
    w = km.factor_weights(prec, N, K, gaussian=True)
    return dtype, theta0, omega, X, w


@pytest.mark.parametrize("N", [1, 255, 256, 257, 600])
@pytest.mark.parametrize("prec,storage", [("f32", ""), ("f64", ""), ("f32", "bf16"), ("f64", "bf16"), ("f64", "f16")])
def test_sym_storage_equals_dense(km, prec, storage, N):
    """Working precision and 16-bit storage (same power-of-two scale and
    rounding as the dense 16-bit matrix)."""
    p = km.F32 if prec == "f32" else km.F64
    dtype, theta0, omega, X, w = _factor_case(km, p, N, 7)
    print(f"case sym storage N={N} {prec} {storage}", flush=True)
    with km.Integrator(N, p, c1=3.0) as it:
        it.set_factor(X, w)
        it.build_adjacency(storage or "native")
        dense = it.adjacency_rows(0, N)
        it.build_adjacency("sym-" + storage if storage else "sym")
        packed = it.adjacency_rows(0, N)
        it.build_adjacency(storage or "native")  # and back: the block buffers are released
        again = it.adjacency_rows(0, N)
    assert np.array_equal(dense, packed) and np.array_equal(dense, again)
    assert np.array_equal(packed, packed.T)


def test_sym_upload_packs_and_checks_symmetry(km):
    N = 300
    rng = np.random.default_rng(5)
    M = rng.standard_normal((N, N))
    M = M + M.T
    print("case sym upload", flush=True)
    with km.Integrator(N, km.F64, c1=1.0) as it:
        it.set_adjacency(M, "sym")
        assert np.array_equal(it.adjacency_rows(0, N), M)
        M[3, 200] += 1e-9
        with pytest.raises(km.KmtError, match="symmetric"):
            it.set_adjacency(M, "sym")


@pytest.mark.parametrize("N", [1, 2, 65, 255, 257, 1000, 4097])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_sym_rhs_matches_oracle(km, orc, N, prec):
    dtype = np.float32 if prec == "f32" else np.float64
    p = km.F32 if prec == "f32" else km.F64
    rng = np.random.default_rng(11 + N)
    M = rng.standard_normal((N, N)) / max(N, 1)
    M = (0.5 * (M + M.T)).astype(dtype)
    omega = rng.standard_normal(N).astype(dtype)
    theta = (rng.uniform(-np.pi, np.pi, N) * 3).astype(dtype)
    c1 = 1.75
    print(f"case sym rhs N={N} {prec}", flush=True)
    with km.Integrator(N, p, c1=c1) as it:
        it.set_omega(omega)
        it.set_adjacency(M, "sym")
        f, cp = it.rhs(theta, couplings=True)
    fo, cso = orc.rhs(orc.Adj(0, M=M), omega, c1, theta, dtype=dtype)
    rs = np.abs(M.astype(np.float64)).sum(axis=1)
    tol = (2.5e-7 if prec == "f32" else 1e-12) * (np.abs(omega.astype(np.float64)) + c1 * rs)
    assert np.all(np.abs(f.astype(np.float64) - fo) <= tol + 1e-30)
    tolc = (1.2e-7 if prec == "f32" else 1e-12) * np.repeat(rs, 2)
    assert np.all(np.abs(cp.astype(np.float64) - cso) <= tolc + 1e-30)
    if prec == "f32":
        assert np.count_nonzero(cp == cso) >= 0.99 * cp.size


@pytest.mark.parametrize("prec,atol", [("f64", 1e-9), ("f32", 1e-6)])
@pytest.mark.parametrize("N", [333, 2000])
def test_sym_integration_matches_oracle(km, orc, prec, atol, N):
    """Same acceptance as the dense path: accepted steps within 2% (or 1),
    r(t) within 1e-3; phases at t=1 within 1e-10 (FP64) / 1e-5 (FP32) of
    the oracle in the same precision (both accumulate the sums in double)."""
    p = km.F32 if prec == "f32" else km.F64
    c1, t1 = 3.0, 1.0
    dtype, theta0, omega, X, w = _factor_case(km, p, N, 2)
    te = _grid(t1, 0.1, dtype)
    print(f"case sym integration N={N} {prec}", flush=True)
    runs = []
    for _ in range(2):
        with km.Integrator(N, p, atol=atol, c1=c1) as it:
            it.set_omega(omega)
            it.set_factor(X, w)
            it.build_adjacency("sym")
            it.set_state(theta0, 0.0, 0.01)
            runs.append(it.run(te, t1, keep_rows=True, ring_rows=8))
    g, g2 = runs
    assert np.array_equal(g["rows"], g2["rows"]) and np.array_equal(g["r"], g2["r"])
    M = orc.adj_rows(X, w, 0, N)
    ref = orc.integrate(orc.Adj(0, M=M), omega, c1, theta0, 0.0, 0.01, t1, atol, t_eval=te, dtype=dtype)
    st = g["status"]
    assert abs(st.accepted - ref["accepted"]) <= max(1, int(0.02 * ref["accepted"]))
    assert np.max(np.abs(g["r"] - ref["r"])) < 1e-3
    ref1 = ref["rows"][-1].astype(np.float64)
    rel = np.abs(g["rows"][-1].astype(np.float64) - ref1) / np.maximum(1.0, np.abs(ref1))
    print(f"  accepted gpu={st.accepted} oracle={ref['accepted']} phase err t=1 {rel.max():.3e}", flush=True)
    assert rel.max() < (1e-10 if prec == "f64" else 1e-5), rel.max()


def test_sym_n50000_fp64_matches_dense_and_invariants(km):
    cfg = PLAIN["fp64-n50000"]
    N, K, c1 = cfg["N"], cfg["K"], cfg["c1"]
    th, om = km.draws(km.F64, N, cfg["seed"], normal=cfg["normal"], init=cfg["init"])
    X = km.factor(N, K, cfg["seed"], gaussian=True)
    w = km.factor_weights(km.F64, N, K, gaussian=True)
    print("case sym N=50000 fp64: RHS vs dense, invariants, 10 attempts", flush=True)
    with km.Integrator(N, km.F64, atol=cfg["atol"], c1=c1) as it:
        it.set_omega(om)
        it.set_factor(X, w)
        it.build_adjacency("native")
        f_dense, cp_dense = it.rhs(th, couplings=True)
        it.build_adjacency("sym")
        f, cp = it.rhs(th, couplings=True)
        # row scale sum_j |M_ij| <= sum_m |w_m| |X_im| sum_j |X_jm|
        rs = np.abs(X) @ (np.abs(w) * np.abs(X).sum(axis=0))
        assert np.all(np.abs(f - f_dense) <= 1e-12 * (np.abs(om) + c1 * rs))
        assert np.all(np.abs(cp - cp_dense) <= 1e-12 * np.repeat(rs, 2))
        assert abs(f.sum() - om.sum()) <= 1e-13 * (np.abs(f).sum() + 1.0)
        it.set_state(th, 0.0, 0.01)
        it.begin(np.arange(1, 11, dtype=np.float64), 10.0, keep_rows=False, ring_rows=64)
        st = it.advance(10)
        y, t, h = it.get_state()
    assert st.error == 0 and st.accepted >= 5
    drift = y.sum() - th.sum() - t * om.sum()
    assert abs(drift) <= 1e-12 * (np.abs(y).sum() + 1.0), drift


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_sym_zero_rank_and_resume(km, prec):
    """K = 0: the stored triangle is zero, so f = omega exactly (closed
    form).  K = 5: 9 attempts in one context equal 4 attempts, a new context
    started from kmt_get_state, and 5 more, bit for bit (the restart path of
    kuramoto.cu:471-494 with symmetric storage)."""
    p = km.F32 if prec == "f32" else km.F64
    N = 700
    dtype, theta0, omega, X, w = _factor_case(km, p, N, 13)
    print(f"case sym K=0 and resume {prec}", flush=True)
    with km.Integrator(N, p, c1=3.0) as it:
        it.set_omega(omega)
        it.set_factor(np.zeros((N, 0), dtype), np.zeros(0, dtype))
        it.build_adjacency("sym")
        assert np.array_equal(it.rhs(theta0), omega)
    te = np.arange(1, 11, dtype=np.float64)

    def run(y0, t0, h0, attempts):
        with km.Integrator(N, p, atol=1e-6 if prec == "f32" else 1e-9, c1=3.0) as it:
            it.set_omega(omega)
            it.set_factor(X, w)
            it.build_adjacency("sym")
            it.set_state(y0, t0, h0)
            it.begin(te[te > t0], 10.0, keep_rows=False, ring_rows=64)
            st = it.advance(attempts)
            assert st.error == 0 and st.active == 1
            return it.get_state()

    y_full, t_full, h_full = run(theta0, 0.0, 0.01, 9)
    y1, t1, h1 = run(theta0, 0.0, 0.01, 4)
    y2, t2, h2 = run(y1, t1, h1, 5)
    assert t2 == t_full and h2 == h_full and np.array_equal(y2, y_full)


@pytest.mark.parametrize("N", [257, 1000, 4097])
@pytest.mark.parametrize("prec,storage", [("f32", "bf16"), ("f32", "f16"), ("f64", "bf16")])
def test_sym16_rhs_matches_dense16(km, prec, storage, N):
    """16-bit symmetric storage, packed from a host matrix: RHS and
    couplings against the dense 16-bit path on the same matrix.  Both
    accumulate in FP32 in different orders, each within 2e-6 of the row
    scale of the exact sums, so they agree within 4e-6."""
    dtype = np.float32 if prec == "f32" else np.float64
    p = km.F32 if prec == "f32" else km.F64
    rng = np.random.default_rng(23 + N)
    M = rng.standard_normal((N, N)) / max(N, 1)
    M = (0.5 * (M + M.T)).astype(dtype)
    omega = rng.standard_normal(N).astype(dtype)
    theta = (rng.uniform(-np.pi, np.pi, N) * 3).astype(dtype)
    c1 = 1.75
    print(f"case sym16 rhs N={N} {prec} {storage}", flush=True)
    with km.Integrator(N, p, c1=c1) as it:
        it.set_omega(omega)
        it.set_adjacency(M, storage)
        f_d, cp_d = it.rhs(theta, couplings=True)
        Mq = it.adjacency_rows(0, N)
        it.set_adjacency(M, "sym-" + storage)
        assert np.array_equal(it.adjacency_rows(0, N), Mq)
        f_s, cp_s = it.rhs(theta, couplings=True)
    rs = np.abs(Mq.astype(np.float64)).sum(axis=1)
    assert np.all(np.abs(f_s.astype(np.float64) - f_d) <= 4e-6 * (np.abs(omega.astype(np.float64)) + c1 * rs))
    assert np.all(np.abs(cp_s.astype(np.float64) - cp_d) <= 4e-6 * np.repeat(rs, 2))


def test_sym_bf16_fp64_integration_follows_dense_bf16(km):
    """kuramoto_64 arithmetic with BF16 symmetric storage against the dense
    BF16 storage of the same factor (FP64 state, FP32 coupling sums in
    another order): r(t) within 1e-5 and accepted steps within 2% (or 1)."""
    N, c1, t1 = 2000, 3.0, 1.0
    dtype, theta0, omega, X, w = _factor_case(km, km.F64, N, 2)
    te = _grid(t1, 0.1, dtype)
    print("case sym-bf16 fp64 integration", flush=True)
    out = {}
    for storage in ("bf16", "sym-bf16"):
        with km.Integrator(N, km.F64, atol=1e-9, c1=c1) as it:
            it.set_omega(omega)
            it.set_factor(X, w)
            it.build_adjacency(storage)
            it.set_state(theta0, 0.0, 0.01)
            out[storage] = it.run(te, t1, keep_rows=False, ring_rows=8)
    a, b = out["bf16"], out["sym-bf16"]
    assert np.max(np.abs(a["r"] - b["r"])) < 1e-5
    assert abs(a["status"].accepted - b["status"].accepted) <= max(1, int(0.02 * a["status"].accepted))

