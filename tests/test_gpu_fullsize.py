"""Parity of the plain path at its full size, N = 50000 with the dense
adjacency in FP64 (20 GB) and in FP32 (10 GB), through checks that do not
need the N x N matrix on the host:

* sampled rows of the RHS against the formula of kuramoto_64.cu:42-55 and
  104-126 (f_i = w_i + c (sin th_i C_i - cos th_i S_i)), with the rows of M
  from the oracle's generator (bit-identical to the device generator) and
  the dot products in numpy in FP64 (another summation order: 1e-12 x row
  scale for FP64, 2e-6 for FP32's float accumulation);
* the symmetric-M invariant sum_i f_i = sum_i w_i (the coupling terms cancel
  pairwise), which a flipped sign, a transposed operand or a row mixed up
  between C and S breaks at O(1);
* the same invariant through whole DP45 steps: every stage k_s sums to
  sum w, and the weights b sum to 1, so sum_i th_i(t) - sum_i th_i(0) =
  (t - t0) sum w up to rounding, however many steps were taken.
"""
import numpy as np
import pytest

from kuramoto_amd.workloads import PLAIN

pytestmark = pytest.mark.gpu


# (workload, RHS row tolerance x row scale, sum invariant, drift; both relative)
CASES = [("fp64-n50000", 1e-12, 1e-13, 1e-12), ("fp32-n50000", 2e-6, 1e-6, 1e-4)]


@pytest.mark.parametrize("name,row_tol,sum_tol,drift_tol", CASES)
def test_n50000_rhs_rows_and_invariants(km, orc, name, row_tol, sum_tol, drift_tol):
    cfg = PLAIN[name]
    N, K, c1 = cfg["N"], cfg["K"], cfg["c1"]
    prec = km.F64 if cfg["prec"] == "f64" else km.F32
    dtype = np.float64 if prec == km.F64 else np.float32
    th, om = km.draws(prec, N, cfg["seed"], normal=cfg["normal"], init=cfg["init"])
    X = km.factor(N, K, cfg["seed"], gaussian=True).astype(dtype)
    w = km.factor_weights(prec, N, K, gaussian=True)
    print(f"case {name} dense: RHS rows, invariant, 10 attempts", flush=True)
    with km.Integrator(N, prec, atol=cfg["atol"], c1=c1) as it:
        it.set_omega(om)
        it.set_factor(X, w)
        it.build_adjacency()
        f = it.rhs(th)
        # sampled rows against the formula on the generated rows
        rows = [0, 1, 2, 777, 12345, 24999, 25000, 31337, 40000, N - 2, N - 1]
        th64, om64, f64 = th.astype(np.float64), om.astype(np.float64), f.astype(np.float64)
        cs, sn = np.cos(th64), np.sin(th64)
        for i in rows:
            Mi = orc.adj_rows(X, w, i, i + 1)[0].astype(np.float64)
            C, S = Mi @ cs, Mi @ sn
            ref = om64[i] + c1 * (sn[i] * C - cs[i] * S)
            scale = c1 * np.abs(Mi).sum() + abs(om64[i])
            assert abs(f64[i] - ref) <= row_tol * scale, i
        # symmetric M: the coupling terms cancel in the sum
        assert abs(f64.sum() - om64.sum()) <= sum_tol * (np.abs(f64).sum() + 1.0)
        # ... and so does every stage of every step
        it.set_state(th, 0.0, 0.01)
        it.begin(np.arange(1, 11, dtype=np.float64), 10.0, keep_rows=False, ring_rows=64)
        st = it.advance(10)
        y, t, h = it.get_state()
    assert st.error == 0 and st.accepted >= 5
    y64 = y.astype(np.float64)
    drift = y64.sum() - th64.sum() - t * om64.sum()
    assert abs(drift) <= drift_tol * (np.abs(y64).sum() + 1.0), drift


def test_fp64_n50000_resume_is_bit_exact(km):
    """Kill and resume at full size (the FILEBASEfs.dat restart, kuramoto.cu:471-494):
    12 attempts in one context equal 5 attempts, a new context started from
    kmt_get_state (y, t, h — the FILEBASEfs.dat contents), and 7 more, bit
    for bit; the 20 GB adjacency is rebuilt by the new context."""
    cfg = PLAIN["fp64-n50000"]
    N, K, c1 = cfg["N"], cfg["K"], cfg["c1"]
    th, om = km.draws(km.F64, N, cfg["seed"], normal=cfg["normal"], init=cfg["init"])
    X = km.factor(N, K, cfg["seed"], gaussian=True)
    w = km.factor_weights(km.F64, N, K, gaussian=True)
    te = np.arange(1, 101, dtype=np.float64)

    def run(y0, t0, h0, attempts):
        with km.Integrator(N, km.F64, atol=cfg["atol"], c1=c1) as it:
            it.set_omega(om)
            it.set_factor(X, w)
            it.build_adjacency()
            it.set_state(y0, t0, h0)
            it.begin(te[te > t0], 100.0, keep_rows=False, ring_rows=64)
            st = it.advance(attempts)
            assert st.error == 0 and st.active == 1
            return it.get_state()

    print("case fp64 N=50000 resume: 12 attempts vs 5 + restart + 7", flush=True)
    y_full, t_full, h_full = run(th, 0.0, 0.01, 12)
    y1, t1, h1 = run(th, 0.0, 0.01, 5)
    y2, t2, h2 = run(y1, t1, h1, 7)
    assert t2 == t_full and h2 == h_full
    assert np.array_equal(y2, y_full)
