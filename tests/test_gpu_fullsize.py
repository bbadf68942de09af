"""Parity of the FP64 plain path at its full size, N = 50000 with the dense
adjacency in FP64 (20 GB), through checks that do not need the N x N matrix
on the host:

* sampled rows of the RHS against the formula of kuramoto_64.cu:42-55 and
  104-126 (f_i = w_i + c (sin th_i C_i - cos th_i S_i)), with the rows of M
  from the oracle's generator (bit-identical to the device generator) and
  the dot products in numpy (another summation order: 1e-12 x row scale);
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


def test_fp64_n50000_rhs_rows_and_invariants(km, orc):
    cfg = PLAIN["fp64-n50000"]
    N, K, c1 = cfg["N"], cfg["K"], cfg["c1"]
    th, om = km.draws(km.F64, N, cfg["seed"], normal=cfg["normal"], init=cfg["init"])
    X = km.factor(N, K, cfg["seed"], gaussian=True)
    w = km.factor_weights(km.F64, N, K, gaussian=True)
    print("case fp64 N=50000 dense (20 GB): RHS rows, invariant, 10 attempts", flush=True)
    with km.Integrator(N, km.F64, atol=cfg["atol"], c1=c1) as it:
        it.set_omega(om)
        it.set_factor(X, w)
        it.build_adjacency()
        f = it.rhs(th)
        # sampled rows against the formula on the generated rows
        rows = [0, 1, 2, 777, 12345, 24999, 25000, 31337, 40000, N - 2, N - 1]
        cs, sn = np.cos(th), np.sin(th)
        for i in rows:
            Mi = orc.adj_rows(X, w, i, i + 1)[0]
            C, S = Mi @ cs, Mi @ sn
            ref = om[i] + c1 * (np.sin(th[i]) * C - np.cos(th[i]) * S)
            scale = c1 * np.abs(Mi).sum() + abs(om[i])
            assert abs(f[i] - ref) <= 1e-12 * scale, i
        # symmetric M: the coupling terms cancel in the sum
        assert abs(f.sum() - om.sum()) <= 1e-13 * (np.abs(f).sum() + 1.0)
        # ... and so does every stage of every step
        it.set_state(th, 0.0, 0.01)
        it.begin(np.arange(1, 11, dtype=np.float64), 10.0, keep_rows=False, ring_rows=64)
        st = it.advance(10)
        y, t, h = it.get_state()
    assert st.error == 0 and st.accepted >= 5
    drift = y.sum() - th.sum() - t * om.sum()
    assert abs(drift) <= 1e-12 * (np.abs(y).sum() + 1.0), drift
