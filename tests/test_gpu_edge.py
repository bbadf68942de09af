"""Edge cases of the single-network path on the GPU, through the C ABI.

* N = 1: the coupling term sin(th) C - cos(th) S vanishes (C = M cos th,
  S = M sin th for one oscillator; up to the rounding of the two products),
  so theta(t) = theta0 + omega t, which DP45 integrates exactly up to
  rounding.
* N = 2: the Adler pair against the oracle (closed-form RHS, kuramoto.cu:45).
* N = 65 and 127 (not multiples of the 64-element padding) against the
  oracle.
* t1 = t0: kmt_begin leaves the run inactive; no step, no output, no error.
* No output grid (n_eval = 0): the run still integrates to t1.
* A single-instance batch equals the single-network engine bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_single_oscillator_is_linear(km, prec):
    p = km.F32 if prec == "f32" else km.F64
    dtype = np.float32 if prec == "f32" else np.float64
    X = np.array([[0.7, -1.1, 0.3]], dtype)
    w = km.factor_weights(p, 1, 3, gaussian=True)
    th0 = np.array([0.4], dtype)
    om = np.array([1.3], dtype)
    print(f"case N=1 {prec}", flush=True)
    with km.Integrator(1, p, atol=1e-6 if prec == "f32" else 1e-10, c1=2.0) as it:
        it.set_omega(om)
        it.set_factor(X, w)
        it.build_adjacency()
        f = it.rhs(th0)
        # sin(th) (M cos th) - cos(th) (M sin th) cancels up to the rounding
        # of the two products
        assert abs(float(f[0]) - float(om[0])) <= 8 * np.finfo(dtype).eps * 2.0
        it.set_state(th0, 0.0, 0.01)
        res = it.run(np.array([0.5, 1.0]), 1.0, keep_rows=True)
    assert res["status"].error == 0 and res["status"].active == 0
    expect = th0[0] + om[0] * np.array([0.5, 1.0])
    tol = 1e-5 if prec == "f32" else 1e-13
    assert np.max(np.abs(res["rows"][:, 0] - expect)) < tol
    assert np.allclose(res["r"], 1.0)


@pytest.mark.parametrize("N", [2, 65, 127])
def test_small_and_ragged_sizes_match_oracle(km, orc, N):
    X = km.factor(N, 5, 31, gaussian=True)
    w = km.factor_weights(km.F64, N, 5, gaussian=True)
    th0, om = km.draws(km.F64, N, 31, normal=True, init=6.283185307)
    te = np.arange(1, 5) * 0.25
    print(f"case N={N}", flush=True)
    with km.Integrator(N, km.F64, atol=1e-10, c1=4.0) as it:
        it.set_omega(om)
        it.set_factor(X, w)
        it.build_adjacency()
        it.set_state(th0, 0.0, 0.01)
        g = it.run(te, 1.0, keep_rows=True)
    ref = orc.integrate(orc.Adj(0, M=orc.adj_rows(X, w, 0, N)), om, 4.0, th0, 0.0, 0.01, 1.0, 1e-10, t_eval=te)
    assert abs(g["status"].accepted - ref["accepted"]) <= 1
    rel = np.abs(g["rows"] - ref["rows"]) / np.maximum(1.0, np.abs(ref["rows"]))
    assert rel.max() < 1e-9


def test_zero_length_run_and_no_output_grid(km):
    N = 300
    X = km.factor(N, 5, 3, gaussian=True)
    w = km.factor_weights(km.F64, N, 5, gaussian=True)
    th0, om = km.draws(km.F64, N, 3, normal=True, init=6.283185307)
    print("case t1 = t0 and n_eval = 0", flush=True)
    with km.Integrator(N, km.F64, atol=1e-9, c1=3.0) as it:
        it.set_omega(om)
        it.set_factor(X, w)
        it.build_adjacency()
        it.set_state(th0, 2.0, 0.01)
        res = it.run(np.zeros(0), 2.0, keep_rows=True)
        assert res["status"].active == 0 and res["status"].accepted == 0 and len(res["r"]) == 0
        y, t, h = it.get_state()
        assert np.array_equal(y, th0) and t == 2.0
        it.set_state(th0, 0.0, 0.01)
        res = it.run(np.zeros(0), 1.0, keep_rows=True)
        assert res["status"].active == 0 and res["status"].accepted > 0 and len(res["r"]) == 0
        assert res["status"].t >= 1.0


def test_single_instance_batch_equals_engine(km):
    N = 777
    th, om = km.draws(km.F32, N, 12, normal=False, init=6.283185307)
    X = km.factor(N, 5, 12, gaussian=True).astype(np.float32)
    w = km.factor_weights(km.F32, N, 5, gaussian=True)
    te = (np.float32(0.25) * np.arange(1, 5)).astype(np.float64)
    print("case batch B=1", flush=True)
    with km.Batch([N], atol=1e-6, shared=False, storage="native") as bt:
        bt.set_instance(0, om, th, 2.5, 0.0, 0.01)
        bt.set_factor(0, X, w)
        res = bt.run(te, 1.0, chunk=4)
        y_b, st_b = bt.get_instance(0)
    with km.Integrator(N, km.F32, atol=1e-6, c1=2.5) as it:
        it.set_omega(om)
        it.set_factor(X, w)
        it.build_adjacency()
        it.set_state(th, 0.0, 0.01)
        single = it.run(te, 1.0, keep_rows=False)
        y, t, h = it.get_state()
    assert np.array_equal(res["r"][0], single["r"])
    assert st_b["accepted"] == single["status"].accepted
    assert np.array_equal(y_b, y)
