"""Whole-run parity of three workloads against the CPU oracle, at their full
sizes, through the C ABI.

* The reference's default run (kuramoto.cu:277-294 defaults with -K 5;
  loop dp45.cu:148-164): N=5000 FP32, rank-5 Gaussian factor from seed 1,
  dense A = X X^T / K with the attractive sign, Cauchy frequencies (f=1),
  theta0 = 0, c = 3, t = 100, output every dt = 1, rtol = 0, atol = 1e-6,
  first step dt/100.  Against the FP32 oracle on the same stored matrix over
  the whole run: phases at t=1 within 1e-5 relative to max(1, |theta|),
  accepted steps within 2%, r(t) within 1e-3.  Against the FP64 oracle of
  the same method (kuramoto_64 on the same inputs, coupled through the
  factor): r(t) within 1e-3 at all 101 outputs (r(0) and t = 1..100),
  accepted steps within 2%.
* The FP64 memory-bound workload (kuramoto_64.cu:118-126, dp45_64.cu:87-164):
  N=50000, dense FP64 A (20 GB on the device, and the same 20 GB matrix in
  host memory for the oracle), normal frequencies, theta0 = pi(2u-1), c = 3,
  seed 2, atol = 1e-9, to t = 1: phases at t=1 within 1e-10, accepted steps
  within 2% (or 1), r(t) within 1e-3.
* -A with adjacency tiles rebuilt in registers (kuramoto.cu:104-108):
  N=200000, K=5, seed 3, Cauchy, theta0 = 0, c = 3, t = 2, dt = 0.5,
  atol = 1e-6.  The rebuilt tiles carry the dense matrix's FP32 rounding of
  every M_ij; the oracle's -A run never forms M_ij (it contracts through
  the factor), so the two FP32 runs see coupling sums that differ by the
  entries' rounding (~1e-7 relative).  With Cauchy frequencies that can
  move phases by ~1e-4 at t = 1 (two FP32 oracle runs of another network,
  dense and factored, N=1500, do); on this network it stays below 1e-5.
  Bounds: accepted steps within 2% (or 1), r(t) within 1e-3 and phases at
  t = 1 and 2 within 1e-5 of the FP32 oracle's -A run; phases no further
  from the FP64 oracle's run of the same method than twice the FP32
  oracle's distance + 1e-5; and sampled RHS rows at full size against the
  formula on the oracle generator's FP32 rows (one float ulp of the row
  scale: both sides sum in double and round once).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _grid(t1, dt, dtype):
    dt = dtype(dt)
    return (dt * np.arange(1, int(dtype(t1) / dt) + 1)).astype(dtype)


def _phase_err(a, b):
    b64 = np.asarray(b, np.float64)
    return float(np.max(np.abs(np.asarray(a, np.float64) - b64) / np.maximum(1.0, np.abs(b64))))


def test_default_run_n5000_fp32_t100(km, orc):
    N, K, seed, c1, t1, dt, atol = 5000, 5, 1, 3.0, 100.0, 1.0, 1e-6
    theta0, omega = km.draws(km.F32, N, seed, normal=False, freq=1.0, init=0.0)
    X64 = km.factor(N, K, seed, gaussian=True)
    X = X64.astype(np.float32)
    w = km.factor_weights(km.F32, N, K, gaussian=True)
    te = _grid(t1, dt, np.float32)
    h0 = float(np.float32(dt) / np.float32(100))
    print("case default run N=5000 FP32 t=100: GPU", flush=True)
    with km.Integrator(N, km.F32, atol=atol, c1=c1) as it:
        it.set_omega(omega)
        it.set_factor(X, w)
        it.build_adjacency()
        r0 = it.order(theta0)
        it.set_state(theta0, 0.0, h0)
        g = it.run(te.astype(np.float64), t1, keep_rows=True, ring_rows=128, chunk=256)
    st = g["status"]
    assert st.error == 0 and st.active == 0 and len(g["r"]) == len(te)
    print("  FP32 oracle (same stored matrix)", flush=True)
    M = orc.adj_rows(X, w, 0, N)
    o32 = orc.integrate(orc.Adj(0, M=M), omega, c1, theta0, 0.0, h0, t1, atol, t_eval=te, dtype=np.float32)
    del M
    print("  FP64 oracle (same method, through the factor)", flush=True)
    w64 = km.factor_weights(km.F64, N, K, gaussian=True)
    o64 = orc.integrate(orc.Adj(2, X=X64, w=w64), omega.astype(np.float64), c1, theta0.astype(np.float64), 0.0,
                        h0, t1, atol, t_eval=te.astype(np.float64), keep_rows=False)
    e1 = _phase_err(g["rows"][0], o32["rows"][0])  # t = 1
    r_gpu = np.concatenate([[r0], g["r"].astype(np.float64)])
    r_64 = np.concatenate([[1.0], o64["r"]])  # theta0 = 0: r(0) = 1 exactly
    dr32 = float(np.max(np.abs(g["r"].astype(np.float64) - o32["r"])))
    dr64 = float(np.max(np.abs(r_gpu - r_64)))
    print(f"  accepted gpu={st.accepted} fp32 oracle={o32['accepted']} fp64 oracle={o64['accepted']}; "
          f"rejected {st.rejected}/{o32['rejected']}/{o64['rejected']}; phase err t=1 {e1:.3e}; "
          f"max|dr| fp32 {dr32:.3e} fp64 {dr64:.3e}; phase err t=100 {_phase_err(g['rows'][-1], o32['rows'][-1]):.3e}",
          flush=True)
    assert e1 < 1e-5
    assert abs(st.accepted - o32["accepted"]) <= 0.02 * o32["accepted"]
    assert abs(st.accepted - o64["accepted"]) <= 0.02 * o64["accepted"]
    assert len(r_gpu) == 101 and dr64 < 1e-3 and dr32 < 1e-3


def test_fp64_n50000_dense_t1_against_cpu_oracle(km, orc):
    N, K, seed, c1, t1, atol = 50000, 5, 2, 3.0, 1.0, 1e-9
    theta0, omega = km.draws(km.F64, N, seed, normal=True, freq=1.0, init=6.283185307)
    X = km.factor(N, K, seed, gaussian=True)
    w = km.factor_weights(km.F64, N, K, gaussian=True)
    te = _grid(t1, 0.1, np.float64)
    print("case FP64 N=50000 dense (20 GB) to t=1: GPU", flush=True)
    with km.Integrator(N, km.F64, atol=atol, c1=c1) as it:
        it.set_omega(omega)
        it.set_factor(X, w)
        it.build_adjacency()
        it.set_state(theta0, 0.0, 0.01)
        g = it.run(te, t1, keep_rows=True, ring_rows=16)
    st = g["status"]
    assert st.error == 0 and st.active == 0
    print("  oracle: building the 20 GB matrix in host memory", flush=True)
    M = orc.adj_rows(X, w, 0, N)
    print("  oracle: integrating", flush=True)
    ref = orc.integrate(orc.Adj(0, M=M), omega, c1, theta0, 0.0, 0.01, t1, atol, t_eval=te)
    del M
    e1 = _phase_err(g["rows"][-1], ref["rows"][-1])
    dr = float(np.max(np.abs(g["r"] - ref["r"])))
    print(f"  accepted gpu={st.accepted} oracle={ref['accepted']}; phase err t=1 {e1:.3e}; max|dr| {dr:.3e}",
          flush=True)
    assert abs(st.accepted - ref["accepted"]) <= max(1, int(0.02 * ref["accepted"]))
    assert e1 < 1e-10 and dr < 1e-3


def test_rebuild_n200000_k5_t2(km, orc):
    N, K, seed, c1, t1, atol = 200000, 5, 3, 3.0, 2.0, 1e-6
    theta0, omega = km.draws(km.F32, N, seed, normal=False, freq=1.0, init=0.0)
    X = km.factor(N, K, seed, gaussian=True).astype(np.float32)
    w = km.factor_weights(km.F32, N, K, gaussian=True)
    te = _grid(t1, 0.5, np.float32)
    h0 = float(np.float32(0.5) / np.float32(100))
    rng = np.random.default_rng(9)
    th = rng.uniform(-40, 40, N).astype(np.float32)
    print("case -A rebuilt tiles N=200000 K=5 to t=2: GPU", flush=True)
    with km.Integrator(N, km.F32, atol=atol, c1=c1) as it:
        it.set_omega(omega)
        it.set_factor(X, w)
        it.use_factored("rebuild")
        f = it.rhs(th)
        it.set_state(theta0, 0.0, h0)
        g = it.run(te.astype(np.float64), t1, keep_rows=True, ring_rows=8)
    st = g["status"]
    assert st.error == 0 and st.active == 0 and len(g["r"]) == len(te)
    # sampled rows of the RHS at full size (64-bit indexing of N^2 = 4e10 pairs)
    c, s = orc.sincos(th, np.float32)
    c64, s64 = c.astype(np.float64), s.astype(np.float64)
    for i in [0, 1, 4095, 65536, 99999, 131071, 150001, N - 2, N - 1]:
        Mi = orc.adj_rows(X, w, i, i + 1)[0].astype(np.float64)
        C, S = np.float32(Mi @ c64), np.float32(Mi @ s64)
        ref = np.float64(omega[i]) + c1 * (np.float64(s[i]) * np.float64(C) - np.float64(c[i]) * np.float64(S))
        scale = abs(float(omega[i])) + c1 * np.abs(Mi).sum()
        assert abs(float(f[i]) - ref) <= 2.5e-7 * scale, i
    print("  oracle: FP32 and FP64 -A runs", flush=True)
    o32 = orc.integrate(orc.Adj(2, X=X, w=w), omega, c1, theta0, 0.0, h0, t1, atol, t_eval=te, dtype=np.float32)
    o64 = orc.integrate(orc.Adj(2, X=X.astype(np.float64), w=w.astype(np.float64)), omega.astype(np.float64), c1,
                        theta0.astype(np.float64), 0.0, h0, t1, atol, t_eval=te.astype(np.float64))
    dr = float(np.max(np.abs(g["r"] - o32["r"])))
    print(f"  accepted gpu={st.accepted} oracle={o32['accepted']}; max|dr| {dr:.3e}", flush=True)
    assert abs(st.accepted - o32["accepted"]) <= max(1, int(0.02 * o32["accepted"]))
    assert dr < 1e-3
    for q in (1, 3):  # t = 1, t = 2
        e_gpu, e_orc = _phase_err(g["rows"][q], o64["rows"][q]), _phase_err(o32["rows"][q], o64["rows"][q])
        e32 = _phase_err(g["rows"][q], o32["rows"][q])
        print(f"  t={te[q]}: phase err vs FP64 gpu {e_gpu:.3e} FP32 oracle {e_orc:.3e}; "
              f"gpu vs FP32 oracle {e32:.3e}", flush=True)
        assert e32 < 1e-5, (float(te[q]), e32)
        assert e_gpu <= 2 * e_orc + 1e-5, (float(te[q]), e_gpu, e_orc)
