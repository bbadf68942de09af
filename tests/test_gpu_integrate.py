"""GPU parity of the fused DP45 integrator against the oracle and the
reference's own K=0 fixtures, through the C ABI.

Reference: dp45_step / dp45_run / dp45_eval (dp45.cu:87-164), step_eval's
output grid and order parameter (kuramoto.cu:164-249), main's final-state
interpolation (kuramoto.cu:630-647).

Tolerances: phases at t=1 within 1e-10 (FP64) / 1e-5 (FP32) of the oracle
in the same precision on the same inputs, relative to max(1, |theta|);
r(t) within 1e-3; accepted-step counts within 2% (or 1).  The FP32 bound
holds because both sides accumulate the coupling sums in double and round
cos / sin once from double, so they see the same FP32 values; with Cauchy
frequencies a one-ulp difference in any of them moves the population by
~1e-4 within t = 1, so the bound is also the check of that rounding.  The
K=0 controller KAT is bit-exact in t and step count.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kat_k0.json")


def _grid(t0, t1, dt, dtype):
    dt = dtype(dt)
    n0 = int(dtype(t0) / dt) + 1
    n1 = int(dtype(t1) / dt) + 1
    return (dt * np.arange(n0, n1)).astype(dtype)


@pytest.mark.parametrize("name", ["5000", "5000_lorentz"])
def test_k0_controller_kat(km, orc, name):
    """data/5000{,_lorentz}/0/1fs.dat: N=5000, K=0, c=1.75, t=100, dt=0.01,
    seed 1.  The GPU run must end at the fixture's t (bit-exact) with the
    same number of accepted steps as the oracle, h within 4 ulp, phases at
    the last accepted step within 8 ulp; and equal the oracle bit-for-bit."""
    fx = json.load(open(GOLD))[name]
    N = fx["N"]
    print(f"case kat {name}", flush=True)
    theta0, omega = km.draws(km.F32, N, fx["seed"], normal=fx["normal"], freq=1.0, init=0.0)
    te = _grid(0.0, fx["t1"], fx["dt"], np.float32)
    h0 = float(np.float32(fx["dt"]) / np.float32(100))
    with km.Integrator(N, km.F32, atol=fx["atol"], c1=fx["c1"]) as it:
        it.set_omega(omega)
        it.set_factor(np.zeros((N, 0), np.float32), np.zeros(0, np.float32))
        it.use_factored(True)
        it.set_state(theta0, 0.0, h0)
        res = it.run(te, fx["t1"], keep_rows=False, ring_rows=4096, chunk=64)
        y, t, h = it.get_state()
    st = res["status"]
    assert np.float32(t).view(np.uint32) == fx["t_bits"]
    assert abs(int(np.float32(h).view(np.int32)) - int(np.uint32(fx["h_bits"]).view(np.int32))) <= 4
    for i, bits in fx["theta_samples"]:
        assert abs(int(y[i:i + 1].view(np.int32)[0]) - int(np.uint32(bits).view(np.int32))) <= 8
    ref = orc.integrate(orc.Adj(2, X=np.zeros((N, 0)), w=np.zeros(0)), omega, fx["c1"], theta0, 0.0, h0, fx["t1"],
                        fx["atol"], t_eval=te, keep_rows=False, dtype=np.float32)
    assert st.accepted == ref["accepted"] and st.rejected == ref["rejected"]
    assert np.float32(h) == np.float32(ref["h"])
    assert np.array_equal(y.view(np.uint32), ref["y"].view(np.uint32))
    assert len(res["r"]) == len(te)
    assert np.max(np.abs(res["r"] - ref["r"])) < 1e-6


def _inputs(km, prec, N, K, seed, normal, init, gauss=True):
    dtype = np.float32 if prec == km.F32 else np.float64
    theta0, omega = km.draws(prec, N, seed, normal=normal, freq=1.0, init=init)
    X = km.factor(N, K, seed, gaussian=gauss).astype(dtype)
    w = km.factor_weights(prec, N, K, gaussian=gauss)
    return dtype, theta0, omega, X, w


def _phase_err(a, b):
    """max |a - b| / max(1, |b|)."""
    b64 = np.asarray(b, np.float64)
    return float(np.max(np.abs(np.asarray(a, np.float64) - b64) / np.maximum(1.0, np.abs(b64))))


PHASE_TOL = {"f32": 1e-5, "f64": 1e-10}


@pytest.mark.parametrize("prec,atol", [("f64", 1e-9), ("f32", 1e-6)])
@pytest.mark.parametrize("N", [333, 2000])
def test_dense_integration_matches_oracle(km, orc, prec, atol, N):
    """Phases at t=1 (and the final kmt_eval interpolation) within 1e-10
    (FP64) / 1e-5 (FP32) of the oracle in the same precision on the same
    inputs; r(t) within 1e-3, accepted steps within 2% (or 1)."""
    p = km.F32 if prec == "f32" else km.F64
    K, seed, c1, t1 = 5, 2, 3.0, 1.0
    dtype, theta0, omega, X, w = _inputs(km, p, N, K, seed, normal=True, init=6.283185307)
    te = _grid(0.0, t1, 0.1, dtype)
    print(f"case dense_integration N={N} prec={prec}", flush=True)
    with km.Integrator(N, p, atol=atol, c1=c1) as it:
        it.set_omega(omega)
        it.set_factor(X, w)
        it.build_adjacency("native")
        it.set_state(theta0, 0.0, 0.01)
        g = it.run(te, t1, keep_rows=True, ring_rows=8)
        fin = it.eval(float(te[-1]))
    M = orc.adj_rows(X, w, 0, N)
    ref = orc.integrate(orc.Adj(0, M=M), omega, c1, theta0, 0.0, 0.01, t1, atol, t_eval=te, dtype=dtype)
    st = g["status"]
    assert abs(st.accepted - ref["accepted"]) <= max(1, int(0.02 * ref["accepted"]))
    assert g["rows"].shape == ref["rows"].shape
    assert np.max(np.abs(g["r"] - ref["r"])) < 1e-3
    assert len(g["times"]) == st.accepted
    e1 = _phase_err(g["rows"][-1], ref["rows"][-1])  # phases at t = 1
    ef = _phase_err(fin, ref["final"])
    print(f"  accepted gpu={st.accepted} oracle={ref['accepted']} phase err t=1 {e1:.3e} final {ef:.3e}", flush=True)
    assert e1 < PHASE_TOL[prec] and ef < PHASE_TOL[prec], (e1, ef)


def test_ring_stall_matches_large_ring(km):
    """A one-row output ring pauses the device at every output; the results
    must equal a run with a large ring."""
    N, K, seed, c1, t1 = 500, 5, 4, 3.0, 2.0
    dtype, theta0, omega, X, w = _inputs(km, km.F64, N, K, seed, normal=False, init=6.283185307)
    te = _grid(0.0, t1, 0.01, dtype)
    outs = []
    for ring in (1, 256):
        with km.Integrator(N, km.F64, atol=1e-8, c1=c1) as it:
            it.set_omega(omega)
            it.set_factor(X, w)
            it.build_adjacency()
            it.set_state(theta0, 0.0, 0.0001)
            outs.append(it.run(te, t1, keep_rows=True, ring_rows=ring, chunk=4))
    a, b = outs
    assert np.array_equal(a["rows"], b["rows"]) and np.array_equal(a["r"], b["r"])
    assert np.array_equal(a["times"], b["times"])
    assert len(a["r"]) == len(te)


def test_fixed_step_matches_oracle(km, orc):
    N, K, seed, c1 = 400, 5, 9, 3.0
    dtype, theta0, omega, X, w = _inputs(km, km.F64, N, K, seed, normal=True, init=1.0)
    te = _grid(0.0, 1.0, 0.25, dtype)
    with km.Integrator(N, km.F64, atol=1e-6, c1=c1, fixed=True) as it:
        it.set_omega(omega)
        it.set_factor(X, w)
        it.build_adjacency()
        it.set_state(theta0, 0.0, 0.0625)
        g = it.run(te, 1.0, keep_rows=True)
    M = orc.adj_rows(X, w, 0, N)
    ref = orc.integrate(orc.Adj(0, M=M), omega, c1, theta0, 0.0, 0.0625, 1.0, 1e-6, fixed=True, t_eval=te)
    assert g["status"].accepted == ref["accepted"] == 16 and g["status"].rejected == 0
    assert np.max(np.abs(g["rows"] - ref["rows"])) < 1e-11


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("K", [5, 12, 20])
def test_factored_integration_matches_dense(km, orc, prec, K):
    """-A (factored) and the dense matrix built from the same factor
    (reference: kuramoto vs kuramoto -A).  Each GPU run against the oracle
    run of the same form: phases at t=1 within 1e-10 (FP64) / 1e-5 (FP32),
    r(t) within 1e-3, accepted steps within 2% (or 1); and the two forms
    against each other: r(t) within 1e-3, steps within 2% (or 1).  (In
    FP32 the two forms differ by the rounding of the stored M_ij, which
    with Cauchy frequencies moves the phases by ~2e-4 at t=1 in the oracle
    too.)  K = 5 and 12 run the one-launch-per-stage low-rank kernels (rank
    tiles of 8 and 16), K = 20 the three-launch path for larger ranks."""
    p = km.F32 if prec == "f32" else km.F64
    N, seed, c1, t1 = 1500, 3, 3.0, 1.0
    atol = 1e-6 if prec == "f32" else 1e-9
    dtype, theta0, omega, X, w = _inputs(km, p, N, K, seed, normal=False, init=0.0)
    te = _grid(0.0, t1, 0.1, dtype)
    print(f"case factored K={K} prec={prec}", flush=True)
    res = {}
    for mode in ("dense", "factored"):
        with km.Integrator(N, p, atol=atol, c1=c1) as it:
            it.set_omega(omega)
            it.set_factor(X, w)
            if mode == "dense":
                it.build_adjacency()
            else:
                it.use_factored(True)
            it.set_state(theta0, 0.0, 0.01)
            res[mode] = it.run(te, t1, keep_rows=True)
    refs = {"dense": orc.integrate(orc.Adj(0, M=orc.adj_rows(X, w, 0, N)), omega, c1, theta0, 0.0, 0.01, t1, atol,
                                   t_eval=te, dtype=dtype),
            "factored": orc.integrate(orc.Adj(2, X=X, w=w), omega, c1, theta0, 0.0, 0.01, t1, atol, t_eval=te,
                                      dtype=dtype)}
    for mode in res:
        g, ref = res[mode], refs[mode]
        e1 = _phase_err(g["rows"][-1], ref["rows"][-1])
        print(f"  {mode}: accepted gpu={g['status'].accepted} oracle={ref['accepted']} phase err t=1 {e1:.3e}",
              flush=True)
        assert abs(g["status"].accepted - ref["accepted"]) <= max(1, int(0.02 * ref["accepted"]))
        assert np.max(np.abs(g["r"] - ref["r"])) < 1e-3
        assert e1 < PHASE_TOL[prec], (mode, e1)
    d, f = res["dense"], res["factored"]
    assert abs(d["status"].accepted - f["status"].accepted) <= max(1, int(0.02 * d["status"].accepted))
    assert np.max(np.abs(d["r"] - f["r"])) < 1e-3


def test_resume_is_bit_exact(km):
    """Stopping after some steps and restarting from kmt_get_state (the
    FILEBASEfs.dat restart of kuramoto.cu:471-494) continues bit-exactly:
    the restart recomputes k1 = f(y), equal to the FSAL k7."""
    N, K, seed, c1, t1 = 700, 5, 6, 3.0, 3.0
    dtype, theta0, omega, X, w = _inputs(km, km.F64, N, K, seed, normal=True, init=6.283185307)
    te = _grid(0.0, t1, 0.5, dtype)

    def make():
        it = km.Integrator(N, km.F64, atol=1e-9, c1=c1)
        it.set_omega(omega)
        it.set_factor(X, w)
        it.build_adjacency()
        return it

    with make() as it:
        it.set_state(theta0, 0.0, 0.01)
        full = it.run(te, t1, keep_rows=False)
        yf, tf, hf = it.get_state()
    with make() as it:
        it.set_state(theta0, 0.0, 0.01)
        part = it.run(te, t1, keep_rows=False, chunk=5, max_attempts=15)
        y1, t1_, h1 = it.get_state()
    te2 = _grid(t1_, t1, 0.5, dtype)
    with make() as it:
        it.set_state(y1, t1_, h1)
        rest = it.run(te2, t1, keep_rows=False)
        y2, t2, h2 = it.get_state()
    assert t2 == tf and h2 == hf
    assert np.array_equal(y2, yf)
    assert np.array_equal(np.concatenate([part["times"], rest["times"]]), full["times"])


def test_step_and_eval_api(km, orc):
    """kmt_step (dp45_step) and kmt_eval (dp45_eval) one attempt at a time."""
    N, K, seed, c1 = 300, 5, 8, 3.0
    dtype, theta0, omega, X, w = _inputs(km, km.F64, N, K, seed, normal=True, init=6.283185307)
    with km.Integrator(N, km.F64, atol=1e-9, c1=c1) as it:
        it.set_omega(omega)
        it.set_factor(X, w)
        it.build_adjacency()
        it.set_state(theta0, 0.0, 0.01)
        it.begin([], 10.0, keep_rows=False, ring_rows=1)
        acc_n = 0
        for _ in range(20):
            acc, t, h = it.step()
            acc_n += acc
        y, t, h = it.get_state()
        mid = it.eval((t + it.advance(0).t_last) / 2)
    M = orc.adj_rows(X, w, 0, N)
    ref = orc.integrate(orc.Adj(0, M=M), omega, c1, theta0, 0.0, 0.01, 10.0, 1e-9, max_steps=acc_n)
    # the controller amplifies summation-order differences (1e-12 in y at
    # atol 1e-9) into 1e-11-level differences of h; compare at that level
    assert abs(ref["t"] - t) < 1e-9 * abs(t)
    assert np.max(np.abs(ref["y"] - y)) < 1e-8
    assert np.all(np.isfinite(mid))


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("K", [5, 12])
def test_rebuilt_tiles_equal_dense(km, prec, K):
    """-A with adjacency tiles rebuilt in registers (kmt_use_factored mode
    KMT_LOWRANK_REBUILD) forms every M_ij with the dense generator's formula
    and sums columns in the dense stage kernel's order: the RHS, the
    couplings and a whole integration equal the dense path on the generated
    matrix bit for bit."""
    p = km.F32 if prec == "f32" else km.F64
    N, seed, c1, t1 = 1337, 4, 3.0, 1.0
    atol = 1e-6 if prec == "f32" else 1e-9
    dtype, theta0, omega, X, w = _inputs(km, p, N, K, seed, normal=True, init=6.283185307)
    te = _grid(0.0, t1, 0.25, dtype)
    print(f"case rebuild K={K} prec={prec}", flush=True)
    out = {}
    for mode in ("dense", "rebuild"):
        with km.Integrator(N, p, atol=atol, c1=c1) as it:
            it.set_omega(omega)
            it.set_factor(X, w)
            if mode == "dense":
                it.build_adjacency()
            else:
                it.use_factored("rebuild")
            f, cpl = it.rhs(theta0, couplings=True)
            it.set_state(theta0, 0.0, 0.01)
            res = it.run(te, t1, keep_rows=True)
            out[mode] = (f, cpl, res, it.get_state())
    (fd, cd, rd, sd), (fr, cr, rr, sr) = out["dense"], out["rebuild"]
    assert np.array_equal(fd, fr) and np.array_equal(cd, cr)
    assert rd["status"].accepted == rr["status"].accepted and rd["status"].rejected == rr["status"].rejected
    assert np.array_equal(rd["rows"], rr["rows"]) and np.array_equal(rd["r"], rr["r"])
    assert np.array_equal(sd[0], sr[0]) and sd[1] == sr[1]


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_fp64_state_with_16bit_adjacency(km, orc, storage):
    """The 64-bit program with a BF16/FP16 adjacency: products accumulated in
    FP32, state update, stage combinations and error norm in FP64.  Against
    the FP64 oracle on the same (dequantized) matrix: accepted steps within
    2% (or 1), r(t) within 1e-6 and phases at t=1 within 1e-6 relative (the
    FP32 accumulation of the coupling is the only reduced-precision step)."""
    N, K, seed, c1, t1, atol = 1500, 5, 7, 3.0, 1.0, 1e-9
    dtype, theta0, omega, X, w = _inputs(km, km.F64, N, K, seed, normal=True, init=6.283185307)
    te = _grid(0.0, t1, 0.1, dtype)
    print(f"case fp64 + {storage}", flush=True)
    with km.Integrator(N, km.F64, atol=atol, c1=c1) as it:
        it.set_omega(omega)
        it.set_factor(X, w)
        it.build_adjacency(storage)
        Mq = it.adjacency_rows(0, N)
        it.set_state(theta0, 0.0, 0.01)
        g = it.run(te, t1, keep_rows=True)
    ref = orc.integrate(orc.Adj(0, M=Mq), omega, c1, theta0, 0.0, 0.01, t1, atol, t_eval=te)
    assert abs(g["status"].accepted - ref["accepted"]) <= max(1, int(0.02 * ref["accepted"]))
    assert np.max(np.abs(g["r"] - ref["r"])) < 1e-6
    rel = np.abs(g["rows"][-1] - ref["rows"][-1]) / np.maximum(1.0, np.abs(ref["rows"][-1]))
    assert rel.max() < 1e-6, rel.max()


@pytest.mark.parametrize("N,prec", [(5000, "f32"), (6000, "f64")])
def test_restart_in_new_context_is_bit_exact(km, N, prec):
    """12 attempts in one context equal 5 attempts, a new context started
    from kmt_get_state (the FILEBASEfs.dat restart, kuramoto.cu:471-494) and
    7 more, bit for bit.  Each context picks its stage-kernel variant in
    kmt_begin by the same fixed rule (N=5000 FP32: 100 MB, 4 rows per
    group; N=6000 FP64: 288 MB, 8 rows), so the error norm's partial sums
    are grouped alike in every context."""
    p = km.F32 if prec == "f32" else km.F64
    dtype, theta0, omega, X, w = _inputs(km, p, N, 5, 4, normal=False, init=6.283185307)
    te = np.arange(1, 101, dtype=np.float64)
    atol = 1e-6 if prec == "f32" else 1e-9
    print(f"case restart N={N} {prec}", flush=True)

    def run(y0, t0, h0, attempts):
        with km.Integrator(N, p, atol=atol, c1=3.0) as it:
            it.set_omega(omega)
            it.set_factor(X, w)
            it.build_adjacency()
            it.set_state(y0, t0, h0)
            it.begin(te[te > t0], 100.0, keep_rows=False, ring_rows=64)
            st = it.advance(attempts)
            assert st.error == 0 and st.active == 1
            return it.get_state() + (it.layout()[2],)

    y_full, t_full, h_full, rows = run(theta0, 0.0, 0.01, 12)
    y1, t1, h1, _ = run(theta0, 0.0, 0.01, 5)
    y2, t2, h2, _ = run(y1, t1, h1, 7)
    assert rows == (4 if prec == "f32" else 8)
    assert t2 == t_full and h2 == h_full and np.array_equal(y2, y_full)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("mode", ["dense", "factored", "signs"])
def test_drained_couplings_equal_rhs_on_rows(km, orc, prec, mode):
    """kmt_read_outputs_couplings (couplings.dat's rows, kuramoto.cu:226-231)
    computes each drained row's couplings on the device from the row in the
    ring: equal bit for bit to kmt_rhs on the rows it returns, for the dense
    matrix, the factored -A contraction and the packed-sign factor; the rows
    and r equal kmt_read_outputs's; the last row's couplings match the
    oracle's within the tolerance of the RHS parity tests (dense only)."""
    p = km.F32 if prec == "f32" else km.F64
    N, K, seed, c1, t1 = 700, 5, 6, 3.0, 0.5
    atol = 1e-6 if prec == "f32" else 1e-9
    dtype, theta0, omega, X, w = _inputs(km, p, N, K, seed, normal=True, init=6.283185307)
    te = _grid(0.0, t1, 0.05, dtype)
    print(f"case drained couplings mode={mode} prec={prec}", flush=True)
    outs = {}
    for drain in ("plain", "couplings"):
        with km.Integrator(N, p, atol=atol, c1=c1) as it:
            it.set_omega(omega)
            if mode == "signs":
                it.set_factor_sign(N, seed)
                it.use_factored(True)
            else:
                it.set_factor(X, w)
                if mode == "dense":
                    it.build_adjacency()
                else:
                    it.use_factored(True)
            it.set_state(theta0, 0.0, 0.01)
            it.begin(te, t1, keep_rows=True, ring_rows=4)
            rows, rs, cps = [], [], []
            while True:
                st = it.advance(3)
                it.read_times()
                if drain == "plain":
                    r_rows, r = it.read_outputs()
                else:
                    r_rows, r, cp = it.read_outputs_couplings()
                    cps.append(cp)
                rows.append(r_rows)
                rs.append(r)
                if not st.active:
                    break
            rows, rs = np.concatenate(rows), np.concatenate(rs)
            assert len(rs) == len(te)
            if drain == "couplings":
                cps = np.concatenate(cps)
                assert cps.shape == (len(te), 2 * N)
                for q in range(len(te)):
                    assert np.array_equal(cps[q], it.couplings(rows[q])), q
            outs[drain] = (rows, rs)
    assert np.array_equal(outs["plain"][0], outs["couplings"][0])
    assert np.array_equal(outs["plain"][1], outs["couplings"][1])
    if mode == "dense":
        M = orc.adj_rows(X, w, 0, N).astype(dtype)
        _, cso = orc.rhs(orc.Adj(0, M=M), omega, c1, rows[-1], dtype=dtype)
        rsum = np.abs(M.astype(np.float64)).sum(axis=1)
        tolc = (1.2e-7 if prec == "f32" else 1e-12) * np.repeat(rsum, 2)
        assert np.all(np.abs(cps[-1].astype(np.float64) - cso) <= tolc + 1e-30)


def test_drained_couplings_need_kept_rows(km):
    N, K, seed = 300, 5, 2
    dtype, theta0, omega, X, w = _inputs(km, km.F64, N, K, seed, normal=True, init=1.0)
    te = _grid(0.0, 0.2, 0.05, dtype)
    with km.Integrator(N, km.F64, atol=1e-8) as it:
        it.set_omega(omega)
        it.set_factor(X, w)
        it.build_adjacency()
        it.set_state(theta0, 0.0, 0.01)
        it.begin(te, 0.2, keep_rows=False, ring_rows=8)
        st = it.advance(64)
        assert st.rows_ready > 0
        with pytest.raises(km.KmtError, match="KMT_ERR_STATE"):
            it.read_outputs_couplings()
