"""GPU parity of the batched integrator (B instances in lockstep stages).

Grouped mode (own adjacency per instance, uneven N) runs the same gemv-pair
loop, epilogue and controller as a single run, so every instance must equal
the single-run engine bit for bit.  Shared mode (one BF16/FP16 adjacency,
MFMA product with hi + lo 16-bit cos/sin) is checked against the single-run
FP32 engine on the dequantized matrix (r within 1e-3, accepted steps within
2% or 1) and against the FP64 oracle on the same matrix.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _grid(t1, dt):
    dt = np.float32(dt)
    return (dt * np.arange(1, int(np.float32(t1) / dt) + 1)).astype(np.float64)


def _instance(km, N, seed, normal, K=5):
    th, om = km.draws(km.F32, N, seed, normal=normal, init=6.283185307)
    X = km.factor(N, K, seed, gaussian=True).astype(np.float32)
    w = km.factor_weights(km.F32, N, K, gaussian=True)
    return th, om, X, w


@pytest.mark.parametrize("storage", ["native", "bf16"])
def test_grouped_equals_single_runs(km, storage):
    Ns = [300, 700, 1000, 64]
    cs = [0.5, 1.5, 3.0, 4.0]
    t1, te = 2.0, _grid(2.0, 0.25)
    insts = [_instance(km, N, 10 + b, normal=b % 2 == 0) for b, N in enumerate(Ns)]
    print(f"case grouped storage={storage}", flush=True)
    with km.Batch(Ns, atol=1e-6, shared=False, storage=storage) as bt:
        for b, (th, om, X, w) in enumerate(insts):
            bt.set_instance(b, om, th, cs[b], 0.0, 0.01)
            bt.set_factor(b, X, w)
        res = bt.run(te, t1, ring_rows=64, chunk=8)
        finals = [bt.get_instance(b) for b in range(len(Ns))]
    assert res["status"].active == 0 and res["status"].errors == 0
    for b, (th, om, X, w) in enumerate(insts):
        with km.Integrator(Ns[b], km.F32, atol=1e-6, c1=cs[b]) as it:
            it.set_omega(om)
            it.set_factor(X, w)
            it.build_adjacency(storage)
            it.set_state(th, 0.0, 0.01)
            single = it.run(te, t1, keep_rows=False, ring_rows=64)
            y, t, h = it.get_state()
        assert np.array_equal(res["r"][b], single["r"]), b
        assert finals[b][1]["accepted"] == single["status"].accepted
        assert finals[b][1]["rejected"] == single["status"].rejected
        assert np.array_equal(finals[b][0], y)
        assert np.float32(finals[b][1]["t"]) == np.float32(t)


@pytest.mark.parametrize("B,N", [(3, 1000), (40, 1000), (64, 2048)])
@pytest.mark.parametrize("storage", ["bf16", "f16"])
@pytest.mark.parametrize("normal", [True, False])
def test_shared_mfma_matches_single_runs(km, orc, B, N, storage, normal):
    """normal=True: moderate phases; accepted-step counts within 2% (or 1) of
    the single run.  normal=False (Cauchy, phases up to ~1e3): FP32 step
    counts sit at the rounding floor of the error estimate and are not
    comparable run to run; r(t) is checked against the FP64 oracle."""
    th_all, om_all = [], []
    for b in range(B):
        th, om = km.draws(km.F32, N, 1 + b, normal=normal, init=6.283185307)
        th_all.append(th)
        om_all.append(om)
    X = km.factor(N, 5, 77, gaussian=True).astype(np.float32)
    w = km.factor_weights(km.F32, N, 5, gaussian=True)
    cs = np.linspace(0.5, 4.0, B)
    t1, te = 1.0, _grid(1.0, 0.25)
    print(f"case shared B={B} N={N} {storage}", flush=True)
    with km.Batch([N] * B, atol=1e-6, shared=True, storage=storage) as bt:
        for b in range(B):
            bt.set_instance(b, om_all[b], th_all[b], cs[b], 0.0, 0.01)
        bt.set_factor(-1, X, w)
        Mq = bt.adjacency_rows(-1, 0, N)
        # the stored (tiled, 16-bit) matrix read back row-major: the FP32
        # factor product rounded once to 8 (bf16) / 11 (f16) bits
        M32 = orc.adj_rows(X, w, 0, N)
        assert np.all(np.abs(Mq - M32) <= np.abs(M32) * 2.0 ** -8 + np.abs(M32).max() * 2.0 ** -20)
        res = bt.run(te, t1, ring_rows=64, chunk=8)
        finals = [bt.get_instance(b) for b in range(B)]
    assert res["status"].active == 0 and res["status"].errors == 0
    check = range(B) if B <= 4 else [0, B // 4 + 1, B // 2 + 8, B - 1]
    for b in check:
        with km.Integrator(N, km.F32, atol=1e-6, c1=cs[b]) as it:
            it.set_omega(om_all[b])
            it.set_adjacency(Mq)
            it.set_state(th_all[b], 0.0, 0.01)
            single = it.run(te, t1, keep_rows=False)
        acc = single["status"].accepted
        if normal:
            assert abs(finals[b][1]["accepted"] - acc) <= max(1, int(0.02 * acc)), b
        assert len(res["r"][b]) == len(te)
        assert np.max(np.abs(res["r"][b] - single["r"])) < 1e-3, b
        # against the FP64 oracle on the same (dequantized) matrix
        ref = orc.integrate(orc.Adj(0, M=Mq.astype(np.float64)), om_all[b].astype(np.float64), cs[b],
                            th_all[b].astype(np.float64), 0.0, 0.01, t1, 1e-9, t_eval=te, keep_rows=False)
        assert np.max(np.abs(res["r"][b] - ref["r"])) < 1e-3, b


def test_shared_autotune_and_split(km):
    """The tile autotuner picks a variant; hi-only vectors stay within the
    order-parameter tolerance of hi + lo."""
    B, N = 8, 3000
    X = km.factor(N, 5, 5, gaussian=True).astype(np.float32)
    w = km.factor_weights(km.F32, N, 5, gaussian=True)
    te = _grid(1.0, 0.5)
    out = {}
    for split in (True, False):
        with km.Batch([N] * B, shared=True, storage="bf16") as bt:
            for b in range(B):
                th, om = km.draws(km.F32, N, 30 + b, normal=True, init=6.283185307)
                bt.set_instance(b, om, th, 3.0, 0.0, 0.01)
            bt.set_factor(-1, X, w)
            var, ks, ms = bt.autotune(2)
            assert var in range(6) and 1 <= ks <= 16 and ms > 0
            tab = bt.tune_table()
            assert tab.shape == (6, 16) and np.all(tab > 0) and ms == tab[var, ks - 1]
            bt.set_vector_split(split)
            out[split] = bt.run(te, 1.0)
    for b in range(B):
        assert np.max(np.abs(out[True]["r"][b] - out[False]["r"][b])) < 1e-3


def test_shared_f16_hi_only_matches_fp64_oracle(km, orc):
    """FP16 adjacency with FP16 cos/sin only in the product (one MFMA per
    tile, the batch64-n20000-f16-hi bench form): the runs complete and r(t)
    of every checked instance is within 1e-3 of the FP64 oracle on the same
    stored matrix.  Step counts are printed, not compared: with normal
    frequencies the steps here are h ~ 0.2, where the 11-bit cos/sin put the
    embedded error at a noise floor above atol = 1e-6 and some instances
    take more (smaller) steps than with hi + lo."""
    B, N = 16, 2048
    X = km.factor(N, 5, 12, gaussian=True).astype(np.float32)
    w = km.factor_weights(km.F32, N, 5, gaussian=True)
    cs = np.linspace(0.5, 4.0, B)
    t1, te = 1.0, _grid(1.0, 0.25)
    draws = [km.draws(km.F32, N, 200 + b, normal=True, init=6.283185307) for b in range(B)]
    print("case shared f16 hi-only", flush=True)
    out = {}
    for split in (True, False):
        with km.Batch([N] * B, atol=1e-6, shared=True, storage="f16") as bt:
            for b in range(B):
                bt.set_instance(b, draws[b][1], draws[b][0], cs[b], 0.0, 0.01)
            bt.set_factor(-1, X, w)
            bt.set_vector_split(split)
            Mq = bt.adjacency_rows(-1, 0, N)
            res = bt.run(te, t1, ring_rows=64, chunk=8)
            out[split] = (res, [bt.get_instance(b)[1] for b in range(B)], Mq)
    res, st, Mq = out[False]
    assert res["status"].active == 0 and res["status"].errors == 0
    print("  accepted hi+lo:", [int(x["accepted"]) for x in out[True][1]], flush=True)
    print("  accepted hi   :", [int(x["accepted"]) for x in st], flush=True)
    print("  rejected hi+lo:", [int(x["rejected"]) for x in out[True][1]], flush=True)
    print("  rejected hi   :", [int(x["rejected"]) for x in st], flush=True)
    for b in (0, 7, 15):
        ref = orc.integrate(orc.Adj(0, M=Mq.astype(np.float64)), draws[b][1].astype(np.float64), cs[b],
                            draws[b][0].astype(np.float64), 0.0, 0.01, t1, 1e-9, t_eval=te, keep_rows=False)
        assert np.max(np.abs(res["r"][b] - ref["r"])) < 1e-3, b


def test_attempt_log_and_proposer_fallback(km):
    """The attempt log records every attempt; a proposer (here: weights that
    propose 2x the standard step) is used after accepts, the embedded error
    still rejects, and rejections fall back to the standard controller —
    the run completes and r stays within 1e-3 of the standard run."""
    B, N = 4, 500
    P = km.proposer_param_count()
    w = np.zeros(P, np.float16)
    w[-1] = np.float16(0.69)  # output bias: log 2
    insts = [_instance(km, N, 50 + b, normal=True) for b in range(B)]
    te = _grid(2.0, 0.5)
    runs = {}
    for mode in ("standard", "proposer"):
        with km.Batch([N] * B, shared=False, storage="native") as bt:
            for b, (th, om, X, wf) in enumerate(insts):
                bt.set_instance(b, om, th, 3.0, 0.0, 0.01)
                bt.set_factor(b, X, wf)
            bt.set_attempt_log(100000)
            if mode == "proposer":
                bt.set_proposer(w, True)
            runs[mode] = bt.run(te, 2.0)
            log = bt.read_attempt_log()
        st = runs[mode]["status"]
        assert len(log) == st.accepted + st.rejected
        if mode == "proposer":
            assert st.proposer_accepts + st.proposer_rejects > 0
    for b in range(B):
        assert np.max(np.abs(runs["standard"]["r"][b] - runs["proposer"]["r"][b])) < 1e-3


@pytest.mark.parametrize("storage,B", [("bf16", 5), ("f16", 5), ("bf16", 40), ("bf16", 64)])
def test_shared_workgroup_variants_agree(km, storage, B):
    """The six workgroup shapes of the MFMA product sum every output in the
    same order (chunk, k-step, hi then lo) within a split-K slab: at a fixed
    split they give bit-identical runs (B = 5: one 32-row tile of V^T,
    B = 40: three, B = 64: four)."""
    N = 1500
    X = km.factor(N, 5, 9, gaussian=True).astype(np.float32)
    w = km.factor_weights(km.F32, N, 5, gaussian=True)
    te = _grid(1.0, 0.25)
    out = []
    for var in range(6):
        print(f"case shared variant {var} {storage} B={B}", flush=True)
        with km.Batch([N] * B, shared=True, storage=storage) as bt:
            for b in range(B):
                th, om = km.draws(km.F32, N, 40 + b, normal=True, init=6.283185307)
                bt.set_instance(b, om, th, 1.0 + b, 0.0, 0.01)
            bt.set_factor(-1, X, w)
            bt.select_tile(var, 4)
            res = bt.run(te, 1.0, chunk=8)
            out.append((res, [bt.get_instance(b) for b in range(B)]))
    for var in range(1, 6):
        for b in range(B):
            assert np.array_equal(out[var][0]["r"][b], out[0][0]["r"][b]), (var, b)
            assert np.array_equal(out[var][1][b][0], out[0][1][b][0]), (var, b)


@pytest.mark.parametrize("storage,B", [("bf16", 20), ("f16", 20), ("bf16", 48)])
def test_shared_row_compaction_is_bit_identical(km, storage, B):
    """Instances with Cauchy frequencies need many more steps than those with
    normal ones; once only the former are left (4 of 20 or 48 -> 8 rows), the
    shared product runs on one 32-row tile instead of two (B = 20) or three
    (B = 48).  Each output row depends only on its own row of V^T, so the run
    equals the uncompacted one bit for bit (same workgroup variant and
    split)."""
    N = 1000
    X = km.factor(N, 5, 21, gaussian=True).astype(np.float32)
    w = km.factor_weights(km.F32, N, 5, gaussian=True)
    te = _grid(2.0, 0.25)
    out = {}
    for on in (True, False):
        print(f"case compaction={on} {storage}", flush=True)
        with km.Batch([N] * B, atol=1e-6, shared=True, storage=storage) as bt:
            for b in range(B):
                th, om = km.draws(km.F32, N, 60 + b, normal=b >= 4, init=6.283185307)
                bt.set_instance(b, om, th, 0.5 + 0.15 * b, 0.0, 0.01)
            bt.set_factor(-1, X, w)
            bt.select_tile(3, 4)
            bt.set_compaction(on)
            assert bt.product_rows() == (32 * ((2 * B + 31) // 32), 0)
            res = bt.run(te, 2.0, chunk=8)
            out[on] = (res, [bt.get_instance(b) for b in range(B)], bt.product_rows())
    res_on, fin_on, rows_on = out[True]
    res_off, fin_off, rows_off = out[False]
    acc = [f[1]["accepted"] for f in fin_off]
    assert min(acc[:4]) > 2 * max(acc[4:])  # the Cauchy instances outlast the others
    assert rows_on[0] == 32 and rows_on[1] >= 1 and rows_off == (32 * ((2 * B + 31) // 32), 0)
    assert res_on["status"].active == 0 and res_on["status"].errors == 0
    for b in range(B):
        assert np.array_equal(res_on["r"][b], res_off["r"][b]), b
        assert np.array_equal(fin_on[b][0], fin_off[b][0]), b
        assert fin_on[b][1] == fin_off[b][1], b
