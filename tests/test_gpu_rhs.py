"""GPU parity of the RHS kernels (fused gemv-pair stage kernel, factored -A
path, adjacency generator) against the oracle, through the C ABI.

Reference: dydt / makecoupling / tot_kuramoto / makey2 (kuramoto.cu:42-55,
:104-126) and getadj (kuramoto.cu:145-161).

Tolerances (stated per test): FP64 sums agree to 1e-12 of the row scale
sum_j |M_ij| (different summation order only); FP32 sums accumulate in
double on both sides and round once to float, so they are bit-identical
except where the two double sums straddle a float rounding boundary (at
most one float ulp, 1.2e-7 of the row scale; at least 99% of the values
bit-identical); 16-bit storage accumulates in FP32 (2e-6 of the row scale);
the adjacency generator is bit-exact (same fma chain on both sides).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 3, 63, 64, 65, 1000, 4097]


def _case(N, dtype, seed=0, symmetric=True):
    rng = np.random.default_rng(seed + N)
    M = rng.standard_normal((N, N)) / max(N, 1)
    if symmetric:
        M = 0.5 * (M + M.T)
    omega = rng.standard_normal(N)
    theta = rng.uniform(-np.pi, np.pi, N) * 3
    return M.astype(dtype), omega.astype(dtype), theta.astype(dtype)


def _scale(M, omega, c1):
    return np.abs(omega).astype(np.float64) + abs(c1) * np.abs(M.astype(np.float64)).sum(axis=1)


@pytest.mark.parametrize("N", SIZES)
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_rhs_dense_matches_oracle(km, orc, N, prec):
    dtype = np.float32 if prec == "f32" else np.float64
    p = km.F32 if prec == "f32" else km.F64
    M, omega, theta = _case(N, dtype)
    c1 = 1.75
    print(f"case rhs_dense N={N} prec={prec}", flush=True)
    with km.Integrator(N, p, c1=c1) as it:
        it.set_omega(omega)
        it.set_adjacency(M)
        f, cp = it.rhs(theta, couplings=True)
    fo, cso = orc.rhs(orc.Adj(0, M=M), omega, c1, theta, dtype=dtype)
    tol = (2.5e-7 if prec == "f32" else 1e-12) * _scale(M, omega, c1)
    assert np.all(np.abs(f.astype(np.float64) - fo) <= tol + 1e-30)
    rs = np.abs(M.astype(np.float64)).sum(axis=1)
    tolc = (1.2e-7 if prec == "f32" else 1e-12) * np.repeat(rs, 2)
    assert np.all(np.abs(cp.astype(np.float64) - cso) <= tolc + 1e-30)
    if prec == "f32":  # double accumulation rounded once: the same floats
        assert np.count_nonzero(cp == cso) >= 0.99 * cp.size
        assert np.count_nonzero(f == fo) >= 0.99 * f.size


def test_rhs_zero_adjacency_is_omega(km):
    """M = 0  =>  f = omega exactly (closed form)."""
    N = 777
    rng = np.random.default_rng(1)
    omega = rng.standard_normal(N).astype(np.float32)
    theta = rng.standard_normal(N).astype(np.float32)
    with km.Integrator(N, km.F32, c1=3.0) as it:
        it.set_omega(omega)
        it.set_adjacency(np.zeros((N, N), np.float32))
        f = it.rhs(theta)
    assert np.array_equal(f, omega)


def test_rhs_adler_pair(km):
    """N=2, M12=M21=a: phi' = dw + 2 c a sin(phi) with phi = th1 - th2
    (reference sign, kuramoto.cu:45)."""
    a, c1 = 0.3, 2.0
    M = np.array([[0.0, a], [a, 0.0]])
    omega = np.array([0.7, -0.2])
    for phi in np.linspace(-3, 3, 13):
        theta = np.array([0.4 + phi, 0.4])
        with km.Integrator(2, km.F64, c1=c1) as it:
            it.set_omega(omega)
            it.set_adjacency(M)
            f = it.rhs(theta)
        assert abs((f[0] - f[1]) - (omega[0] - omega[1] + 2 * c1 * a * np.sin(phi))) < 1e-14


@pytest.mark.parametrize("gauss", [True, False])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_build_adjacency_bit_exact(km, orc, gauss, prec):
    N, K, seed = 3001, 5, 7
    dtype = np.float32 if prec == "f32" else np.float64
    p = km.F32 if prec == "f32" else km.F64
    X = km.factor(N, K, seed, gaussian=gauss).astype(dtype)
    w = km.factor_weights(p, N, K, gaussian=gauss)
    print(f"case build_adjacency gauss={gauss} prec={prec}", flush=True)
    with km.Integrator(N, p) as it:
        it.set_factor(X, w)
        it.build_adjacency("native")
        rows = it.adjacency_rows(0, N)
    ref = orc.adj_rows(X, w, 0, N)
    assert np.array_equal(rows.view(np.uint8), ref.view(np.uint8))
    assert np.array_equal(rows, rows.T)  # exactly symmetric


@pytest.mark.parametrize("storage", ["bf16", "f16"])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_reduced_storage_rhs(km, orc, storage, prec):
    """BF16/FP16 storage with FP32 accumulation of the products equals the
    working-precision kernel applied to the dequantized matrix up to FP32
    accumulation (checked against the oracle on that matrix); the FP64
    program keeps the RHS epilogue in FP64."""
    N, K, seed, c1 = 2049, 5, 3, 3.0
    p = km.F32 if prec == "f32" else km.F64
    dtype = np.float32 if prec == "f32" else np.float64
    X = km.factor(N, K, seed, gaussian=True).astype(dtype)
    w = km.factor_weights(p, N, K, gaussian=True)
    rng = np.random.default_rng(5)
    omega = rng.standard_normal(N).astype(dtype)
    theta = rng.uniform(-np.pi, np.pi, N).astype(dtype)
    print(f"case reduced storage={storage} prec={prec}", flush=True)
    with km.Integrator(N, p, c1=c1) as it:
        it.set_omega(omega)
        it.set_factor(X, w)
        it.build_adjacency(storage)
        Mq = it.adjacency_rows(0, N)
        scale = it.adjacency_scale()
        f = it.rhs(theta)
    assert Mq.dtype == dtype
    assert scale > 0 and np.log2(scale) == np.round(np.log2(scale))
    fo, _ = orc.rhs(orc.Adj(0, M=Mq.astype(np.float64)), omega.astype(np.float64), c1, theta.astype(np.float64))
    tol = 2e-6 * _scale(Mq, omega, c1)
    assert np.all(np.abs(f.astype(np.float64) - fo) <= tol)
    # quantization error against the exact matrix stays at the format's epsilon
    M = orc.adj_rows(X, w, 0, N)
    rel = np.abs(Mq - M).max() / np.abs(M).max()
    assert rel < (2 ** -8 if storage == "bf16" else 2 ** -10)


@pytest.mark.parametrize("K", [0, 1, 5, 9, 20])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_factored_rhs_matches_oracle(km, orc, K, prec):
    N = 5003
    dtype = np.float32 if prec == "f32" else np.float64
    p = km.F32 if prec == "f32" else km.F64
    X = km.factor(N, K, 11, gaussian=False).astype(dtype)
    w = km.factor_weights(p, N, K, gaussian=False)
    rng = np.random.default_rng(K)
    omega = rng.standard_normal(N).astype(dtype)
    theta = rng.uniform(-np.pi, np.pi, N).astype(dtype)
    print(f"case factored K={K} prec={prec}", flush=True)
    with km.Integrator(N, p, c1=1.75) as it:
        it.set_omega(omega)
        it.set_factor(X, w)
        it.use_factored(True)
        f = it.rhs(theta)
    fo, _ = orc.rhs(orc.Adj(2, X=X, w=w), omega, 1.75, theta, dtype=dtype)
    tol = (2.5e-7 if prec == "f32" else 1e-13) * (np.abs(omega) + 1.75 * K)
    assert np.all(np.abs(f.astype(np.float64) - fo) <= tol + 1e-12 * (prec == "f32"))
    if prec == "f32":  # G and C, S in double on both sides, rounded once
        assert np.count_nonzero(f == fo) >= 0.99 * N


def test_order_parameter(km):
    N = 10001
    rng = np.random.default_rng(2)
    th = rng.uniform(-np.pi, np.pi, N)
    with km.Integrator(N, km.F64) as it:
        r = it.order(th)
    exact = abs(np.mean(np.exp(1j * th)))
    assert abs(r - exact) < 1e-12
    with km.Integrator(N, km.F64) as it:
        assert abs(it.order(np.full(N, 0.3)) - 1.0) < 1e-14


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("N", [1337, 6000])
def test_stage_variants_agree(km, orc, prec, N):
    """All 12 dense stage-kernel variants (rows per group 4/8/16 x
    non-temporal loads x balanced grid) sum every row in the same order: the
    RHS and couplings are bit-identical, and whole integrations agree (the
    balanced grid groups the error partials differently, so the norm may move
    in its last bit)."""
    p = km.F32 if prec == "f32" else km.F64
    dtype = np.float32 if prec == "f32" else np.float64
    X = km.factor(N, 5, 21, gaussian=True).astype(dtype)
    w = km.factor_weights(p, N, 5, gaussian=True)
    th, om = km.draws(p, N, 21, normal=True, init=6.283185307)
    te = np.arange(1, 5, dtype=np.float64) * 0.25
    print(f"case variants N={N} prec={prec}", flush=True)
    with km.Integrator(N, p, atol=1e-6 if prec == "f32" else 1e-9, c1=3.0) as it:
        it.set_omega(om)
        it.set_factor(X, w)
        it.build_adjacency()
        ref = None
        runs = {}
        for rows in (4, 8, 16):
            for nt in (False, True):
                for bal in (False, True):
                    it.select_variant(rows, nt, bal)
                    assert it.layout()[2:] == (rows, nt, bal)
                    f, cp = it.rhs(th, couplings=True)
                    if ref is None:
                        ref = (f, cp)
                    assert np.array_equal(f, ref[0]) and np.array_equal(cp, ref[1]), (rows, nt, bal)
                    if rows == 8 and nt:
                        it.set_state(th, 0.0, 0.01)
                        runs[bal] = it.run(te, 1.0, keep_rows=True)
    a, b = runs[False], runs[True]
    assert abs(a["status"].accepted - b["status"].accepted) <= 1
    assert np.max(np.abs(a["rows"][-1] - b["rows"][-1])) < (1e-4 if prec == "f32" else 1e-10)


@pytest.mark.parametrize("prec,storage", [("f32", "native"), ("f64", "native"), ("f32", "bf16"), ("f64", "bf16"),
                                          ("f32", "f16")])
@pytest.mark.parametrize("N", [700, 3333, 12345])
def test_stage_ring_is_bit_identical(km, prec, storage, N):
    """The chunk-ring stage kernel (variant flag 4: each row's next chunk
    loads while the current one is reduced; the last, partial chunk read as
    zeros through buffer bounds) sums every row in the stage kernel's order
    and groups the error partials the same way: RHS, couplings and a whole
    integration are the same bits as the plain variant of the same rows per
    group.  N = 700, 3333, 12345 leave a partial last chunk and a partial
    last row group."""
    p = km.F32 if prec == "f32" else km.F64
    dtype = np.float32 if prec == "f32" else np.float64
    X = km.factor(N, 5, 23, gaussian=True).astype(dtype)
    w = km.factor_weights(p, N, 5, gaussian=True)
    th, om = km.draws(p, N, 23, normal=True, init=6.283185307)
    te = np.arange(1, 5, dtype=np.float64) * 0.25
    print(f"case ring N={N} prec={prec} storage={storage}", flush=True)
    with km.Integrator(N, p, atol=1e-6 if prec == "f32" else 1e-9, c1=3.0) as it:
        it.set_omega(om)
        it.set_factor(X, w)
        it.build_adjacency(storage)
        for rows in (4, 8, 16):
            out = {}
            for ring in (False, True):
                it.select_variant(rows, True, False, ring)
                f, cp = it.rhs(th, couplings=True)
                it.set_state(th, 0.0, 0.01)
                g = it.run(te, 1.0, keep_rows=True)
                out[ring] = (f, cp, g)
            (f0, c0, g0), (f1, c1, g1) = out[False], out[True]
            assert np.array_equal(f0, f1) and np.array_equal(c0, c1), rows
            assert np.array_equal(g0["rows"], g1["rows"]) and np.array_equal(g0["times"], g1["times"]), rows


def test_autotune_keeps_a_valid_variant(km):
    """kmt_autotune times every variant (one workgroup per group with and
    without the chunk ring) on the context's matrix and keeps the fastest:
    the table is complete, the kept flags are a valid combination, and the
    RHS is the same bits as under the fixed default."""
    N = 3000
    X = km.factor(N, 5, 31, gaussian=True).astype(np.float32)
    w = km.factor_weights(km.F32, N, 5, gaussian=True)
    th, om = km.draws(km.F32, N, 31, normal=True, init=6.283185307)
    print("case autotune", flush=True)
    with km.Integrator(N, km.F32, c1=3.0) as it:
        it.set_omega(om)
        it.set_factor(X, w)
        it.build_adjacency()
        f0 = it.rhs(th)
        ms = it.autotune(3)
        assert ms.shape == (3, 2, 2) and np.all(ms > 0) and np.all(ms < 1e3)
        fl = it.variant_flags()
        assert fl & ~7 == 0 and (fl & 6) != 6 and it.layout()[2] in (4, 8, 16)
        assert np.array_equal(it.rhs(th), f0)
