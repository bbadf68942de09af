"""Pinning the oracle (CPU, no GPU).

Every oracle function is checked against something it does not itself
compute: the reference's own fixtures (data/*/0/1fs.dat, 1frequencies.dat,
reduced to the constants in tests/golden/kat_k0.json), scipy's published
DOPRI5 coefficients and solve_ivp trajectories (tests/golden/scipy_rk45.npz),
closed forms, and brute force on small inputs.
"""
import hashlib
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------- generators
def test_philox_known_answers(orc):
    """Random123 kat_vectors for philox4x32-10."""
    assert orc.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert orc.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert orc.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


@pytest.mark.parametrize("name", ["5000", "5000_lorentz"])
def test_frequency_kat(orc, name):
    """kuramoto.cu:416, :498-500, :526-536: srand(seed), N burned IC draws,
    then N frequency draws, typed exactly as the reference; the bytes of
    data/<name>/0/1frequencies.dat are pinned by their SHA-256."""
    fx = json.load(open(os.path.join(GOLD, "kat_k0.json")))[name]
    _, om = orc.draws(fx["N"], fx["seed"], fx["normal"], 1.0, 0.0, dtype=np.float32)
    assert hashlib.sha256(om.tobytes()).hexdigest() == fx["frequencies_sha256"]


def test_frequency_kat_discriminates_typing(orc):
    """A plausible mistake (double log instead of logf in the normal draw)
    changes the bytes: the FP64 generator's values rounded to float differ
    from the FP32 generator's in some entries."""
    _, om32 = orc.draws(5000, 1, True, 1.0, 0.0, dtype=np.float32)
    _, om64 = orc.draws(5000, 1, True, 1.0, 0.0, dtype=np.float64)
    assert not np.array_equal(om32, om64.astype(np.float32))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_initial_phase_draw_range_and_moments(orc, dtype):
    """kuramoto.cu:499 (kuramoto_64.cu the same): theta_j = I/2 (2 u - 1),
    u = rand()/RAND_MAX, i.e. uniform on [-I/2, I/2]: every draw inside it,
    the extremes reached, mean 0 and variance I^2/12 (the I/2 scale: a
    dropped halving would double the range and quadruple the variance),
    and I = 0 gives zeros."""
    N, I = 400000, 6.283185307
    th, _ = orc.draws(N, 11, False, 1.0, I, dtype=dtype)
    th = th.astype(np.float64)
    half = np.float64(dtype(I) / 2) if dtype == np.float32 else I / 2
    assert th.min() >= -half * (1 + 1e-7) and th.max() <= half * (1 + 1e-7)
    assert th.min() < -0.999 * half and th.max() > 0.999 * half
    sd = I / np.sqrt(12.0)
    assert abs(th.mean()) < 5 * sd / np.sqrt(N)
    assert abs(th.var() / (I * I / 12) - 1) < 0.01
    # uniform: a fourth moment of 9/5 sigma^4 (a normal would give 3)
    assert abs(np.mean(th ** 4) / sd ** 4 - 1.8) < 0.02
    th0, _ = orc.draws(1000, 11, False, 1.0, 0.0, dtype=dtype)
    assert not np.any(th0)


def test_gaussian_factor_moments(orc):
    X = orc.gaussian_factor(200000, 5, 3)
    assert abs(X.mean()) < 0.005 and abs(X.std() - 1) < 0.005
    assert abs(np.mean(X ** 4) - 3) < 0.05  # kurtosis of a normal
    assert abs(np.corrcoef(X[:, 0], X[:, 1])[0, 1]) < 0.01


def test_rademacher_factor(orc):
    X = orc.rademacher_factor(100003, 6, 1)
    assert set(np.unique(X)) == {-1.0, 1.0}
    assert abs(X.mean()) < 0.01


def test_adjacency_symmetric_and_weights(orc):
    N, K = 300, 5
    for gauss, make in ((True, orc.gaussian_factor), (False, orc.rademacher_factor)):
        X = make(N, K, 4)
        w = orc.factor_weights(N, K, gauss)
        M = orc.adj_rows(X, w, 0, N)
        assert np.array_equal(M, M.T)
        assert np.allclose(M, X @ np.diag(w) @ X.T, rtol=1e-12, atol=1e-16)
    # volcano weights (makeu2, kuramoto.cu:141): (-1)^m / N
    assert np.array_equal(orc.factor_weights(50, 4, False), np.array([1, -1, 1, -1]) / 50.0)


# ------------------------------------------------------------------ tableau
def test_tableau_equals_scipy_rk45(orc):
    from scipy.integrate._ivp.rk import RK45

    a, c, e, p = orc.tableau(np.float64)
    for s in range(5):
        assert np.array_equal(a[s, : s + 1], RK45.A[s + 1, : s + 1])
    assert np.array_equal(a[5, :6], RK45.B)
    assert np.array_equal(c, RK45.C)
    assert np.array_equal(e, RK45.E)
    assert np.array_equal(p, RK45.P)


def test_tableau_order_conditions(orc):
    a, c, e, p = orc.tableau(np.float64)
    assert abs(a[5].sum() - 1) < 1e-15               # sum b = 1
    assert abs(e.sum()) < 1e-17                       # sum e = 0 (b - bhat)
    for s in range(5):                                # c_{s+1} = sum_j a_{s+1,j}
        assert abs(a[s, : s + 1].sum() - c[s + 1]) < 1e-15
    b7 = np.append(a[5], 0.0)
    assert np.allclose(p.sum(axis=1), b7, atol=1e-15)  # dense output at x=1 = y_new
    assert abs(np.dot(a[5], c) - 0.5) < 1e-15          # second-order condition


def test_tableau_f32_is_rounded_f64(orc):
    a32, c32, e32, p32 = orc.tableau(np.float32)
    a64, c64, e64, p64 = orc.tableau(np.float64)
    assert np.array_equal(a32, a64.astype(np.float32))
    assert np.array_equal(e32, e64.astype(np.float32))
    assert np.array_equal(p32, p64.astype(np.float32))


# -------------------------------------------------------------------- RHS
def _brute(M, om, c1, th):
    """c1 * sum_j M_ij sin(th_i - th_j) + om_i (reference sign, kuramoto.cu:45)."""
    return om + c1 * (M * np.sin(th[:, None] - th[None, :])).sum(axis=1)


@pytest.mark.parametrize("N", [1, 2, 7, 64, 301])
def test_rhs_brute_force(orc, N):
    rng = np.random.default_rng(N)
    M = rng.standard_normal((N, N)) / N
    om = rng.standard_normal(N)
    th = rng.uniform(-4, 4, N)
    f, cs = orc.rhs(orc.Adj(0, M=M), om, 1.75, th)
    assert np.allclose(f, _brute(M, om, 1.75, th), rtol=0, atol=1e-13)
    assert np.allclose(cs[0::2], M @ np.cos(th), atol=1e-14)
    assert np.allclose(cs[1::2], M @ np.sin(th), atol=1e-14)


def test_rhs_sign_is_pinned(orc):
    """Adler pair: phi' = dw + 2 c a sin(phi).  A flipped sign (the
    textbook c/N sin(theta_j - theta_i) form without the generator's sign
    flip) fails this."""
    a, c1 = 0.4, 2.0
    M = np.array([[0.0, a], [a, 0.0]])
    om = np.array([0.3, -0.5])
    for phi in np.linspace(-3, 3, 7):
        th = np.array([phi + 0.2, 0.2])
        f, _ = orc.rhs(orc.Adj(0, M=M), om, c1, th)
        assert abs((f[0] - f[1]) - (0.8 + 2 * c1 * a * np.sin(phi))) < 1e-14
        f_flip, _ = orc.rhs(orc.Adj(0, M=M), om, -c1, th)
        if abs(np.sin(phi)) > 1e-3:
            assert abs((f_flip[0] - f_flip[1]) - (0.8 + 2 * c1 * a * np.sin(phi))) > 1e-3


def test_rhs_zero_and_symmetric_invariant(orc):
    N = 400
    rng = np.random.default_rng(3)
    om = rng.standard_normal(N)
    th = rng.uniform(-3, 3, N)
    f, _ = orc.rhs(orc.Adj(0, M=np.zeros((N, N))), om, 3.0, th)
    assert np.array_equal(f, om)
    X = orc.gaussian_factor(N, 5, 1)
    w = orc.factor_weights(N, 5, True)
    M = orc.adj_rows(X, w, 0, N)
    f, _ = orc.rhs(orc.Adj(0, M=M), om, 3.0, th)
    assert abs(f.sum() - om.sum()) < 1e-12  # symmetric M: sum_i f_i = sum_i w_i


def test_fp32_sincos_is_the_rounded_double_function(orc):
    """FP32 cos / sin of the model are the float32 roundings of the float64
    functions (numpy's), over the phase range of Cauchy runs (|theta| up to
    1e5, where one float ulp is 8e-3 rad).  Float-library cosf would differ
    from these in some last bits."""
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(-4, 4, 100000), rng.uniform(-1e5, 1e5, 100000)]).astype(np.float32)
    c, s = orc.sincos(x, np.float32)
    x64 = x.astype(np.float64)
    assert np.array_equal(c, np.cos(x64).astype(np.float32))
    assert np.array_equal(s, np.sin(x64).astype(np.float32))
    c64, s64 = orc.sincos(x64, np.float64)
    assert np.max(np.abs(c64 - np.cos(x64))) <= 2.3e-16 and np.max(np.abs(s64 - np.sin(x64))) <= 2.3e-16


def test_fp32_coupling_sums_round_once(orc):
    """FP32 coupling sums C_i, S_i accumulate in double and round once to
    float: they equal the float32 rounding of numpy's float64 sums (another
    summation order) except where the two double sums straddle a float
    rounding boundary (none here).  A float accumulation would miss this at
    the 1e-7 level in most rows."""
    N = 3000
    X = orc.gaussian_factor(N, 5, 4).astype(np.float32)
    w = orc.factor_weights(N, 5, True, np.float32)
    M = orc.adj_rows(X, w, 0, N)
    rng = np.random.default_rng(5)
    th = rng.uniform(-50, 50, N).astype(np.float32)
    om = rng.standard_normal(N).astype(np.float32)
    f, cs = orc.rhs(orc.Adj(0, M=M), om, 3.0, th, dtype=np.float32)
    c, s = orc.sincos(th, np.float32)
    M64 = M.astype(np.float64)
    assert np.array_equal(cs[0::2], (M64 @ c.astype(np.float64)).astype(np.float32))
    assert np.array_equal(cs[1::2], (M64 @ s.astype(np.float64)).astype(np.float32))
    f32acc = np.zeros(N, np.float32)
    for j in range(N):  # float accumulation, column order
        f32acc += M[:, j] * c[j]
    assert np.count_nonzero(f32acc != cs[0::2]) > N // 2


@pytest.mark.parametrize("K", [0, 1, 5, 12])
def test_rhs_factored_equals_dense(orc, K):
    N = 257
    X = orc.rademacher_factor(N, K, 2)
    w = orc.factor_weights(N, K, False)
    M = orc.adj_rows(X, w, 0, N) if K else np.zeros((N, N))
    rng = np.random.default_rng(K)
    om = rng.standard_normal(N)
    th = rng.uniform(-3, 3, N)
    fd, _ = orc.rhs(orc.Adj(0, M=M), om, 1.75, th)
    ff, _ = orc.rhs(orc.Adj(2, X=X, w=w), om, 1.75, th)
    fr, _ = orc.rhs(orc.Adj(1, X=X, w=w), om, 1.75, th)
    assert np.allclose(fd, ff, atol=1e-13) and np.array_equal(fd, fr)


# ------------------------------------------------------- integrator + KAT
@pytest.mark.parametrize("name,steps", [("5000", 12), ("5000_lorentz", 43)])
def test_k0_controller_kat(orc, name, steps):
    """data/<name>/0/1fs.dat (N=5000, K=0, c=1.75, t=100, dt=0.01, seed 1):
    final t bit-exact, h within 4 ulp, phases at the last accepted step
    within 8 ulp.  The accepted-step count (12 / 43) is what t pins."""
    fx = json.load(open(os.path.join(GOLD, "kat_k0.json")))[name]
    N = fx["N"]
    th0, om = orc.draws(N, fx["seed"], fx["normal"], 1.0, 0.0, dtype=np.float32)
    dt = np.float32(fx["dt"])
    te = (dt * np.arange(1, int(np.float32(fx["t1"]) / dt) + 1)).astype(np.float32)
    res = orc.integrate(orc.Adj(2, X=np.zeros((N, 0)), w=np.zeros(0)), om, fx["c1"], th0, 0.0,
                        float(dt / np.float32(100)), fx["t1"], fx["atol"], t_eval=te, keep_rows=False,
                        dtype=np.float32)
    assert np.float32(res["t"]).view(np.uint32) == fx["t_bits"]
    assert abs(int(np.float32(res["h"]).view(np.int32)) - int(np.uint32(fx["h_bits"]).view(np.int32))) <= 4
    for i, bits in fx["theta_samples"]:
        assert abs(int(res["y"][i:i + 1].view(np.int32)[0]) - int(np.uint32(bits).view(np.int32))) <= 8
    assert res["accepted"] == steps and res["rejected"] == 0
    assert res["n_out"] == len(te)


def test_k0_order_parameter_closed_form(orc):
    """K=0: theta_i(t) = w_i t exactly up to the integrator's error, so
    r(t) = |mean exp(i w t)|."""
    N = 2000
    th0, om = orc.draws(N, 3, True, 1.0, 0.0, dtype=np.float64)
    te = np.arange(1, 11) * 0.5
    res = orc.integrate(orc.Adj(2, X=np.zeros((N, 0)), w=np.zeros(0)), om, 1.0, th0, 0.0, 1e-3, 5.0, 1e-10,
                        t_eval=te)
    r_exact = np.abs(np.exp(1j * np.outer(te, om)).mean(axis=1))
    assert np.allclose(res["r"], r_exact, atol=1e-9)
    assert np.allclose(res["rows"], np.outer(te, om), atol=1e-8)


def test_dense_output_exact_for_constant_rhs(orc):
    """f = const: every interpolated phase is ylast + (t_eval - t_last) w."""
    N = 50
    om = np.linspace(-2, 2, N)
    te = np.linspace(0.05, 3.0, 60)
    res = orc.integrate(orc.Adj(2, X=np.zeros((N, 0)), w=np.zeros(0)), om, 1.0, np.zeros(N), 0.0, 0.01, 3.0, 1e-8,
                        t_eval=te)
    assert np.allclose(res["rows"], np.outer(te, om), rtol=0, atol=1e-13)


def test_fixed_step_count(orc):
    N = 20
    om = np.ones(N)
    res = orc.integrate(orc.Adj(2, X=np.zeros((N, 0)), w=np.zeros(0)), om, 1.0, np.zeros(N), 0.0, 0.125, 1.0, 1e-6,
                        fixed=True)
    assert res["accepted"] == 8 and res["rejected"] == 0 and res["t"] == 1.0
    # no clamp of the last step (dp45.cu:153-155): 10 x 0.1 sums to just
    # below 1, so an 11th step is taken
    res = orc.integrate(orc.Adj(2, X=np.zeros((N, 0)), w=np.zeros(0)), om, 1.0, np.zeros(N), 0.0, 0.1, 1.0, 1e-6,
                        fixed=True)
    assert res["accepted"] == 11 and res["t"] > 1.0


@pytest.mark.parametrize("case", ["n8", "n32", "n64_nonsym"])
def test_trajectory_against_scipy(orc, case):
    """tests/golden/scipy_rk45.npz: the reference's own cross-check
    (solve_ivp RK45, volcano.ipynb).  The oracle's DP45 at the same atol
    lands as close to the DOP853 truth as scipy's RK45 does (same method;
    scipy differs in the initial step, the cap after a reject and the end
    clamp, so only trajectory-level agreement is expected)."""
    g = np.load(os.path.join(GOLD, "scipy_rk45.npz"))
    M, om, th0, te = g[f"{case}_M"], g[f"{case}_omega"], g[f"{case}_theta0"], g[f"{case}_t_eval"]
    c1, atol, t1 = g[f"{case}_params"]
    res = orc.integrate(orc.Adj(0, M=M), om, c1, th0, 0.0, 1e-3, t1, atol, t_eval=te)
    truth, rk = g[f"{case}_truth"], g[f"{case}_rk45"]
    e_orc = np.abs(res["rows"] - truth).max()
    e_rk = np.abs(rk - truth).max()
    assert e_orc < max(20 * e_rk, 50 * atol), (e_orc, e_rk)
    assert res["rows"].shape == truth.shape
    assert abs(res["accepted"] - int(g[f"{case}_rk45_steps"][0])) < 0.5 * res["accepted"] + 5


def test_resume_bit_exact(orc):
    """Restarting from (y, t, h) of an accepted step continues bit-exactly
    (the reference recomputes k1 = f(y) on -R, equal to the FSAL k7)."""
    N = 100
    X = orc.gaussian_factor(N, 5, 2)
    w = orc.factor_weights(N, 5, True)
    M = orc.adj_rows(X, w, 0, N)
    th0, om = orc.draws(N, 2, True, 1.0, 6.283185307, dtype=np.float64)
    full = orc.integrate(orc.Adj(0, M=M), om, 3.0, th0, 0.0, 0.01, 2.0, 1e-9)
    part = orc.integrate(orc.Adj(0, M=M), om, 3.0, th0, 0.0, 0.01, 2.0, 1e-9, max_steps=7)
    rest = orc.integrate(orc.Adj(0, M=M), om, 3.0, part["y"], part["t"], part["h"], 2.0, 1e-9)
    assert rest["t"] == full["t"] and np.array_equal(rest["y"], full["y"])
    assert part["accepted"] + rest["accepted"] == full["accepted"]


# ------------------------------------------------------- proposer network
def test_mlp_oracle_gradient_finite_differences(orc):
    """Back-propagated gradient of the FP64 network against central finite
    differences on a sample of parameters of every layer."""
    rng = np.random.default_rng(0)
    P = 4801
    w = rng.standard_normal(P) * 0.2
    X = rng.standard_normal((40, 8))
    y = rng.standard_normal(40) * 0.3
    loss, g, _ = orc.mlp_loss_grad(w, X, y)
    idx = [0, 7, 300, 511, 512, 575, 576, 1000, 4671, 4672, 4735, 4736, 4799, 4800]
    for p in idx:
        e = np.zeros(P)
        e[p] = 1e-6
        fd = (orc.mlp_loss_grad(w + e, X, y)[0] - orc.mlp_loss_grad(w - e, X, y)[0]) / 2e-6
        assert abs(fd - g[p]) < 1e-7 + 1e-5 * abs(fd), (p, fd, g[p])


def test_mlp_oracle_zero_network(orc):
    """All-zero weights except b3: out = b3, dL/db3 = 2 mean(b3 - y), the
    output-layer gradient is zero (h2 = 0) and nothing flows back."""
    w = np.zeros(4801)
    w[-1] = 0.25
    X = np.random.default_rng(1).standard_normal((10, 8))
    y = np.linspace(-1, 1, 10)
    loss, g, out = orc.mlp_loss_grad(w, X, y)
    assert np.allclose(out, 0.25)
    assert abs(loss - np.mean((0.25 - y) ** 2)) < 1e-15
    assert abs(g[-1] - 2 * np.mean(0.25 - y)) < 1e-15
    assert np.all(g[:-1] == 0)
