"""GPU training kernels of the step-size proposer against the FP64 oracle
(oracle/mlp_ref.hpp) on the same FP16-rounded weights."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, 8)).astype(np.float32)
    y = (0.3 * np.tanh(X[:, 0]) - 0.2 * X[:, 3] + 0.1 * X[:, 6] * X[:, 7]).astype(np.float32)
    return X, y


def test_forward_and_gradient_match_oracle(km, orc):
    rng = np.random.default_rng(3)
    w0 = (rng.standard_normal(km.proposer_param_count()) * 0.15).astype(np.float32)
    X, y = _data(517, 4)
    with km.Mlp(w0) as m:
        w32, w16 = m.weights()
        out = m.forward(X)
        loss, g = m.loss_grad(X, y)
    wd = w16.astype(np.float64)  # the GPU computes with the FP16 copy
    lo, go, oo = orc.mlp_loss_grad(wd, X.astype(np.float64), y.astype(np.float64))
    assert np.max(np.abs(out - oo)) < 2e-5
    assert abs(loss - lo) < 1e-5 * max(1.0, lo)
    scale = np.abs(go).max()
    assert np.max(np.abs(g - go)) < 2e-4 * scale, np.max(np.abs(g - go)) / scale


def test_training_reduces_loss(km):
    X, y = _data(8192, 5)
    with km.Mlp(seed=7) as m:
        l0, _ = m.loss_grad(X, y)
        hist = m.train(X, y, epochs=30, batch=512, lr=3e-3)
        l1, _ = m.loss_grad(X, y)
        w32, w16 = m.weights()
    assert hist[-1] < 0.2 * hist[0] and l1 < 0.2 * l0
    assert np.all(np.isfinite(w32)) and np.array_equal(w16, w32.astype(np.float16))
