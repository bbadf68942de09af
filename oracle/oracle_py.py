"""ORACLE — test infrastructure only.

ctypes binding of oracle/_build/liboracle.so (oracle_capi.cpp) and the path
of the oracle command-line programs.  Imported only by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, always as the
checker or the reported CPU baseline, never as the product.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "liboracle.so")
CLI32 = os.path.join(BUILD, "kuramoto_oracle")
CLI64 = os.path.join(BUILD, "kuramoto_oracle_64")

_P = ctypes.c_void_p
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "-j4"], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_P)


def _real(dtype):
    return ctypes.c_float if np.dtype(dtype) == np.float32 else ctypes.c_double


def philox(ctr, key, rounds=10):
    out = (ctypes.c_uint32 * 4)()
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    load().orc_philox(c, k, rounds, out)
    return list(out)


def gaussian_factor(N, K, seed):
    X = np.zeros((N, K), np.float64)
    load().orc_gaussian_factor(ctypes.c_long(N), K, ctypes.c_uint32(seed), _p(X))
    return X


def rademacher_factor(N, K, seed):
    X = np.zeros((N, K), np.float64)
    load().orc_rademacher_factor(ctypes.c_long(N), K, ctypes.c_uint32(seed), _p(X))
    return X


def factor_weights(N, K, gauss, dtype=np.float64):
    w = np.zeros(max(K, 1), dtype)
    fn = load().orc_factor_weights_f32 if np.dtype(dtype) == np.float32 else load().orc_factor_weights_f64
    fn(ctypes.c_long(N), K, int(gauss), _p(w))
    return w[:K]


def adj_rows(X, w, i0, i1):
    X = np.ascontiguousarray(X)
    w = np.ascontiguousarray(w, dtype=X.dtype)
    N, K = X.shape
    out = np.zeros((i1 - i0, N), X.dtype)
    fn = load().orc_adj_rows_f32 if X.dtype == np.float32 else load().orc_adj_rows_f64
    fn(ctypes.c_long(N), K, _p(X), _p(w), ctypes.c_long(i0), ctypes.c_long(i1), _p(out))
    return out


def draws(N, seed, normal, freq=1.0, init=0.0, dtype=np.float32):
    ic = np.zeros(N, dtype)
    om = np.zeros(N, dtype)
    R = _real(dtype)
    fn = load().orc_draws_f32 if np.dtype(dtype) == np.float32 else load().orc_draws_f64
    fn(ctypes.c_long(N), seed, int(normal), R(freq), R(init), _p(ic), _p(om))
    return ic, om


def sincos(x, dtype=np.float32):
    """(cos, sin) of x in the working precision, as the oracle's model uses them."""
    x = np.ascontiguousarray(x, dtype)
    c = np.zeros_like(x)
    s = np.zeros_like(x)
    fn = load().orc_sincos_f32 if np.dtype(dtype) == np.float32 else load().orc_sincos_f64
    fn(ctypes.c_long(x.size), _p(x), _p(c), _p(s))
    return c, s


def tableau(dtype=np.float64):
    a = np.zeros(36, dtype)
    c = np.zeros(6, dtype)
    e = np.zeros(7, dtype)
    p = np.zeros(28, dtype)
    fn = load().orc_tableau_f32 if np.dtype(dtype) == np.float32 else load().orc_tableau_f64
    fn(_p(a), _p(c), _p(e), _p(p))
    return a.reshape(6, 6), c, e, p.reshape(7, 4)


class Adj:
    """Adjacency description for the oracle: kind 0 dense (M, mscale),
    1 dense rows rebuilt from the factor, 2 factored (-A)."""

    def __init__(self, kind, M=None, mscale=1.0, X=None, w=None):
        self.kind, self.M, self.mscale, self.X, self.w = kind, M, mscale, X, w

    def args(self, dtype):
        M = None if self.M is None else np.ascontiguousarray(self.M, dtype=dtype)
        X = None if self.X is None else np.ascontiguousarray(self.X, dtype=dtype)
        w = None if self.w is None else np.ascontiguousarray(self.w, dtype=dtype)
        K = 0 if w is None else w.size
        if X is not None and X.size == 0:
            X = np.zeros(1, dtype)
        if w is not None and w.size == 0:
            w = np.zeros(1, dtype)
        return M, X, w, K


def rhs(adj: Adj, omega, c1, theta, dtype=np.float64):
    dtype = np.dtype(dtype)
    R = _real(dtype)
    M, X, w, K = adj.args(dtype)
    om = np.ascontiguousarray(omega, dtype)
    th = np.ascontiguousarray(theta, dtype)
    N = th.size
    f = np.zeros(N, dtype)
    cs = np.zeros(2 * N, dtype)
    fn = load().orc_rhs_f32 if dtype == np.float32 else load().orc_rhs_f64
    fn(ctypes.c_long(N), adj.kind, _p(M), ctypes.c_double(adj.mscale), _p(X), _p(w), K, _p(om), R(c1), _p(th), _p(f),
       _p(cs))
    return f, cs


def integrate(adj: Adj, omega, c1, theta0, t0, h0, t1, atol, rtol=0.0, fixed=False, t_eval=(), keep_rows=True,
              times_cap=100000, max_steps=-1, dtype=np.float64, norms_cap=0):
    """dp45_run + step_eval's output walk on the CPU.  Returns a dict."""
    dtype = np.dtype(dtype)
    R = _real(dtype)
    M, X, w, K = adj.args(dtype)
    om = np.ascontiguousarray(omega, dtype)
    y = np.array(theta0, dtype=dtype, copy=True)
    N = y.size
    te = np.ascontiguousarray(t_eval, dtype)
    ne = te.size
    rows = np.zeros((ne, N), dtype) if keep_rows else None
    r = np.zeros(max(ne, 1), dtype)
    times = np.zeros(times_cap, dtype)
    counts = (ctypes.c_long * 4)()
    final = np.zeros(N, dtype)
    norms = np.zeros(max(norms_cap, 1), np.float64)
    t = R(t0)
    h = R(h0)
    fn = load().orc_integrate_f32 if dtype == np.float32 else load().orc_integrate_f64
    fn(ctypes.c_long(N), adj.kind, _p(M), ctypes.c_double(adj.mscale), _p(X), _p(w), K, _p(om), R(c1), _p(y),
       ctypes.byref(t), ctypes.byref(h), R(t1), R(atol), R(rtol), int(fixed), _p(te) if ne else None, ne, _p(rows),
       _p(r), _p(times), times_cap, max_steps, counts, _p(final), _p(norms), norms_cap)
    acc = counts[0]
    out = {
        "y": y, "t": t.value, "h": h.value, "accepted": acc, "rejected": counts[1], "rhs_evals": counts[2],
        "n_out": counts[3], "r": r[: counts[3]], "times": times[: min(acc, times_cap)], "final": final,
    }
    if keep_rows:
        out["rows"] = rows[: counts[3]]
    if norms_cap:
        out["norms"] = norms[: min(counts[0] + counts[1], norms_cap)]
    return out


def num_threads():
    return load().orc_num_threads()


def mlp_loss_grad(w, X, y):
    """FP64 proposer network: (loss, grad, out)."""
    w = np.ascontiguousarray(w, np.float64)
    X = np.ascontiguousarray(X, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    n = X.shape[0]
    g = np.zeros(load().orc_mlp_param_count(), np.float64)
    out = np.zeros(n, np.float64)
    fn = load().orc_mlp_loss_grad
    fn.restype = ctypes.c_double
    loss = fn(_p(w), _p(X), _p(y), n, _p(g), _p(out))
    return loss, g, out
