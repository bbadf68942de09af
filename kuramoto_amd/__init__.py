"""kuramoto_amd — Python bindings (ctypes) of the MI355X Kuramoto/DP45 library.

The library itself is C++/HIP behind the C ABI of ``include/kuramoto_amd.h``
(built into ``kuramoto_amd/lib/libkuramoto_amd.so``); this module only
declares that ABI for Python callers (tests, the benchmark).  It never falls
back to a CPU implementation: if the shared library is missing it raises.

Reference counterparts: the ``parameters`` context and callbacks of
kuramoto.cu:14-265 and the integrator API of dp45.h:7-11.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libkuramoto_amd.so")
BIN_DIR = os.path.join(_HERE, "bin")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "kuramoto_amd.h")

F32, F64 = 0, 1
ADJ_NATIVE, ADJ_BF16, ADJ_F16, ADJ_SYM = 0, 2, 3, 4
STORAGE = {"native": ADJ_NATIVE, "bf16": ADJ_BF16, "f16": ADJ_F16, "sym": ADJ_SYM,
           "sym-bf16": ADJ_SYM | ADJ_BF16, "sym-f16": ADJ_SYM | ADJ_F16}

ERRORS = {-1: "KMT_ERR_ARG", -2: "KMT_ERR_HIP", -3: "KMT_ERR_STATE", -4: "KMT_ERR_NUMERIC", -5: "KMT_ERR_MEMORY"}


class KmtError(RuntimeError):
    def __init__(self, code: int, what: str, msg: str):
        super().__init__(f"{what}: {ERRORS.get(code, code)}: {msg}")
        self.code = code


class Status(ctypes.Structure):
    _fields_ = [
        ("t", ctypes.c_double),
        ("h", ctypes.c_double),
        ("t_last", ctypes.c_double),
        ("accepted", ctypes.c_long),
        ("rejected", ctypes.c_long),
        ("attempts", ctypes.c_long),
        ("rhs_evals", ctypes.c_long),
        ("active", ctypes.c_int),
        ("error", ctypes.c_int),
        ("rows_ready", ctypes.c_int),
        ("times_ready", ctypes.c_int),
        ("eval_i", ctypes.c_int),
        ("pending", ctypes.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
_P = ctypes.c_void_p
_L = ctypes.c_long
_I = ctypes.c_int
_D = ctypes.c_double

_SIGNATURES = {
    "kmt_last_error": ([], ctypes.c_char_p),
    "kmt_device_count": ([_P], _I),
    "kmt_mem_info": ([_I, _P, _P], _I),
    "kmt_host_alloc": ([ctypes.c_size_t, _P], _I),
    "kmt_host_free": ([_P], _I),
    "kmt_draws": ([_I, _L, _I, _I, _D, _D, _P, _P], _I),
    "kmt_factor": ([_L, _I, ctypes.c_uint32, _I, _P], _I),
    "kmt_factor_weights": ([_I, _L, _I, _I, _P], _I),
    "kmt_create": ([_P, _I, _I, _L, _D, _D, _I, _D], _I),
    "kmt_destroy": ([_P], _I),
    "kmt_set_omega": ([_P, _P], _I),
    "kmt_set_state": ([_P, _P, _D, _D], _I),
    "kmt_get_state": ([_P, _P, _P, _P], _I),
    "kmt_set_adjacency": ([_P, _P, _I, _L, _I], _I),
    "kmt_set_factor": ([_P, _P, _P, _I], _I),
    "kmt_set_factor_sign": ([_P, _I, ctypes.c_uint32], _I),
    "kmt_build_adjacency": ([_P, _I], _I),
    "kmt_use_factored": ([_P, _I], _I),
    "kmt_get_adjacency_rows": ([_P, _L, _L, _P], _I),
    "kmt_get_adjacency_scale": ([_P, _P], _I),
    "kmt_rhs": ([_P, _P, _P, _P], _I),
    "kmt_rhs_noalloc": ([_P, _P, _P, _P, _P], _I),
    "kmt_order": ([_P, _P, _P], _I),
    "kmt_begin": ([_P, _P, _I, _D, _I, _I], _I),
    "kmt_advance": ([_P, _I, _P], _I),
    "kmt_read_outputs": ([_P, _I, _P, _P, _P], _I),
    "kmt_read_outputs_couplings": ([_P, _I, _P, _P, _P, _P], _I),
    "kmt_read_times": ([_P, _I, _P, _P, _P], _I),
    "kmt_step": ([_P, _P, _P, _P], _I),
    "kmt_eval": ([_P, _D, _P], _I),
    "kmt_time_stage_kernel": ([_P, _I, _P], _I),
    "kmt_time_attempts": ([_P, _I, _P, _P, _P], _I),
    "kmt_layout": ([_P, _P, _P, _P, _P], _I),
    "kmt_autotune": ([_P, _I, _P], _I),
    "kmt_select_variant": ([_P, _I, _I], _I),
    "kmt_batch_create": ([_P, _I, _I, _P, _D, _D, _I, _I, _I], _I),
    "kmt_batch_destroy": ([_P], _I),
    "kmt_batch_set_instance": ([_P, _I, _P, _P, _D, _D, _D], _I),
    "kmt_batch_set_factor": ([_P, _I, _P, _P, _I], _I),
    "kmt_batch_get_adjacency_rows": ([_P, _I, _L, _L, _P], _I),
    "kmt_batch_begin": ([_P, _P, _I, _D, _I], _I),
    "kmt_batch_advance": ([_P, _I, _P], _I),
    "kmt_batch_read_order": ([_P, _I, _I, _P, _P], _I),
    "kmt_batch_get_instance": ([_P, _I, _P, _P], _I),
    "kmt_batch_autotune": ([_P, _I, _P, _P, _P], _I),
    "kmt_batch_time_attempts": ([_P, _I, _P, _P, _P], _I),
    "kmt_batch_set_vector_split": ([_P, _I], _I),
    "kmt_batch_tune_table": ([_P, _P], _I),
    "kmt_batch_set_compaction": ([_P, _I], _I),
    "kmt_batch_product_rows": ([_P, _P, _P], _I),
    "kmt_batch_select_tile": ([_P, _I, _I], _I),
    "kmt_batch_set_proposer": ([_P, _P, _I, _I], _I),
    "kmt_proposer_param_count": ([], _I),
    "kmt_batch_set_attempt_log": ([_P, _I], _I),
    "kmt_batch_read_attempt_log": ([_P, _I, _P, _P], _I),
    "kmt_mlp_create": ([_P, _I, _P, ctypes.c_uint32], _I),
    "kmt_mlp_destroy": ([_P], _I),
    "kmt_mlp_get_weights": ([_P, _P, _P], _I),
    "kmt_mlp_forward": ([_P, _P, _I, _P], _I),
    "kmt_mlp_loss_grad": ([_P, _P, _P, _I, _P, _P], _I),
    "kmt_mlp_train": ([_P, _P, _P, _I, _I, _I, _D, _P], _I),
    "kmt_bandwidth_probe": ([_I, ctypes.c_size_t, _I, _P, _P], _I),
}


def load():
    """Load the shared library (raises ImportError when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() or make -C kuramoto_amd/csrc")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (args, res) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def _check(rc, what):
    if rc != 0:
        raise KmtError(rc, what, load().kmt_last_error().decode())


def _dtype(precision):
    return np.float32 if precision == F32 else np.float64


def device_count() -> int:
    n = ctypes.c_int(0)
    _check(load().kmt_device_count(ctypes.byref(n)), "kmt_device_count")
    return n.value


def bandwidth_probe(nbytes=8 << 30, reps=5, device=0):
    """(copy GB/s, read GB/s) of plain 16-byte streaming kernels."""
    c, r = ctypes.c_double(), ctypes.c_double()
    _check(load().kmt_bandwidth_probe(device, nbytes, reps, ctypes.byref(c), ctypes.byref(r)), "kmt_bandwidth_probe")
    return c.value, r.value


def draws(precision, N, seed, normal=False, freq=1.0, init=0.0, ics=True, freqs=True):
    """main()'s rand() draws: (theta0, omega) (kuramoto.cu:416, :495-541)."""
    dt = _dtype(precision)
    th = np.zeros(N, dt) if ics else None
    om = np.zeros(N, dt) if freqs else None
    _check(load().kmt_draws(precision, N, seed, int(normal), freq, init, _ptr(th), _ptr(om)), "kmt_draws")
    return th, om


def factor(N, K, seed, gaussian=True):
    X = np.zeros((N, K), np.float64)
    _check(load().kmt_factor(N, K, seed, int(gaussian), _ptr(X) if X.size else _ptr(np.zeros(1))), "kmt_factor")
    return X


def factor_weights(precision, N, K, gaussian=True):
    w = np.zeros(max(K, 1), _dtype(precision))
    _check(load().kmt_factor_weights(precision, N, K, int(gaussian), _ptr(w)), "kmt_factor_weights")
    return w[:K]


class Integrator:
    """One integrator context on one GPU (kmt_create ... kmt_destroy)."""

    def __init__(self, N, precision=F32, atol=1e-6, rtol=0.0, fixed=False, c1=3.0, device=0):
        self.lib = load()
        self.N = int(N)
        self.precision = precision
        self.dtype = _dtype(precision)
        self.ctx = ctypes.c_void_p()
        _check(self.lib.kmt_create(ctypes.byref(self.ctx), device, precision, self.N, atol, rtol, int(fixed), c1),
               "kmt_create")

    def close(self):
        if self.ctx:
            self.lib.kmt_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _vec(self, a, n=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if n is not None and a.size != n:
            raise ValueError(f"expected {n} values, got {a.size}")
        return a

    # ---- inputs ----
    def set_omega(self, omega):
        om = self._vec(omega, self.N)
        _check(self.lib.kmt_set_omega(self.ctx, _ptr(om)), "kmt_set_omega")

    def set_state(self, theta, t=0.0, h=1e-2):
        th = self._vec(theta, self.N)
        _check(self.lib.kmt_set_state(self.ctx, _ptr(th), t, h), "kmt_set_state")

    def get_state(self):
        th = np.zeros(self.N, self.dtype)
        t, h = ctypes.c_double(), ctypes.c_double()
        _check(self.lib.kmt_get_state(self.ctx, _ptr(th), ctypes.byref(t), ctypes.byref(h)), "kmt_get_state")
        return th, t.value, h.value

    def set_adjacency(self, M, storage="native"):
        M = np.ascontiguousarray(M, dtype=self.dtype)
        if M.shape != (self.N, self.N):
            raise ValueError("adjacency must be N x N")
        _check(self.lib.kmt_set_adjacency(self.ctx, _ptr(M), self.precision, self.N, STORAGE[storage]),
               "kmt_set_adjacency")

    def set_factor(self, X, w):
        X = np.ascontiguousarray(X, dtype=self.dtype)
        w = np.ascontiguousarray(w, dtype=self.dtype)
        K = w.size
        if X.shape != (self.N, K):
            raise ValueError("factor must be N x K")
        _check(self.lib.kmt_set_factor(self.ctx, _ptr(X) if K else None, _ptr(w) if K else None, K),
               "kmt_set_factor")

    def set_factor_sign(self, K, seed):
        """The volcano +-1 factor of rank K drawn on the device (packed bits),
        w_m = (-1)^m / N."""
        _check(self.lib.kmt_set_factor_sign(self.ctx, int(K), int(seed) & 0xFFFFFFFF), "kmt_set_factor_sign")

    def build_adjacency(self, storage="native"):
        _check(self.lib.kmt_build_adjacency(self.ctx, STORAGE[storage]), "kmt_build_adjacency")

    def use_factored(self, on=True):
        """-A coupling: True / "contract" = X diag(w) (X^T v); "rebuild" =
        adjacency tiles rebuilt in registers every stage; False = dense."""
        mode = {"contract": 1, "rebuild": 2}.get(on, None) if isinstance(on, str) else int(on)
        if mode is None:
            raise ValueError(f"unknown low-rank mode {on!r}")
        _check(self.lib.kmt_use_factored(self.ctx, mode), "kmt_use_factored")

    def adjacency_rows(self, i0, i1):
        out = np.zeros((i1 - i0, self.N), self.dtype)
        _check(self.lib.kmt_get_adjacency_rows(self.ctx, i0, i1, _ptr(out)), "kmt_get_adjacency_rows")
        return out

    def adjacency_scale(self):
        s = ctypes.c_double()
        _check(self.lib.kmt_get_adjacency_scale(self.ctx, ctypes.byref(s)), "kmt_get_adjacency_scale")
        return s.value

    # ---- evaluations ----
    def rhs(self, theta, couplings=False):
        th = self._vec(theta, self.N)
        f = np.zeros(self.N, self.dtype)
        cp = np.zeros(2 * self.N, self.dtype) if couplings else None
        _check(self.lib.kmt_rhs(self.ctx, _ptr(th), _ptr(f), _ptr(cp)), "kmt_rhs")
        return (f, cp) if couplings else f

    def rhs_noalloc(self, theta_dev, f_dev=None, couplings_dev=None, stream=None):
        """kmt_rhs_noalloc on device pointers (ints): no allocation and no
        synchronisation; the caller synchronises `stream` before reading."""
        _check(self.lib.kmt_rhs_noalloc(self.ctx, theta_dev, f_dev, couplings_dev, stream), "kmt_rhs_noalloc")

    def couplings(self, theta):
        th = self._vec(theta, self.N)
        cp = np.zeros(2 * self.N, self.dtype)
        _check(self.lib.kmt_rhs(self.ctx, _ptr(th), None, _ptr(cp)), "kmt_rhs")
        return cp

    def order(self, theta):
        th = self._vec(theta, self.N)
        r = ctypes.c_double()
        _check(self.lib.kmt_order(self.ctx, _ptr(th), ctypes.byref(r)), "kmt_order")
        return r.value

    # ---- integration ----
    def begin(self, t_eval, t1, keep_rows=True, ring_rows=64):
        te = np.ascontiguousarray(t_eval, dtype=np.float64)
        self.keep_rows = bool(keep_rows)
        self.ring_rows = int(ring_rows)
        _check(self.lib.kmt_begin(self.ctx, _ptr(te) if te.size else None, te.size, t1, int(keep_rows), ring_rows),
               "kmt_begin")

    def advance(self, max_attempts):
        st = Status()
        _check(self.lib.kmt_advance(self.ctx, max_attempts, ctypes.byref(st)), "kmt_advance")
        return st

    def read_outputs(self, max_rows=None):
        n_max = self.ring_rows if max_rows is None else max_rows
        rows = np.zeros((n_max, self.N), self.dtype) if self.keep_rows else None
        r = np.zeros(n_max, self.dtype)
        n = ctypes.c_int()
        _check(self.lib.kmt_read_outputs(self.ctx, n_max, _ptr(rows), _ptr(r), ctypes.byref(n)), "kmt_read_outputs")
        return (rows[: n.value] if rows is not None else None), r[: n.value]

    def read_outputs_couplings(self, max_rows=None):
        """read_outputs plus each row's couplings (C_i, S_i), shape (n, 2N),
        computed on the device from the rows in the ring (needs keep_rows)."""
        n_max = self.ring_rows if max_rows is None else max_rows
        rows = np.zeros((n_max, self.N), self.dtype)
        r = np.zeros(n_max, self.dtype)
        cpl = np.zeros((n_max, 2 * self.N), self.dtype)
        n = ctypes.c_int()
        _check(self.lib.kmt_read_outputs_couplings(self.ctx, n_max, _ptr(rows), _ptr(r), _ptr(cpl), ctypes.byref(n)),
               "kmt_read_outputs_couplings")
        return rows[: n.value], r[: n.value], cpl[: n.value]

    def read_times(self, max_n=4096):
        t = np.zeros(max_n, self.dtype)
        h = np.zeros(max_n, self.dtype)
        n = ctypes.c_int()
        _check(self.lib.kmt_read_times(self.ctx, max_n, _ptr(t), _ptr(h), ctypes.byref(n)), "kmt_read_times")
        return t[: n.value], h[: n.value]

    def step(self):
        acc = ctypes.c_int()
        t, h = ctypes.c_double(), ctypes.c_double()
        _check(self.lib.kmt_step(self.ctx, ctypes.byref(acc), ctypes.byref(t), ctypes.byref(h)), "kmt_step")
        return bool(acc.value), t.value, h.value

    def eval(self, t_eval):
        out = np.zeros(self.N, self.dtype)
        _check(self.lib.kmt_eval(self.ctx, t_eval, _ptr(out)), "kmt_eval")
        return out

    def time_stage_kernel(self, reps=10):
        ms = ctypes.c_double()
        _check(self.lib.kmt_time_stage_kernel(self.ctx, reps, ctypes.byref(ms)), "kmt_time_stage_kernel")
        return ms.value

    def time_attempts(self, attempts):
        region, stage = ctypes.c_double(), ctypes.c_double()
        st = Status()
        _check(self.lib.kmt_time_attempts(self.ctx, attempts, ctypes.byref(region), ctypes.byref(stage),
                                          ctypes.byref(st)), "kmt_time_attempts")
        return region.value, stage.value, st

    def autotune(self, reps=3):
        """Times the 12 stage-kernel variants; returns their ms as an array
        indexed [rows 4/8/16][non-temporal off/on][balanced grid off/on]."""
        ms = np.zeros(12)
        _check(self.lib.kmt_autotune(self.ctx, reps, _ptr(ms)), "kmt_autotune")
        return ms.reshape(3, 2, 2)

    def select_variant(self, rows, nontemporal=True, balanced=False, ring=False):
        _check(self.lib.kmt_select_variant(self.ctx, rows, int(nontemporal) | (2 if balanced else 0)
                                           | (4 if ring else 0)), "kmt_select_variant")

    def variant_flags(self):
        """kmt_layout's variant_flags: bit 0 non-temporal loads, bit 1
        balanced grid, bit 2 chunk ring."""
        fl = ctypes.c_int()
        _check(self.lib.kmt_layout(self.ctx, None, None, None, ctypes.byref(fl)), "kmt_layout")
        return fl.value

    def layout(self):
        """(lda, bytes per stored element, rows per group, non-temporal, balanced)."""
        lda, eb, rows, fl = ctypes.c_long(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(self.lib.kmt_layout(self.ctx, ctypes.byref(lda), ctypes.byref(eb), ctypes.byref(rows),
                                   ctypes.byref(fl)), "kmt_layout")
        return lda.value, eb.value, rows.value, bool(fl.value & 1), bool(fl.value & 2)

    def run(self, t_eval, t1, keep_rows=True, ring_rows=64, chunk=32, max_attempts=None):
        """dp45_run + step_eval's output walk: returns a dict with the phase
        rows at t_eval (if kept), r at t_eval, accepted-step times and h, and
        the final status."""
        self.begin(t_eval, t1, keep_rows, ring_rows)
        rows, rs, ts, hs = [], [], [], []
        done = 0
        while True:
            n = chunk if max_attempts is None else min(chunk, max_attempts - done)
            st = self.advance(n)
            done += n
            t, h = self.read_times()
            ts.append(t)
            hs.append(h)
            r_rows, r = self.read_outputs()
            if r_rows is not None:
                rows.append(r_rows)
            rs.append(r)
            if not st.active or (max_attempts is not None and done >= max_attempts):
                break
        out = {
            "times": np.concatenate(ts) if ts else np.zeros(0, self.dtype),
            "hs": np.concatenate(hs) if hs else np.zeros(0, self.dtype),
            "r": np.concatenate(rs) if rs else np.zeros(0, self.dtype),
            "status": st,
        }
        if keep_rows:
            out["rows"] = np.concatenate(rows) if rows else np.zeros((0, self.N), self.dtype)
        return out


class BatchStatus(ctypes.Structure):
    _fields_ = [
        ("accepted", ctypes.c_long),
        ("rejected", ctypes.c_long),
        ("attempts", ctypes.c_long),
        ("rhs_evals", ctypes.c_long),
        ("active", ctypes.c_int),
        ("errors", ctypes.c_int),
        ("rows_ready", ctypes.c_int),
        ("pad_", ctypes.c_int),
        ("t_min", ctypes.c_double),
        ("t_max", ctypes.c_double),
        ("osc_steps", ctypes.c_double),
        ("proposer_accepts", ctypes.c_long),
        ("proposer_rejects", ctypes.c_long),
    ]


def proposer_param_count():
    return load().kmt_proposer_param_count()


class Batch:
    """B FP32 instances advanced in lockstep (kmt_batch_*).  shared=True: one
    adjacency for all (BF16/FP16, MFMA); shared=False: one per instance."""

    def __init__(self, Ns, atol=1e-6, rtol=0.0, shared=True, storage="bf16", fixed=False, device=0):
        self.lib = load()
        self.Ns = [int(n) for n in Ns]
        self.B = len(self.Ns)
        self.ptr = ctypes.c_void_p()
        arr = np.ascontiguousarray(self.Ns, dtype=np.int64)
        _check(self.lib.kmt_batch_create(ctypes.byref(self.ptr), device, self.B, _ptr(arr), atol, rtol, int(shared),
                                         STORAGE[storage], int(fixed)), "kmt_batch_create")

    def close(self):
        if self.ptr:
            self.lib.kmt_batch_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_instance(self, b, omega, theta0, c1, t0=0.0, h0=0.01):
        om = np.ascontiguousarray(omega, np.float32)
        th = np.ascontiguousarray(theta0, np.float32)
        assert om.size == self.Ns[b] and th.size == self.Ns[b]
        _check(self.lib.kmt_batch_set_instance(self.ptr, b, _ptr(om), _ptr(th), c1, t0, h0), "kmt_batch_set_instance")

    def set_factor(self, b, X, w):
        X = np.ascontiguousarray(X, np.float32)
        w = np.ascontiguousarray(w, np.float32)
        _check(self.lib.kmt_batch_set_factor(self.ptr, b, _ptr(X), _ptr(w), w.size), "kmt_batch_set_factor")

    def adjacency_rows(self, b, i0, i1):
        n = self.Ns[0 if b < 0 else b]
        out = np.zeros((i1 - i0, n), np.float32)
        _check(self.lib.kmt_batch_get_adjacency_rows(self.ptr, b, i0, i1, _ptr(out)), "kmt_batch_get_adjacency_rows")
        return out

    def begin(self, t_eval, t1, ring_rows=256):
        te = np.ascontiguousarray(t_eval, np.float64)
        self.ring_rows = ring_rows
        _check(self.lib.kmt_batch_begin(self.ptr, _ptr(te) if te.size else None, te.size, t1, ring_rows),
               "kmt_batch_begin")

    def advance(self, n):
        st = BatchStatus()
        _check(self.lib.kmt_batch_advance(self.ptr, n, ctypes.byref(st)), "kmt_batch_advance")
        return st

    def read_order(self, b):
        r = np.zeros(self.ring_rows, np.float32)
        n = ctypes.c_int()
        _check(self.lib.kmt_batch_read_order(self.ptr, b, self.ring_rows, _ptr(r), ctypes.byref(n)),
               "kmt_batch_read_order")
        return r[: n.value]

    def get_instance(self, b):
        th = np.zeros(self.Ns[b], np.float32)
        st = np.zeros(8)
        _check(self.lib.kmt_batch_get_instance(self.ptr, b, _ptr(th), _ptr(st)), "kmt_batch_get_instance")
        keys = ("t", "h", "accepted", "rejected", "active", "error", "proposer_accepts", "proposer_rejects")
        return th, dict(zip(keys, st.tolist()))

    def autotune(self, reps=3):
        nt, ks, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(self.lib.kmt_batch_autotune(self.ptr, reps, ctypes.byref(nt), ctypes.byref(ks), ctypes.byref(ms)),
               "kmt_batch_autotune")
        return nt.value, ks.value, ms.value

    def select_tile(self, variant, split_k):
        _check(self.lib.kmt_batch_select_tile(self.ptr, variant, split_k), "kmt_batch_select_tile")

    def set_compaction(self, on):
        _check(self.lib.kmt_batch_set_compaction(self.ptr, 1 if on else 0), "kmt_batch_set_compaction")

    def product_rows(self):
        """(rows the shared product computes now, compactions so far)."""
        rows, n = ctypes.c_int(), ctypes.c_int()
        _check(self.lib.kmt_batch_product_rows(self.ptr, ctypes.byref(rows), ctypes.byref(n)),
               "kmt_batch_product_rows")
        return rows.value, n.value

    def tune_table(self):
        """ms per product, [workgroup shape 0..5][split-K 1..16]."""
        ms = np.zeros(96)
        _check(self.lib.kmt_batch_tune_table(self.ptr, _ptr(ms)), "kmt_batch_tune_table")
        return ms.reshape(6, 16)

    def time_attempts(self, attempts):
        region, coup = ctypes.c_double(), ctypes.c_double()
        st = BatchStatus()
        _check(self.lib.kmt_batch_time_attempts(self.ptr, attempts, ctypes.byref(region), ctypes.byref(coup),
                                                ctypes.byref(st)), "kmt_batch_time_attempts")
        return region.value, coup.value, st

    def set_vector_split(self, hi_lo=True):
        _check(self.lib.kmt_batch_set_vector_split(self.ptr, int(hi_lo)), "kmt_batch_set_vector_split")

    def set_proposer(self, weights=None, enabled=True):
        w = None if weights is None else np.ascontiguousarray(weights, np.float16).view(np.uint16)
        _check(self.lib.kmt_batch_set_proposer(self.ptr, _ptr(w), 0 if w is None else w.size, int(enabled)),
               "kmt_batch_set_proposer")

    def set_attempt_log(self, capacity):
        _check(self.lib.kmt_batch_set_attempt_log(self.ptr, capacity), "kmt_batch_set_attempt_log")

    def read_attempt_log(self, max_n=1 << 20):
        rec = np.zeros((max_n, 14), np.float32)
        n = ctypes.c_int()
        _check(self.lib.kmt_batch_read_attempt_log(self.ptr, max_n, _ptr(rec), ctypes.byref(n)),
               "kmt_batch_read_attempt_log")
        return rec[: n.value]

    def run(self, t_eval, t1, ring_rows=256, chunk=32, max_attempts=None):
        self.begin(t_eval, t1, ring_rows)
        rs = [[] for _ in range(self.B)]
        done = 0
        while True:
            n = chunk if max_attempts is None else min(chunk, max_attempts - done)
            st = self.advance(n)
            done += n
            for b in range(self.B):
                rs[b].append(self.read_order(b))
            if st.active == 0 or (max_attempts is not None and done >= max_attempts):
                break
        return {"r": [np.concatenate(x) for x in rs], "status": st}


class Mlp:
    """The proposer network on the GPU (kmt_mlp_*)."""

    def __init__(self, w_init=None, seed=1, device=0):
        self.lib = load()
        self.ptr = ctypes.c_void_p()
        w = None if w_init is None else np.ascontiguousarray(w_init, np.float32)
        _check(self.lib.kmt_mlp_create(ctypes.byref(self.ptr), device, _ptr(w), seed), "kmt_mlp_create")
        self.P = proposer_param_count()

    def close(self):
        if self.ptr:
            self.lib.kmt_mlp_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def weights(self):
        w32 = np.zeros(self.P, np.float32)
        w16 = np.zeros(self.P, np.uint16)
        _check(self.lib.kmt_mlp_get_weights(self.ptr, _ptr(w32), _ptr(w16)), "kmt_mlp_get_weights")
        return w32, w16.view(np.float16)

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros(X.shape[0], np.float32)
        _check(self.lib.kmt_mlp_forward(self.ptr, _ptr(X), X.shape[0], _ptr(out)), "kmt_mlp_forward")
        return out

    def loss_grad(self, X, y):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        loss = ctypes.c_double()
        g = np.zeros(self.P, np.float32)
        _check(self.lib.kmt_mlp_loss_grad(self.ptr, _ptr(X), _ptr(y), X.shape[0], ctypes.byref(loss), _ptr(g)),
               "kmt_mlp_loss_grad")
        return loss.value, g

    def train(self, X, y, epochs=10, batch=1024, lr=1e-3):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        hist = np.zeros(epochs)
        _check(self.lib.kmt_mlp_train(self.ptr, _ptr(X), _ptr(y), X.shape[0], epochs, batch, lr, _ptr(hist)),
               "kmt_mlp_train")
        return hist
