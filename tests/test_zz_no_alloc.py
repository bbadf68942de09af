"""kmt_rhs_noalloc: kmt_rhs on device vectors with no allocation, against
kmt_rhs (host vectors, the library's own buffers) through the C ABI.

The device buffers and the caller's stream come from the HIP runtime the
library itself uses (ctypes on libamdhip64.so.7), as a caller that manages
its own memory would provide them.  Two calls in a row, the second with a
smaller N (the shared scratch then holds the longer previous vector past
the new N, so the zero padding is exercised); outputs bit-identical to
kmt_rhs (same kernels, same summation order); the device's free memory is
unchanged across the second call (the first launch of a kernel may load its
code); a context too large for the scratch is refused with KMT_ERR_ARG.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so.7")
        for name, args in {"hipMalloc": [ctypes.c_void_p, ctypes.c_size_t], "hipFree": [ctypes.c_void_p],
                           "hipMemcpy": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int],
                           "hipStreamCreate": [ctypes.c_void_p], "hipStreamDestroy": [ctypes.c_void_p],
                           "hipStreamSynchronize": [ctypes.c_void_p],
                           "hipMemGetInfo": [ctypes.c_void_p, ctypes.c_void_p]}.items():
            fn = getattr(self.lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int

    def check(self, rc, what):
        assert rc == 0, f"{what} returned {rc}"

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        self.check(self.lib.hipMalloc(ctypes.byref(p), nbytes), "hipMalloc")
        return p.value

    def h2d(self, dst, a):
        self.check(self.lib.hipMemcpy(dst, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")

    def d2h(self, a, src):
        self.check(self.lib.hipMemcpy(a.ctypes.data, src, a.nbytes, 2), "hipMemcpy D2H")

    def free_bytes(self):
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        self.check(self.lib.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)), "hipMemGetInfo")
        return f.value


def _setup(km, it, N, dtype, mode, seed):
    rng = np.random.default_rng(seed)
    omega = rng.standard_normal(N).astype(dtype)
    theta = rng.uniform(-20, 20, N).astype(dtype)
    X = km.factor(N, 5, seed, gaussian=True).astype(dtype)
    w = km.factor_weights(km.F32 if dtype == np.float32 else km.F64, N, 5, gaussian=True)
    it.set_omega(omega)
    it.set_factor(X, w)
    if mode == "dense":
        it.build_adjacency()
    else:
        it.use_factored(mode)
    return theta


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("mode", ["dense", "contract", "rebuild"])
def test_rhs_noalloc_equals_rhs(km, prec, mode):
    hip = _Hip()
    p = km.F32 if prec == "f32" else km.F64
    dtype = np.float32 if prec == "f32" else np.float64
    s = ctypes.c_void_p()
    hip.check(hip.lib.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
    try:
        for call, N in enumerate((3001, 1000)):  # the second call is smaller
            print(f"case rhs_noalloc {prec} {mode} call {call} N={N}", flush=True)
            with km.Integrator(N, p, c1=1.75) as it:
                theta = _setup(km, it, N, dtype, mode, 40 + call)
                f_ref, cp_ref = it.rhs(theta, couplings=True)
                cp_only = it.couplings(theta)
                es = np.dtype(dtype).itemsize
                th_d, f_d, cp_d = hip.malloc(N * es), hip.malloc(N * es), hip.malloc(2 * N * es)
                try:
                    hip.h2d(th_d, theta)
                    free0 = hip.free_bytes()
                    it.rhs_noalloc(th_d, f_d, cp_d, s.value)
                    hip.check(hip.lib.hipStreamSynchronize(s), "hipStreamSynchronize")
                    free1 = hip.free_bytes()
                    f, cp = np.zeros(N, dtype), np.zeros(2 * N, dtype)
                    hip.d2h(f, f_d)
                    hip.d2h(cp, cp_d)
                    assert np.array_equal(f, f_ref) and np.array_equal(cp, cp_ref)
                    if call == 1:  # (a kernel's first launch may load its code)
                        assert free1 == free0, (free0, free1)
                    # couplings only, on the context's own stream
                    cp2 = np.zeros(2 * N, dtype)
                    hip.d2h(cp2, cp_d)  # (previous values)
                    it.rhs_noalloc(th_d, None, cp_d, None)
                    it.order(theta)  # synchronises the context's stream
                    hip.d2h(cp2, cp_d)
                    assert np.array_equal(cp2, cp_only)
                finally:
                    for ptr in (th_d, f_d, cp_d):
                        hip.lib.hipFree(ptr)
    finally:
        hip.lib.hipStreamDestroy(s)


def test_rhs_noalloc_refuses_contexts_beyond_its_scratch(km):
    for p, N in ((km.F64, 262145), (km.F32, 524289)):
        with km.Integrator(N, p) as it:
            with pytest.raises(km.KmtError) as e:
                it.rhs_noalloc(1, 1, None, None)
            assert e.value.code == -1 and "scratch" in str(e.value)
    with km.Integrator(262144, km.F64) as it:  # the largest FP64 size passes the size check
        with pytest.raises(km.KmtError) as e:
            it.rhs_noalloc(1, 1, None, None)
        assert e.value.code == -3  # no adjacency yet: refused before any launch
