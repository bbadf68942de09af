"""The algorithmic bytes per RHS that bench.py's roofline divides by the
measured launch time, for every storage form of the single-context
workloads (no GPU): the dense matrix s_A N^2, its upper triangle
s_A N (N + 1) / 2, the factor read twice (2 N K s_y), the packed signs read
twice (2 N K / 8), each plus the per-row vectors N (2 s_y + s_w)."""
import numpy as np
import pytest

import bench


def _r(N, dtype, elem_bytes):
    return {"N": N, "dtype": dtype, "elem_bytes": elem_bytes}


@pytest.mark.parametrize("name", sorted(bench.CONFIGS))
def test_rhs_bytes_every_config(name):
    cfg = bench.CONFIGS[name]
    dtype = np.float32 if cfg["prec"] == "f32" else np.float64
    s_y = np.dtype(dtype).itemsize
    N, K = cfg["N"], cfg["K"]
    st = cfg["storage"]
    elem = {"native": s_y, "sym": s_y, "bf16": 2, "f16": 2, "sym-bf16": 2, "sym-f16": 2}.get(st, s_y)
    got = bench.rhs_bytes(cfg, _r(N, dtype, elem))
    vec = N * 3 * s_y
    if st == "factored":
        want = 2 * N * K * s_y + vec
    elif st == "signs":
        want = N * K // 4 + vec
    elif st.startswith("sym"):
        want = elem * (N * N + N) // 2 + vec
    else:
        want = elem * N * N + vec
    assert got == want


def test_rhs_bytes_headline_value():
    # FP64 N = 50000 dense: 8 * 50000^2 + 50000 * 24 = 20,001,200,000 bytes
    cfg = bench.CONFIGS["fp64-n50000"]
    assert bench.rhs_bytes(cfg, _r(50000, np.float64, 8)) == 20_001_200_000
    # the same workload on its upper triangle: 10,001,400,000 bytes
    cfg = bench.CONFIGS["fp64-n50000-sym"]
    assert bench.rhs_bytes(cfg, _r(50000, np.float64, 8)) == 10_001_400_000
    # N = K = 10^5 packed signs, FP32 state: 2,501,200,000 bytes
    cfg = bench.CONFIGS["fp32-n100000-k100000-signs"]
    assert bench.rhs_bytes(cfg, _r(100000, np.float32, 4)) == 2_501_200_000
