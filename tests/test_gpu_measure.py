"""GPU checks of the measurement helpers that bench.py relies on (no
reference counterpart): the bandwidth probe, HIP-event timing of attempts,
and that kmt_advance / kmt_batch_advance run exactly the requested number of
attempts through the replayed graphs (8 and 1 attempts)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_bandwidth_probe(km):
    print("case probe 64 MiB", flush=True)
    copy, read = km.bandwidth_probe(64 << 20, 2)
    assert 100 < copy < 20000 and 100 < read < 20000


def test_advance_and_time_attempts_count_attempts(km):
    N = 3000
    X = km.factor(N, 5, 4, gaussian=True).astype(np.float32)
    th, om = km.draws(km.F32, N, 4, normal=True, init=6.283185307)
    print("case plain N=3000 advance / time_attempts", flush=True)
    with km.Integrator(N, km.F32, atol=1e-6, c1=2.0) as it:
        it.set_omega(om)
        it.set_factor(X, km.factor_weights(km.F32, N, 5, gaussian=True))
        it.build_adjacency()
        it.set_state(th, 0.0, 0.01)
        it.begin(np.arange(1, 101, dtype=np.float64), 100.0, keep_rows=False, ring_rows=256)
        st0 = it.advance(0)
        st1 = it.advance(13)  # one graph of 8 attempts, five of 1
        assert st1.attempts - st0.attempts == 13 and st1.active == 1
        region, stage, st2 = it.time_attempts(4)
        assert st2.attempts - st1.attempts == 4
        assert st2.rhs_evals - st1.rhs_evals == 6 * 4  # FSAL: six new RHS per attempt
        assert region > 0 and 0 < stage < region


def test_batch_advance_and_time_attempts(km):
    N, B = 1000, 3
    X = km.factor(N, 5, 6, gaussian=True).astype(np.float32)
    w = km.factor_weights(km.F32, N, 5, gaussian=True)
    print("case batch B=3 advance / time_attempts", flush=True)
    with km.Batch([N] * B, atol=1e-6, shared=True, storage="bf16") as bt:
        for b in range(B):
            th, om = km.draws(km.F32, N, 70 + b, normal=True, init=6.283185307)
            bt.set_instance(b, om, th, 1.0 + b, 0.0, 0.01)
        bt.set_factor(-1, X, w)
        bt.begin(np.arange(1, 101, dtype=np.float64), 100.0, ring_rows=256)
        s0 = bt.advance(0)
        s1 = bt.advance(9)
        assert s1.attempts - s0.attempts == 9 and s1.active == B
        region, coup, s2 = bt.time_attempts(3)
        assert s2.attempts - s1.attempts == 3
        assert region > 0 and 0 < coup < region
