"""The C-ABI boundary on the CPU: the shared library loads, exports every
function include/kuramoto_amd.h declares, and its host-side generators
(no GPU work) agree bit-for-bit with the oracle.  No GPU compute calls."""
import hashlib
import json
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "kuramoto_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kmt_[a-z_0-9]+)\s*\(", text)))


def test_header_symbols_exported(km):
    names = _declared()
    assert len(names) >= 25
    lib = km.load()
    for n in names:
        assert hasattr(lib, n), n
    # the Python binding declares exactly the header's functions
    assert sorted(km._SIGNATURES) == names


def test_host_draws_bit_exact(km, orc):
    for prec, dt in ((km.F32, np.float32), (km.F64, np.float64)):
        for normal in (False, True):
            th, om = km.draws(prec, 3000, 5, normal=normal, freq=1.5, init=6.283185307)
            th2, om2 = orc.draws(3000, 5, normal, 1.5, 6.283185307, dtype=dt)
            assert np.array_equal(th.view(np.uint8), th2.view(np.uint8))
            assert np.array_equal(om.view(np.uint8), om2.view(np.uint8))
    # -R reload semantics: without the IC draws the frequencies use the
    # first rand() values
    _, om_skip = km.draws(km.F32, 100, 1, normal=False, ics=False)
    _, om_full = km.draws(km.F32, 100, 1, normal=False)
    assert not np.array_equal(om_skip, om_full)


def test_host_draws_match_reference_frequency_files(km):
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "kat_k0.json")))
    for name in ("5000", "5000_lorentz"):
        fx = gold[name]
        _, om = km.draws(km.F32, 5000, 1, normal=fx["normal"])
        assert hashlib.sha256(om.tobytes()).hexdigest() == fx["frequencies_sha256"]


@pytest.mark.parametrize("gauss", [True, False])
def test_host_factor_bit_exact(km, orc, gauss):
    X = km.factor(4001, 5, 9, gaussian=gauss)
    Xo = (orc.gaussian_factor if gauss else orc.rademacher_factor)(4001, 5, 9)
    assert np.array_equal(X, Xo)
    for prec, dt in ((km.F32, np.float32), (km.F64, np.float64)):
        assert np.array_equal(km.factor_weights(prec, 4001, 5, gauss), orc.factor_weights(4001, 5, gauss, dtype=dt))


def test_cli_help_without_gpu(km):
    for prog in ("kuramoto", "kuramoto_64"):
        out = subprocess.run([os.path.join(km.BIN_DIR, prog), "-h"], capture_output=True, text=True, timeout=60)
        assert out.returncode == 0
        assert out.stdout.startswith("usage:\tkuramoto [-hvnRFA] [-N N] [-K K] [-D D]")
        assert "filebase is base file name for output." in out.stdout


def test_errors_are_reported(km):
    """Bad arguments fail with a status code and a message (no GPU needed:
    kmt_create checks N before touching a device)."""
    with pytest.raises(km.KmtError) as e:
        km.Integrator(0, km.F32)
    assert "N must be >= 1" in str(e.value)
    with pytest.raises(km.KmtError):
        km.draws(7, 10, 1)
    with pytest.raises(km.KmtError) as e:
        km.bandwidth_probe(1000, 1)
    assert ">= 1 MiB" in str(e.value)
