"""Write tests/golden/kat_k0.json from the reference's own fixtures.

Runs in the development container only (it reads /root/reference/data).
The fixture keeps only test constants, not the reference files:
  * SHA-256 of data/5000/0/1frequencies.dat (normal, seed 1) and
    data/5000_lorentz/0/1frequencies.dat (Cauchy, seed 1) — the frequency
    generator of kuramoto.cu:526-536 must reproduce these bit-exactly;
  * from data/*/0/1fs.dat (N=5000, K=0, c=1.75, t=100, dt=0.01, seed 1,
    -D 0; produced by job1.sh:24-25): the final t and h bits and 16 sampled
    (index, theta bits) pairs.  These files hold N+2 floats: the phases at
    the last accepted step, then t and h.
Provenance of the command line: job1.sh:24 (normal) and :25 (Cauchy).
"""
import hashlib
import json
import os

import numpy as np

REF = "/root/reference/data"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kat_k0.json")


def main():
    fx = {}
    for name, normal in (("5000", True), ("5000_lorentz", False)):
        freq = open(os.path.join(REF, name, "0", "1frequencies.dat"), "rb").read()
        fs = np.fromfile(os.path.join(REF, name, "0", "1fs.dat"), dtype=np.float32)
        N = 5000
        assert fs.size == N + 2
        idx = np.linspace(0, N - 1, 16).astype(int)
        fx[name] = {
            "cmd": f"-N 5000 -K 0 -s 1 -c 1.75 -t 100 -d 0.01 -D 0 {'-nvR' if normal else '-vR'}",
            "N": N,
            "K": 0,
            "seed": 1,
            "normal": normal,
            "c1": 1.75,
            "t1": 100.0,
            "dt": 0.01,
            "atol": 1e-6,
            "frequencies_sha256": hashlib.sha256(freq).hexdigest(),
            "t_bits": int(fs[N:N + 1].view(np.uint32)[0]),
            "h_bits": int(fs[N + 1:N + 2].view(np.uint32)[0]),
            "t": float(fs[N]),
            "h": float(fs[N + 1]),
            "theta_samples": [[int(i), int(fs[i:i + 1].view(np.uint32)[0])] for i in idx],
        }
    with open(OUT, "w") as f:
        json.dump(fx, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
