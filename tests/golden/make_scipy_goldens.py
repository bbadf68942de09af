"""Write tests/golden/scipy_rk45.npz: small Kuramoto problems integrated by
scipy's solve_ivp, the reference's own independent CPU check
(volcano.ipynb cells at JSON lines 1100-1115: RK45, rtol=0, atol=1e-6, the
notebook's RHS omegas + sigma*(s*adj.dot(c) - c*adj.dot(s))).

Inputs are drawn with numpy (independent of the oracle's generators).  For
each case the file holds the inputs, scipy RK45 at the case's atol, and a
DOP853 solution at atol=1e-13 as the truth.  Regenerate with
    python tests/golden/make_scipy_goldens.py
"""
import os

import numpy as np
from scipy.integrate import solve_ivp

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scipy_rk45.npz")

CASES = [
    # name, N, coupling c1, atol, t1, seed, symmetric
    ("n8", 8, 1.75, 1e-8, 1.0, 11, True),
    ("n32", 32, 3.0, 1e-9, 1.0, 12, True),
    ("n64_nonsym", 64, 2.0, 1e-8, 2.0, 13, False),
]


def rhs(t, th, om, c1, M):
    c, s = np.cos(th), np.sin(th)
    return om + c1 * (s * (M @ c) - c * (M @ s))


def main():
    data = {}
    for name, N, c1, atol, t1, seed, sym in CASES:
        rng = np.random.default_rng(seed)
        M = rng.standard_normal((N, N)) / N
        if sym:
            M = 0.5 * (M + M.T)
        om = rng.standard_normal(N)
        th0 = rng.uniform(-np.pi, np.pi, N)
        te = np.linspace(0.1, t1, int(round(t1 / 0.1)))
        rk = solve_ivp(rhs, (0, t1), th0, method="RK45", rtol=1e-13, atol=atol, t_eval=te, args=(om, c1, M),
                       first_step=1e-3)
        tr = solve_ivp(rhs, (0, t1), th0, method="DOP853", rtol=1e-13, atol=1e-13, t_eval=te, args=(om, c1, M))
        for k, v in (("M", M), ("omega", om), ("theta0", th0), ("t_eval", te), ("rk45", rk.y.T), ("truth", tr.y.T),
                     ("params", np.array([c1, atol, t1]))):
            data[f"{name}_{k}"] = v
        data[f"{name}_rk45_steps"] = np.array([len(rk.t)])
    np.savez_compressed(OUT, **data)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
