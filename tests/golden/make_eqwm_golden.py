"""Writes tests/golden/eqwm_ref.json: input/output vectors of the reference's
own solve_u_eqwm (eqwm.f90:90-223), compiled from its source by
`make -C oracle ref` into oracle/_ref/libref_eqwm.so.  Run in the container
that has the reference sources; the JSON is the committed fixture."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_lib import ref_eqwm, ref_eqwm_available  # noqa: E402


def main():
    if not ref_eqwm_available():
        raise SystemExit("oracle/_ref/libref_eqwm.so missing: make -C oracle ref")
    rng = np.random.default_rng(2018)
    rows = []
    for ncv in (30, 100):
        for _ in range(60):
            u = float(10 ** rng.uniform(-1, 1.5))
            nu = float(10 ** rng.uniform(-5.5, -3))
            h = float(rng.uniform(0.02, 0.3))
            tin = float(10 ** rng.uniform(-4, -1))
            st = float(rng.choice([1.05, 1.10, 1.15]))
            tau, it = ref_eqwm(u, nu, h, tin, ncv=ncv, stretch=st)
            rows.append(dict(u=u, nu=nu, h=h, tau_in=tin, ncv=ncv, stretch=st, tau=tau, iter=it))
    # a turbulent probe point (u = 1, nu = 1e-4, h = 0.1) and a laminar point
    for (u, nu, h, tin, ncv) in [(1.0, 1e-4, 0.1, 1e-3, 30), (1e-3, 1e-3, 0.1, 1e-6, 30)]:
        tau, it = ref_eqwm(u, nu, h, tin, ncv=ncv, stretch=1.1)
        rows.append(dict(u=u, nu=nu, h=h, tau_in=tin, ncv=ncv, stretch=1.1, tau=tau, iter=it))
    with open(os.path.join(HERE, "eqwm_ref.json"), "w") as f:
        json.dump({"source": "eqwm.f90 solve_u_eqwm, rho=1, tol=1e-7, max_iter=100 (hard-coded)",
                   "rows": rows}, f, indent=0)
    print("wrote", len(rows), "rows")


if __name__ == "__main__":
    main()
