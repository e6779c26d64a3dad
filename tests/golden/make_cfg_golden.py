"""Write the oracle fixtures used by the GPU parity tests (run on the CPU, commits data only).

cfg1_n20.npz: 3x1 cantilever, N=20 (60x20 cells), E=1, nu=0.3, p=3, filter radius 0.05, Left
fixed, traction (0,-1) on Right at 0.5 / length 0.1; rho = 0.5 + 0.1 U(-1,1) per vertex for
seeds 0..2.  Full xi, u, w vectors and J (P1, FP64, SuperLU direct solves).

cfg2_n400.json: the same cantilever at N=400 (1200x400 cells, 963,202 DOFs), rho =
Gaussian-smoothed (sigma 5 cells) seeded noise mapped to [1e-3, 1], p=3: J, norms of u and w,
and 2000 seeded samples of u and w (the full vectors would be 11 MB).

    python tests/golden/make_cfg_golden.py
"""

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tests.cases import CANTILEVER, oracle_problem, smoothed_density, uniform_density  # noqa: E402


def cfg1():
    nx, ny = 60, 20
    prob = oracle_problem(CANTILEVER, nx, ny)
    prob.set_penalization(3)
    out = {}
    for seed in range(3):
        rho = uniform_density((nx + 1) * (ny + 1), seed)
        J = prob.calculate_objective(rho)
        w = prob.calculate_objective_gradient()
        out[f"rho_{seed}"] = rho
        out[f"xi_{seed}"] = prob.filtered_rho
        out[f"u_{seed}"] = prob.u
        out[f"w_{seed}"] = w
        out[f"J_{seed}"] = np.array(J)
    np.savez_compressed(os.path.join(HERE, "cfg1_n20.npz"), **out)


def cfg2():
    N = 400
    nx, ny = 3 * N, N
    t0 = time.time()
    prob = oracle_problem(CANTILEVER, nx, ny)
    prob.set_penalization(3)
    rho = smoothed_density(nx, ny, seed=0, sigma=5.0)
    J = prob.calculate_objective(rho)
    w = prob.calculate_objective_gradient()
    u, xi = prob.u, prob.filtered_rho
    rng = np.random.default_rng(2024)
    iu = np.sort(rng.choice(u.size, 2000, replace=False))
    iw = np.sort(rng.choice(w.size, 2000, replace=False))
    data = {
        "workload": "cantilever 3x1, N=400, smoothed density seed 0 sigma 5 in [1e-3,1], p=3",
        "solver": "oracle P1 FP64, SuperLU direct",
        "seconds": time.time() - t0,
        "J": J, "u_norm": float(np.linalg.norm(u)), "u_max": float(np.abs(u).max()),
        "w_norm": float(np.linalg.norm(w)), "xi_sum": float(xi.sum()),
        "u_index": iu.tolist(), "u_values": u[iu].tolist(),
        "w_index": iw.tolist(), "w_values": w[iw].tolist(),
    }
    with open(os.path.join(HERE, "cfg2_n400.json"), "w") as fh:
        json.dump(data, fh)
    print(f"cfg2: J={J!r} in {data['seconds']:.1f}s")


if __name__ == "__main__":
    cfg1()
    cfg2()
