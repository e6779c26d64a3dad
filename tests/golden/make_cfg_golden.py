"""Write the oracle fixtures used by the GPU parity tests (run on the CPU, commits data only).

cfg1_n20.npz: 3x1 cantilever, N=20 (60x20 cells), E=1, nu=0.3, p=3, filter radius 0.05, Left
fixed, traction (0,-1) on Right at 0.5 / length 0.1; rho = 0.5 + 0.1 U(-1,1) per vertex for
seeds 0..2.  Full xi, u, w vectors and J (P1, FP64, SuperLU direct solves).

cfg2_n400.json: the same cantilever at N=400 (1200x400 cells, 963,202 DOFs), rho =
Gaussian-smoothed (sigma 5 cells) seeded noise mapped to [1e-3, 1], p=3: J, norms of u and w,
and 2000 seeded samples of u and w (the full vectors would be 11 MB).

cfg3_n1000.json: the same at N=1000 (3000x1000 cells, 6,008,002 DOFs), p=3 (the benchmark's
workload), solved with SuperLU in the grid's nested-dissection order (the generic orderings
run out of memory there; the process is capped at 56 GiB of address space).

    python tests/golden/make_cfg_golden.py [cfg1 cfg2 cfg3]
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


def cfg2(N=400, name="cfg2_n400.json", permc_spec=None):
    nx, ny = 3 * N, N
    t0 = time.time()
    if permc_spec:
        from oracle.fem import ElasticityOracle
        ElasticityOracle.permc_spec = permc_spec
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
        "workload": f"cantilever 3x1, N={N}, smoothed density seed 0 sigma 5 in [1e-3,1], p=3",
        "solver": "oracle P1 FP64, SuperLU direct" + (f" ({permc_spec} order)" if permc_spec
                                                      else ""),
        "seconds": time.time() - t0,
        "J": J, "u_norm": float(np.linalg.norm(u)), "u_max": float(np.abs(u).max()),
        "w_norm": float(np.linalg.norm(w)), "xi_sum": float(xi.sum()),
        "u_index": iu.tolist(), "u_values": u[iu].tolist(),
        "w_index": iw.tolist(), "w_values": w[iw].tolist(),
    }
    with open(os.path.join(HERE, name), "w") as fh:
        json.dump(data, fh)
    print(f"{name}: J={J!r} in {data['seconds']:.1f}s")


def cfg3():
    import resource
    cap = 56 << 30
    resource.setrlimit(resource.RLIMIT_AS, (cap, cap))
    cfg2(N=1000, name="cfg3_n1000.json", permc_spec="ND")


if __name__ == "__main__":
    which = sys.argv[1:] or ["cfg1", "cfg2"]
    for name in which:
        {"cfg1": cfg1, "cfg2": cfg2, "cfg3": cfg3}[name]()
