"""Shared workload definitions for the tests (seeded, small enough for the oracle)."""

from __future__ import annotations

import numpy as np

# the GPU-side workload builders live in the package (the benchmark uses them too)
from topomax_amd.workloads import CANTILEVER, batch_cases as cfg5_cases, smoothed_density  # noqa: E402,F401


def uniform_density(n_vertices: int, seed: int):
    """0.5 + 0.1 U(-1, 1) per vertex."""
    rng = np.random.default_rng(seed)
    return 0.5 + 0.1 * rng.uniform(-1.0, 1.0, n_vertices)


def oracle_problem(case: dict, nx: int, ny: int, degree: int = 1):
    from oracle.fem import ElasticityOracle, StructuredMesh
    mesh = StructuredMesh(case["width"], case["height"], nx, ny)
    return ElasticityOracle(mesh, young_modulus=case["young_modulus"],
                            poisson_ratio=case["poisson_ratio"],
                            filter_radius=case["filter_radius"],
                            fixed_sides=case["fixed_sides"], tractions=case.get("tractions"),
                            body_force=case.get("body_force"), degree=degree)


def gpu_problem(case: dict, nx: int, ny: int, precision: str = "fp64", rtol: float = 1e-12,
                degree: int = 1):
    from topomax_amd.workloads import problem_from_case
    return problem_from_case(case, nx, ny, precision=precision, rtol=rtol, degree=degree)


def case_from_design(design: dict) -> dict:
    (key,) = design.keys()
    dom = design[key]["domain_parameters"]
    par = design[key]["problem_parameters"]
    body = None
    if par.get("body_force") is not None:
        bf = par["body_force"]
        body = {"center": tuple(bf["region"]["center"]), "radius": bf["region"]["radius"],
                "value": tuple(bf["value"])}
    tractions = None
    if par.get("tractions") is not None:
        tractions = [{"side": t["side"], "center": t["center"], "length": t["length"],
                      "value": tuple(t["value"])} for t in par["tractions"]]
    return dict(width=dom["width"], height=dom["height"], young_modulus=par["young_modulus"],
                poisson_ratio=par["poisson_ratio"], filter_radius=par["filter_radius"],
                fixed_sides=par["fixed_sides"], tractions=tractions, body_force=body)
