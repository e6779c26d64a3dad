"""Design-consistency re-evaluation on the GPU path.

Restates ``misc/get_consistent_objectives.py:110-145``: a design found on a coarse mesh
(parameter N) is interpolated onto the big_N = 320 mesh of the same design file
(``df.interpolate`` of the CG1 design onto the fine CG1 space) and its objective is evaluated
there with the design's largest penalty.
"""

from __future__ import annotations

import pickle

import numpy as np

from .elasticity_problem import StructuredMesh, problem_from_design


def load_design(path: str):
    """Read a design file written by ``topomax_amd.emd.EMDSolver.save_rho`` (a pickled dict
    with N, domain_size, problem and the vertex-ordered nodal vector)."""
    with open(path, "rb") as fh:
        data = pickle.load(fh)
    if data.get("order") != "vertex":
        raise ValueError("design vector is not in vertex order (v = iy*(nx+1)+ix)")
    return data["N"], tuple(data["domain_size"]), np.asarray(data["vector"], dtype=np.float64)


def consistent_objective(design, N: int, rho_coarse, *, big_N: int = 320, degree: int = 1,
                         precision: str = "fp64", rtol: float = 1e-10, device=None):
    """Objective of the coarse design (mesh parameter N, nodal values in vertex order) on the
    big_N mesh with the largest penalty of the design.  Returns (objective, fine design)."""
    problem, domain = problem_from_design(design, big_N, precision=precision, rtol=rtol,
                                          degree=degree, device=device)
    problem.set_penalization(max(domain.penalties))
    coarse = StructuredMesh.from_N(domain.width, domain.height, N)
    rho = problem.interpolate_design(rho_coarse, coarse)
    return problem.calculate_objective(rho), rho
