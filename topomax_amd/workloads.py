"""Seeded synthetic workloads of the elasticity path (the benchmark's designs).

* ``CANTILEVER``: the 3 x 1 cantilever of the N=20 / N=400 / N=1000 workloads (Left side
  fixed, traction (0, -1) on the Right side centred at 0.5 with length 0.1, E = 1, nu = 0.3,
  filter radius 0.05);
* ``smoothed_density``: Gaussian-smoothed (sigma in cells) seeded uniform noise on the node
  grid, min-max mapped to [lo, 1];
* ``batch_cases``: the batched workload's random designs (N in {40, ..., 320}, aspect 1-3,
  p 1-5, 1-3 tractions on random sides, optional circular body force, random fixed sides);
* ``problem_from_case``: an ``ElasticityProblem`` on cuda:0 for one of these case dicts.

The case dicts use the design files' field names (``designs/definitions.py:150-223``).
"""

from __future__ import annotations

import numpy as np
from scipy.ndimage import gaussian_filter

CANTILEVER = dict(width=3.0, height=1.0, young_modulus=1.0, poisson_ratio=0.3, filter_radius=0.05,
                  fixed_sides=["Left"],
                  tractions=[{"side": "Right", "center": 0.5, "length": 0.1, "value": (0.0, -1.0)}],
                  body_force=None)


def smoothed_density(nx: int, ny: int, seed: int, sigma: float = 5.0, lo: float = 1e-3):
    """Gaussian-smoothed (sigma in cells) uniform noise on the node grid, min-max mapped to
    [lo, 1], flattened in vertex order."""
    rng = np.random.default_rng(seed)
    a = gaussian_filter(rng.uniform(0.0, 1.0, (ny + 1, nx + 1)), sigma, mode="nearest")
    a = (a - a.min()) / (a.max() - a.min())
    return (lo + (1.0 - lo) * a).ravel()


def problem_from_case(case: dict, nx: int, ny: int, precision: str = "fp64",
                      rtol: float = 1e-12, degree: int = 1, device="cuda:0"):
    """``ElasticityProblem`` of an nx x ny mesh for a case dict (penalty left unset)."""
    from .designs import (CircularRegion, DomainParameters, ElasticityParameters, Force,
                          ProblemType, Side, Traction)
    from .elasticity_problem import ControlSpace, ElasticityProblem, StructuredMesh
    mesh = StructuredMesh(case["width"], case["height"], nx, ny)
    space = ControlSpace(mesh, device)
    tractions = None
    if case.get("tractions"):
        tractions = [Traction(Side.from_string(t["side"]), t["center"], t["length"],
                              tuple(t["value"])) for t in case["tractions"]]
    body = None
    if case.get("body_force"):
        bf = case["body_force"]
        body = Force(CircularRegion(tuple(bf["center"]), bf["radius"]), tuple(bf["value"]))
    params = ElasticityParameters([Side.from_string(s) for s in case["fixed_sides"]], body,
                                  tractions, case["filter_radius"], case["young_modulus"],
                                  case["poisson_ratio"])
    domain = DomainParameters(case["width"], case["height"], ProblemType.ELASTICITY, 25.0, 25.0,
                              [3.0], 0.5)
    return ElasticityProblem(mesh, space, domain, params, precision=precision, rtol=rtol,
                             degree=degree, device=device)


def batch_cases(n_designs: int = 64, seed: int = 2024):
    """The batched workload: designs with N uniform in {40, 80, ..., 320}, aspect W in {1, 2, 3}
    (H = 1), penalty p in {1..5}, 1-3 tractions on random sides, an optional circular body force
    (radius 0.1, random centre) and a random non-empty set of fixed sides.  Returns a list of
    (case, N, p) with seeded parameters."""
    rng = np.random.default_rng(seed)
    sides = ["Left", "Right", "Top", "Bottom"]
    out = []
    for _ in range(n_designs):
        N = int(rng.choice(np.arange(40, 321, 40)))
        W = float(rng.choice([1.0, 2.0, 3.0]))
        p = int(rng.integers(1, 6))
        tractions = []
        for _ in range(int(rng.integers(1, 4))):
            side = str(rng.choice(sides))
            extent = W if side in ("Top", "Bottom") else 1.0
            length = float(rng.uniform(0.05, 0.3)) * extent
            center = float(rng.uniform(length / 2, extent - length / 2))
            tractions.append({"side": side, "center": center, "length": length,
                              "value": tuple(float(v) for v in rng.uniform(-1, 1, 2))})
        body = None
        if rng.uniform() < 0.5:
            body = {"center": (float(rng.uniform(0.1, W - 0.1)), float(rng.uniform(0.1, 0.9))),
                    "radius": 0.1, "value": tuple(float(v) for v in rng.uniform(-1, 1, 2))}
        mask = int(rng.integers(1, 16))
        fixed = [s for i, s in enumerate(sides) if mask & (1 << i)]
        case = dict(width=W, height=1.0, young_modulus=1.0, poisson_ratio=0.3,
                    filter_radius=0.05, fixed_sides=fixed, tractions=tractions, body_force=body)
        out.append((case, N, p))
    return out
