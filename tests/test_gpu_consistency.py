"""Design transfer between meshes and the consistency re-evaluation
(misc/get_consistent_objectives.py:110-145) on the GPU.

The interpolation is checked against a brute-force restatement on the oracle mesh: for every
fine vertex, the coarse triangle of the oracle's triangle list that contains it (barycentric
coordinates >= -1e-12) and its barycentric combination; a wrong diagonal would fail it.  The
re-evaluated objective is checked against the oracle's objective of the same fine design.
"""

import json

import numpy as np
import pytest
import torch

from tests.cases import case_from_design, oracle_problem, smoothed_density
from tests.conftest import design_path
from topomax_amd import StructuredMesh
from topomax_amd.consistency import consistent_objective

pytestmark = pytest.mark.gpu


def brute_force_interp(coarse, values, fine):
    from oracle.fem import StructuredMesh as OMesh
    m = OMesh(coarse.width, coarse.height, coarse.nx, coarse.ny)
    P = m.coords[m.triangles]                    # (T, 3, 2)
    out = np.empty(fine.n_vertices)
    xs = (fine.width / fine.nx) * np.arange(fine.nx + 1)
    ys = (fine.height / fine.ny) * np.arange(fine.ny + 1)
    X, Y = np.meshgrid(xs, ys)
    pts = np.stack([X.ravel(), Y.ravel()], axis=1)
    a, b, c = P[:, 0], P[:, 1], P[:, 2]
    det = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (c[:, 0] - a[:, 0]) * (b[:, 1] - a[:, 1])
    for i, (x, y) in enumerate(pts):
        l1 = ((x - a[:, 0]) * (c[:, 1] - a[:, 1]) - (c[:, 0] - a[:, 0]) * (y - a[:, 1])) / det
        l2 = ((b[:, 0] - a[:, 0]) * (y - a[:, 1]) - (x - a[:, 0]) * (b[:, 1] - a[:, 1])) / det
        l0 = 1 - l1 - l2
        t = int(np.nonzero((l0 >= -1e-12) & (l1 >= -1e-12) & (l2 >= -1e-12))[0][0])
        tri = m.triangles[t]
        out[i] = l0[t] * values[tri[0]] + l1[t] * values[tri[1]] + l2[t] * values[tri[2]]
    return out


@pytest.mark.parametrize("name,N,big", [("cantilever", 5, 20), ("triangle", 4, 12),
                                        ("short_cantilever", 10, 30)])
def test_design_interpolation(name, N, big):
    from topomax_amd import problem_from_design
    design = json.load(open(design_path(name)))
    prob, dom = problem_from_design(design, big)
    coarse = StructuredMesh.from_N(dom.width, dom.height, N)
    vals = np.random.default_rng(5).uniform(0, 1, coarse.n_vertices)
    got = prob.interpolate_design(torch.as_tensor(vals, device="cuda:0"), coarse).cpu().numpy()
    ref = brute_force_interp(coarse, vals, prob.mesh)
    print(name, np.abs(got - ref).max(), flush=True)
    assert np.abs(got - ref).max() < 1e-13


def test_consistent_objective_matches_oracle():
    design = json.load(open(design_path("cantilever")))
    N, big = 10, 40
    coarse = StructuredMesh.from_N(3.0, 1.0, N)
    rho_c = smoothed_density(coarse.nx, coarse.ny, seed=9, sigma=1.5, lo=0.05)
    J, rho_f = consistent_objective(design, N, rho_c, big_N=big, rtol=1e-12)
    rho_f = rho_f.cpu().numpy()
    fine = StructuredMesh.from_N(3.0, 1.0, big)
    assert np.abs(rho_f - brute_force_interp(coarse, rho_c, fine)).max() < 1e-13
    orc = oracle_problem(case_from_design(design), fine.nx, fine.ny)
    orc.set_penalization(max(design["Elasticity"]["domain_parameters"]["penalties"]))
    J_ref = orc.calculate_objective(rho_f)
    print("J", J, J_ref, flush=True)
    assert abs(J - J_ref) < 1e-8 * abs(J_ref)
