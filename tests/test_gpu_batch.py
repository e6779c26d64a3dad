"""Batched objective + gradient (tmx_batch_*) over uneven designs: different domains, loads,
supports, mesh sizes and penalties in one call.

The batch runs the same kernel bodies with the same tiles and deterministic reductions as a
solo call (as single-design calls on worker threads for handles that have not run yet, as one
merged step for the others), so every design's J and w must equal its solo result bit for bit.
One design is also checked against the oracle, and a NaN design must come back as NaN without
disturbing the others.
"""

import json
import os

import numpy as np
import pytest
import torch

from tests.cases import case_from_design, gpu_problem, oracle_problem, smoothed_density
from tests.conftest import GOLDEN
from topomax_amd import ProblemBatch

pytestmark = pytest.mark.gpu

# (design file, N, penalty): uneven sizes, loads (tractions, body force) and supports
DESIGNS = [("cantilever.json", 100, 3), ("bridge.json", 60, 4), ("triangle.json", 40, 3),
           ("short_cantilever.json", 30, 5), ("cantilever.json", 40, 1), ("bridge.json", 20, 2)]


def _case(name):
    with open(os.path.join(GOLDEN, "designs", name)) as fh:
        return case_from_design(json.load(fh))


def _build(precision, rtol):
    probs, rhos = [], []
    for k, (name, N, p) in enumerate(DESIGNS):
        case = _case(name)
        nx, ny = round(case["width"] * N), round(case["height"] * N)
        gp = gpu_problem(case, nx, ny, precision=precision, rtol=rtol)
        gp.set_penalization(p)
        probs.append(gp)
        rhos.append(torch.as_tensor(smoothed_density(nx, ny, seed=k, sigma=3.0, lo=0.05),
                                    device="cuda:0"))
    return probs, rhos


@pytest.mark.parametrize("precision,rtol", [("fp64", 1e-12), ("mixed", 1e-10)])
def test_batch_equals_solo_bitwise(precision, rtol):
    probs, rhos = _build(precision, rtol)
    solo = [gp.objective_and_gradient(r) for gp, r in zip(probs, rhos)]
    batch = ProblemBatch(probs, workers=4)
    for rep in range(2):  # the second call reuses every handle's graphs and operators
        J, w = batch.objective_and_gradient(rhos)
        for k, (Js, ws, _) in enumerate(solo):
            print(precision, rep, DESIGNS[k], J[k], Js, batch.last_stats[k], flush=True)
            assert J[k] == Js
            assert torch.equal(w[k], ws)


def test_batch_matches_oracle_and_isolates_nan():
    probs, rhos = _build("fp64", 1e-12)
    name, N, p = DESIGNS[2]
    case = _case(name)
    nx, ny = round(case["width"] * N), round(case["height"] * N)
    orc = oracle_problem(case, nx, ny)
    orc.set_penalization(p)
    Jo = orc.calculate_objective(rhos[2].cpu().numpy())
    wo = orc.calculate_objective_gradient()
    bad = rhos[1].clone()
    bad[11] = float("nan")
    rhos[1] = bad
    batch = ProblemBatch(probs, workers=3)
    J, w = batch.objective_and_gradient(rhos)
    print("oracle", Jo, "batch", J[2], flush=True)
    assert abs(J[2] - Jo) < 1e-8 * abs(Jo)
    assert np.linalg.norm(w[2].cpu().numpy() - wo) < 1e-8 * np.linalg.norm(wo)
    assert np.isnan(J[1])
    assert all(np.isfinite(J[k]) for k in range(len(J)) if k != 1)


def test_batch_rejects_duplicates_and_size_mismatch():
    probs, rhos = _build("fp64", 1e-12)
    with pytest.raises(ValueError, match="twice"):
        ProblemBatch([probs[0], probs[1], probs[0]])
    batch = ProblemBatch(probs[:2])
    with pytest.raises(ValueError):
        batch.objective_and_gradient([rhos[1], rhos[0]])


def test_batch_64_random_designs():
    """The batched workload: 64 seeded designs (N in {40..320}, aspect 1-3, p 1-5, 1-3 random
    tractions, optional body force, random fixed sides), mixed precision with autotuned tiles,
    in one batch call; every result bitwise equal to its solo call, and the two smallest
    designs against the oracle."""
    from tests.cases import cfg5_cases
    cases = cfg5_cases(64)
    probs, rhos = [], []
    for k, (case, N, p) in enumerate(cases):
        nx, ny = int(case["width"] * N), N
        gp = gpu_problem(case, nx, ny, precision="mixed", rtol=1e-10)
        gp.set_penalization(p)
        probs.append(gp)
        rhos.append(torch.as_tensor(smoothed_density(nx, ny, seed=k), device="cuda:0"))
    for gp in probs:
        gp.autotune()
    batch = ProblemBatch(probs, workers=4)
    J, w = batch.objective_and_gradient(rhos)  # (fresh handles: single-design calls)
    assert all(np.isfinite(J))
    J2, w2 = batch.objective_and_gradient(rhos)  # merged: every handle has run once
    print(batch.last_report, flush=True)
    assert batch.last_report["merged_designs"] == 64
    for k in range(64):
        assert J2[k] == J[k] and torch.equal(w2[k], w[k])
    for k in range(0, 64, 7):  # solo spot checks, bitwise
        Js, ws, _ = probs[k].objective_and_gradient(rhos[k])
        assert J[k] == Js and torch.equal(w[k], ws)
    small = sorted(range(64), key=lambda k: cases[k][1] * cases[k][0]["width"])[:2]
    for k in small:
        case, N, p = cases[k]
        nx, ny = int(case["width"] * N), N
        orc = oracle_problem(case, nx, ny)
        orc.set_penalization(p)
        Jo = orc.calculate_objective(rhos[k].cpu().numpy())
        print("design", k, case["fixed_sides"], N, p, J[k], Jo, flush=True)
        assert abs(J[k] - Jo) <= 1e-6 * abs(Jo)  # 0 when every load sits on a fixed side
