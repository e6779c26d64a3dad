"""The merged batch step (tmx_batch_*, TMX_BATCH_MERGED): one kernel launch per PCG phase and
V-cycle stage covers every qualifying design of the batch, the designs' PCG iterations in
lockstep.

Each design goes through the same kernel bodies, with the same arguments and reductions, as
its single-design call, and a design that has converged sits out the remaining iterations with
zero step sizes, so J, w and the solve statistics must equal the single-design call's bit for
bit.  Designs that do not qualify (never run before, another precision scheme, the P2 element,
an option just changed) and designs whose merged solve stops without converging (a NaN design)
run as single-design calls instead, which the report records; the launches per step depend on
the hierarchies and iteration counts, not on the number of designs.
"""

import json
import os

import numpy as np
import pytest
import torch

from tests.cases import case_from_design, gpu_problem, smoothed_density
from tests.conftest import GOLDEN
from topomax_amd import ProblemBatch

pytestmark = pytest.mark.gpu

# (design file, N, penalty): uneven sizes, loads, supports, odd grids (bridge N=25)
DESIGNS = [("cantilever.json", 100, 3), ("bridge.json", 60, 4), ("triangle.json", 40, 3),
           ("short_cantilever.json", 30, 5), ("cantilever.json", 40, 1), ("bridge.json", 25, 2)]


def _case(name):
    with open(os.path.join(GOLDEN, "designs", name)) as fh:
        return case_from_design(json.load(fh))


def _problem(k, name, N, p, precision, rtol, degree=1):
    case = _case(name)
    nx, ny = round(case["width"] * N), round(case["height"] * N)
    gp = gpu_problem(case, nx, ny, precision=precision, rtol=rtol, degree=degree)
    gp.set_penalization(p)
    rho = torch.as_tensor(smoothed_density(nx, ny, seed=k, sigma=3.0, lo=0.05), device="cuda:0")
    return gp, rho


def _build(precision, rtol, designs=DESIGNS):
    probs, rhos = [], []
    for k, (name, N, p) in enumerate(designs):
        gp, rho = _problem(k, name, N, p, precision, rtol)
        probs.append(gp)
        rhos.append(rho)
    return probs, rhos


def _check_equal(batch, J, w, solo, label):
    for k, (Js, ws, sts) in enumerate(solo):
        print(label, k, J[k], Js, batch.last_stats[k], sts, flush=True)
        assert J[k] == Js
        assert torch.equal(w[k], ws)
        assert batch.last_stats[k]["iterations"] == sts["iterations"]
        assert batch.last_stats[k]["rel_residual"] == sts["rel_residual"]


@pytest.mark.parametrize("precision,rtol", [("fp64", 1e-12), ("mixed", 1e-10), ("mixed", 1e-8)])
def test_merged_equals_solo_bitwise(precision, rtol):
    probs, rhos = _build(precision, rtol)
    solo = []
    for gp, r in zip(probs, rhos):
        J, w, _ = gp.objective_and_gradient(r)
        solo.append((J, w, dict(gp.last_stats)))
    batch = ProblemBatch(probs, workers=4)
    for rep in range(2):
        J, w = batch.objective_and_gradient(rhos)
        rep_ = batch.last_report
        print(precision, rtol, rep, rep_, flush=True)
        assert rep_["merged_designs"] == len(probs) and rep_["solo_designs"] == 0
        assert rep_["launches"] > 0
        if rep == 1:
            assert rep_["record_ms"] == 0.0  # the recorded plan was reused
        _check_equal(batch, J, w, solo, f"{precision} {rep}")


def test_launches_do_not_grow_with_designs():
    """The same design replicated 2 and 8 times: the merged step issues the same launches and
    iterations; the results equal the single-design call's."""
    counts = {}
    name, N, p = DESIGNS[1]
    for copies in (2, 8):
        built = [_problem(0, name, N, p, "mixed", 1e-8) for _ in range(copies)]
        probs, rhos = [b[0] for b in built], [b[1] for b in built]
        Js, ws, _ = probs[0].objective_and_gradient(rhos[0])
        for gp, r in zip(probs[1:], rhos[1:]):
            gp.objective_and_gradient(r)  # (each handle runs once alone before merging)
        batch = ProblemBatch(probs)
        J, w = batch.objective_and_gradient(rhos)
        rep = batch.last_report
        print(copies, rep, flush=True)
        assert rep["merged_designs"] == copies
        counts[copies] = (rep["launches"], tuple(rep["iterations"]))
        for k in range(copies):
            assert J[k] == Js and torch.equal(w[k], ws)
    assert counts[2] == counts[8]


def test_nan_design_runs_alone():
    probs, rhos = _build("mixed", 1e-10)
    solo = []
    for gp, r in zip(probs, rhos):
        J, w, _ = gp.objective_and_gradient(r)
        solo.append((J, w, dict(gp.last_stats)))
    bad = rhos[2].clone()
    bad[7] = float("nan")
    batch = ProblemBatch(probs)
    J, w = batch.objective_and_gradient(rhos[:2] + [bad] + rhos[3:])
    rep = batch.last_report
    print(rep, J, flush=True)
    assert np.isnan(J[2])
    assert rep["solo_designs"] == 1 and rep["merged_designs"] == len(probs) - 1
    for k in range(len(probs)):
        if k != 2:
            assert J[k] == solo[k][0] and torch.equal(w[k], solo[k][1])
    # the NaN design's handle recovers: merged again, equal to its single-design result
    J, w = batch.objective_and_gradient(rhos)
    assert batch.last_report["merged_designs"] == len(probs)
    assert J[2] == solo[2][0] and torch.equal(w[2], solo[2][1])


def test_unqualified_designs_run_alone_alongside():
    """A handle that has not run yet, a P2 problem and a mixed-ir handle run as single-design
    calls next to the merged ones; a changed option sends a handle alone for one call."""
    probs, rhos = _build("mixed", 1e-10, DESIGNS[:3])
    for gp, r in zip(probs, rhos):
        gp.objective_and_gradient(r)
    fresh, rf = _problem(7, "cantilever.json", 50, 3, "mixed", 1e-10)
    p2, r2 = _problem(8, "bridge.json", 20, 3, "fp64", 1e-10, degree=2)
    ir, ri = _problem(9, "triangle.json", 30, 3, "mixed-ir", 1e-10)
    allp = probs + [fresh, p2, ir]
    allr = rhos + [rf, r2, ri]
    batch = ProblemBatch(allp, workers=3)
    J, w = batch.objective_and_gradient(allr, rtol=1e-10)
    print(batch.last_report, flush=True)
    assert batch.last_report["merged_designs"] == 3 and batch.last_report["solo_designs"] == 3
    J2, w2 = batch.objective_and_gradient(allr, rtol=1e-10)
    print(batch.last_report, flush=True)
    assert batch.last_report["merged_designs"] == 4  # the fresh handle has run once now
    for k in range(len(allp)):
        assert J2[k] == J[k] and torch.equal(w2[k], w[k])
        Js, ws, _ = allp[k].objective_and_gradient(allr[k], rtol=1e-10)
        assert J[k] == Js and torch.equal(w[k], ws)
    probs[1].set_option("fused", 1)  # (the default value: results unchanged)
    J3, _ = batch.objective_and_gradient(allr, rtol=1e-10)
    assert batch.last_report["merged_designs"] == 3
    assert J3 == J
    batch.objective_and_gradient(allr, rtol=1e-10)
    assert batch.last_report["merged_designs"] == 4
