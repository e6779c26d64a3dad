"""The GPU EMD loop (topomax_amd.emd) against the oracle's restatement of the reference loop
(oracle.emd.run_emd, itself pinned by the reference's triangle golden at P2) on the same P1
problem: the objective history, iteration count, exit condition and final design must agree.

Both loops solve to 1e-12 (GPU PCG vs SuperLU), so the histories agree to about 1e-10; the
bound is 1e-8 relative per objective and 1e-8 in the design.  The output-file test checks the
reference's file layout (src/solver.py:304-349, FEM_src/utils.py:47-69) without unpickling
the dataclass records (pickletools), and loads only the design dict this code wrote.
"""

import json
import os
import pickle
import pickletools

import numpy as np
import pytest

from tests.cases import case_from_design, oracle_problem
from tests.conftest import design_path
from topomax_amd.emd import EMDSolver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,N", [("triangle", 10), ("cantilever", 10)])
def test_emd_matches_oracle_loop(name, N):
    from oracle.emd import run_emd
    design = json.load(open(design_path(name)))
    dom = design["Elasticity"]["domain_parameters"]
    solver = EMDSolver(N, design_path(name), rtol=1e-12, filter_rtol=1e-13)
    res = solver.solve()
    case = case_from_design(design)
    orc = oracle_problem(case, solver.mesh.nx, solver.mesh.ny, degree=1)
    ref = run_emd(orc, width=dom["width"], height=dom["height"],
                  volume_fraction=dom["volume_fraction"], step_size=dom["fem_step_size"],
                  penalties=dom["penalties"])
    assert len(res) == len(ref)
    for r, o in zip(res, ref):
        print(name, r.penalty, r.iteration, r.exit_condition, r.objectives[-1],
              o["iteration"], o["exit_condition"], o["objectives"][-1], flush=True)
        assert r.iteration == o["iteration"]
        assert r.exit_condition == o["exit_condition"]
        a, b = np.array(r.objectives), np.array(o["objectives"])
        assert a.shape == b.shape
        assert np.abs(a - b).max() < 1e-8 * np.abs(b).max()
    rho = solver.rho.cpu().numpy()
    assert np.abs(rho - ref[-1]["rho"]).max() < 1e-8
    # the volume constraint holds on the device result
    assert abs(solver.integrate(solver.rho) - solver.volume) < 1e-10


def test_emd_outputs_reference_layout(tmp_path):
    solver = EMDSolver(10, design_path("triangle"), data_path=str(tmp_path), max_iterations=3)
    res = solver.solve()
    folder = tmp_path / "GPU" / "triangle" / "data"
    names = sorted(os.listdir(folder))
    print(names, flush=True)
    assert "N=10_p=3.0_result.dat" in names
    assert "N=10_p=3.0_k=0.dat" in names and "N=10_p=3.0_k=0_rho.dat" in names
    globals_seen = [arg for op, arg, _ in pickletools.genops(open(folder / "N=10_p=3.0_result.dat", "rb").read())
                    if op.name in ("SHORT_BINUNICODE", "BINUNICODE", "GLOBAL")]
    assert "src.utils" in globals_seen and "SolverResult" in globals_seen
    assert "exit_condition" in globals_seen and res[0].exit_condition in globals_seen
    it = [arg for op, arg, _ in pickletools.genops(open(folder / "N=10_p=3.0_k=0.dat", "rb").read())
          if op.name in ("SHORT_BINUNICODE", "BINUNICODE", "GLOBAL")]
    assert "IterationData" in it and "N=10_p=3.0_k=0_rho.dat" in it
    with open(folder / "N=10_p=3.0_k=0_rho.dat", "rb") as fh:  # written by this library
        data = pickle.load(fh)
    assert data["N"] == 10 and data["problem"] == "design" and data["order"] == "vertex"
    assert np.allclose(data["vector"], 0.5)
