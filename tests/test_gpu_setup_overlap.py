"""The elasticity solve's coarse-operator build on the hierarchy's side stream
(TMX_OPT_SETUP_OVERLAP) against the same build in order on the solve's stream.

Both run the same kernels on the same inputs; only the order between the build and the first
V-cycle's finer levels changes, and every reader of the coarse operators waits for the build.
So displacement, compliance, filtered design and gradient must be bitwise equal, also when
designs change between calls on one handle, when a V-cycle alone (tmx_precondition) follows a
solve, and for a right-hand side of zero (no iteration joins the build inside the solve).
"""

import numpy as np
import pytest
import torch

from tests.cases import CANTILEVER, gpu_problem, smoothed_density

pytestmark = pytest.mark.gpu


def designs(nx, ny, count):
    return [torch.as_tensor(0.2 + 0.8 * smoothed_density(nx, ny, seed, sigma=2.0, lo=0.0),
                            device="cuda:0") for seed in range(count)]


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed-ru"])
@pytest.mark.parametrize("nx,ny", [(120, 40), (250, 90)])
def test_overlapped_setup_is_bitwise_equal(precision, nx, ny):
    runs = {}
    for overlap in (0, 1):
        gp = gpu_problem(CANTILEVER, nx, ny, precision=precision, rtol=1e-10)
        gp.set_penalization(3)
        gp.set_option("setup_overlap", overlap)
        out = []
        for rho in designs(nx, ny, 3):
            J, w, mw = gp.objective_and_gradient(rho)
            out.append((J, gp.filtered_rho.clone(), gp.u.clone(), w.clone(), mw.clone(),
                        gp.last_stats["iterations"]))
            u = gp.forward(rho).clone()
            out.append((u, gp.last_stats["iterations"]))
            r = torch.linspace(-1.0, 1.0, 2 * (nx + 1) * (ny + 1), dtype=torch.float64,
                               device="cuda:0")
            out.append((gp.precondition(r, "elasticity", gp.filtered_rho, precision="fp32"),))
        torch.cuda.synchronize()
        runs[overlap] = out
        print(precision, nx, ny, "overlap", overlap, "levels", gp.levels,
              [o[-1] for o in out if len(o) > 1 and isinstance(o[-1], int)], flush=True)
    for a, b in zip(runs[0], runs[1]):
        for x, y in zip(a, b):
            if isinstance(x, torch.Tensor):
                assert torch.equal(x, y)
            else:
                assert x == y


def test_overlapped_setup_zero_load_then_solve():
    """A zero load converges before any V-cycle, so the build is joined only when the solve
    returns; a second handle with the real load and the same design, and the zero-load handle
    itself afterwards, must solve exactly as in order."""
    nx, ny = 120, 40
    rho = designs(nx, ny, 1)[0]
    zero_case = dict(CANTILEVER, tractions=[dict(CANTILEVER["tractions"][0], value=(0.0, 0.0))])
    res = {}
    for overlap in (0, 1):
        gz = gpu_problem(zero_case, nx, ny, precision="mixed", rtol=1e-10)
        gz.set_penalization(3)
        gz.set_option("setup_overlap", overlap)
        uz = gz.forward(rho).clone()
        assert torch.count_nonzero(uz) == 0 and gz.last_stats["iterations"] == 0
        r = torch.linspace(-1.0, 1.0, 2 * (nx + 1) * (ny + 1), dtype=torch.float64,
                           device="cuda:0")
        z = gz.precondition(r, "elasticity", rho, precision="fp32").clone()
        gp = gpu_problem(CANTILEVER, nx, ny, precision="mixed", rtol=1e-10)
        gp.set_penalization(3)
        gp.set_option("setup_overlap", overlap)
        u = gp.forward(rho).clone()
        res[overlap] = (uz, z, u, gp.last_stats["iterations"])
    for a, b in zip(res[0], res[1]):
        assert torch.equal(a, b) if isinstance(a, torch.Tensor) else a == b
