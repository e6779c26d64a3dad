"""PCG iterations replayed from captured hipGraphs (TMX_OPT_GRAPHS = 1) against the same
iterations launched one by one (the default): the same kernels with the same arguments, so
every replay must give the direct launches' results bit for bit, with the program's own
eager work and synchronisations between the replays.  The filter's DCT preconditioner (library
GEMMs) is not captured: with graphs on the filter keeps the V-cycle, so both runs here use it."""

import pytest
import torch

from tests.cases import CANTILEVER, gpu_problem, smoothed_density

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed-ir"])
def test_graph_replays_equal_direct_launches(precision):
    nx, ny = 240, 80
    gp = gpu_problem(CANTILEVER, nx, ny, precision=precision, rtol=1e-10)
    gp.set_penalization(3)
    gp.set_option("filter_dct", 0)
    rho = torch.as_tensor(smoothed_density(nx, ny, 5), device="cuda:0")
    gp.set_option("graphs", 0)
    J0, w0, _ = gp.objective_and_gradient(rho)
    gp.set_option("graphs", 1)
    for rep in range(3):
        J, w, _ = gp.objective_and_gradient(rho)
        torch.cuda.synchronize()
        print(precision, rep, J, J0, gp.last_stats, flush=True)
        assert J == J0
        assert torch.equal(w, w0)
        gp.precondition(torch.ones(2 * (nx + 1) * (ny + 1), dtype=torch.float64,
                                   device="cuda:0"), "elasticity", rho)  # eager work between
