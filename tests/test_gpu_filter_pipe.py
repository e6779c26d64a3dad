"""The filter's finest level as two register column pipelines (nu sweeps from zero + residual +
restriction; prolongation + nu sweeps + b.y) against the band kernels and against the
launch-per-sweep V-cycle.  The same sweeps in the same order, so the preconditioner agrees to
round-off: 1e-12 relative in FP64, 1e-5 in FP32.  Odd grids (no nested level below the finest)
keep the band kernels, which the comparison then covers trivially."""

import pytest
import torch

from tests.cases import CANTILEVER, gpu_problem, smoothed_density

pytestmark = pytest.mark.gpu

SHAPES = [(60, 20), (120, 40), (45, 15), (250, 90), (2, 2), (300, 100), (118, 6)]


def rel(a, b):
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


def rhs(nx, ny, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    return torch.randn((nx + 1) * (ny + 1), generator=g, dtype=torch.float64).cuda()


@pytest.mark.parametrize("nx,ny", SHAPES)
@pytest.mark.parametrize("nu", [1, 2, 3, 4])
def test_pipe_equals_band_and_per_sweep(nx, ny, nu):
    gp = gpu_problem(CANTILEVER, nx, ny)
    gp.set_penalization(3)
    gp.set_option("filter_dct", 0)  # the FP32 filter preconditioner as a V-cycle
    gp.set_option("nu_fine_filter", nu)
    r = rhs(nx, ny, 7)
    for prec, tol in (("fp64", 1e-12), ("fp32", 1e-5)):
        gp.set_option("filter_pipe", 0)
        gp.set_option("band_fine_filter", 0)
        f_sweep = gp.precondition(r, "filter", precision=prec)
        gp.set_option("band_fine_filter", 1)
        f_band = gp.precondition(r, "filter", precision=prec)
        for rows in (16, 1, 3, 64):
            gp.set_option("filter_pipe", rows)
            f = gp.precondition(r, "filter", precision=prec)
            print(nx, ny, nu, prec, rows, rel(f, f_band), rel(f, f_sweep), flush=True)
            assert torch.isfinite(f).all()
            assert rel(f, f_band) < tol
            assert rel(f, f_sweep) < tol


def test_pipe_repeated_calls_identical():
    nx, ny = 300, 100
    gp = gpu_problem(CANTILEVER, nx, ny)
    gp.set_penalization(3)
    gp.set_option("filter_dct", 0)
    r = rhs(nx, ny, 8)
    for prec in ("fp64", "fp32"):
        a = gp.precondition(r, "filter", precision=prec)
        b = gp.precondition(r, "filter", precision=prec)
        assert torch.equal(a, b)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pipe_solve_matches(precision):
    """Objective and gradient with the filter's finest level in pipelines or band kernels
    (solves to 1e-11; the filter's r.z partials come from the up pipeline)."""
    nx, ny = 240, 80
    gp = gpu_problem(CANTILEVER, nx, ny, precision=precision, rtol=1e-11)
    gp.set_penalization(3)
    gp.set_option("filter_dct", 0)
    rho = torch.as_tensor(smoothed_density(nx, ny, 3), device="cuda:0")
    gp.set_option("filter_pipe", 0)
    J0, w0, _ = gp.objective_and_gradient(rho)
    gp.set_option("filter_pipe", 16)
    J1, w1, _ = gp.objective_and_gradient(rho)
    print(precision, J0, J1, flush=True)
    assert abs(J1 - J0) < 1e-9 * abs(J0)
    assert rel(w1, w0) < 1e-8
