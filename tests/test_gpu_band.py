"""The tiled multi-sweep V-cycle kernels (band_down / band_up) against the launch-per-sweep
V-cycle they replace on the middle levels.

Both compute the same sequence of damped block-Jacobi sweeps, residuals and transfers, only
grouped into fewer launches, so the preconditioner output must agree to round-off: 1e-12
relative in FP64 and 1e-5 in FP32 (where the element levels form A x in a different order).
The shapes cover nested P1 levels, 9-point levels with odd node counts, a direction that is
not coarsened (one cell wide), tile edges that do not divide the grid, and every side fixed.
"""

import pytest
import torch

from tests.cases import CANTILEVER, gpu_problem, smoothed_density

pytestmark = pytest.mark.gpu

SHAPES = [(60, 20), (120, 40), (45, 15), (64, 64), (250, 90), (1, 200), (300, 100)]
FIXED = [["Left"], ["Bottom"], ["Left", "Right", "Top", "Bottom"], ["Right", "Top"]]


def rel(a, b):
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


def vectors(nx, ny, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    r = torch.randn(2 * (nx + 1) * (ny + 1), generator=g, dtype=torch.float64).cuda()
    rf = torch.randn((nx + 1) * (ny + 1), generator=g, dtype=torch.float64).cuda()
    return r, rf


@pytest.mark.parametrize("nx,ny", SHAPES)
@pytest.mark.parametrize("fixed", FIXED, ids=["-".join(f) for f in FIXED])
def test_band_equals_per_sweep(nx, ny, fixed):
    case = dict(CANTILEVER, fixed_sides=fixed)
    gp = gpu_problem(case, nx, ny)
    gp.set_penalization(3)
    gp.set_option("filter_pipe", 0)  # the filter's finest level in the band kernels too
    gp.set_option("filter_dct", 0)  # and its FP32 preconditioner the V-cycle
    xi = torch.as_tensor(smoothed_density(nx, ny, 1, sigma=2.0), device="cuda:0")
    r, rf = vectors(nx, ny, 5)
    for prec, tol in (("fp64", 1e-12), ("fp32", 1e-5)):
        gp.set_option("band", 0)
        z_ref = gp.precondition(r, "elasticity", xi, precision=prec)
        f_ref = gp.precondition(rf, "filter", precision=prec)
        for band, fine in ((1, 1), (2, 1), (1, 0)):
            gp.set_option("band", band)
            gp.set_option("band_fine_filter", fine)
            z = gp.precondition(r, "elasticity", xi, precision=prec)
            f = gp.precondition(rf, "filter", precision=prec)
            print(nx, ny, fixed, prec, "levels", gp.levels, "band", band, fine, rel(z, z_ref),
                  rel(f, f_ref), flush=True)
            assert torch.isfinite(z).all() and torch.isfinite(f).all()
            assert rel(z, z_ref) < tol
            assert rel(f, f_ref) < tol


def test_band_repeated_calls_identical():
    """Deterministic: no atomics or races, so two V-cycles give the same bits."""
    nx, ny = 300, 100
    gp = gpu_problem(CANTILEVER, nx, ny)
    gp.set_penalization(3)
    gp.set_option("filter_pipe", 0)
    gp.set_option("filter_dct", 0)
    xi = torch.as_tensor(smoothed_density(nx, ny, 2, sigma=2.0), device="cuda:0")
    r, rf = vectors(nx, ny, 6)
    for prec in ("fp64", "fp32"):
        a = gp.precondition(r, "elasticity", xi, precision=prec)
        b = gp.precondition(r, "elasticity", xi, precision=prec)
        assert torch.equal(a, b)
        fa = gp.precondition(rf, "filter", precision=prec)
        fb = gp.precondition(rf, "filter", precision=prec)
        assert torch.equal(fa, fb)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_band_solve_matches(precision):
    """Objective and gradient with and without the band kernels (PCG to 1e-11): the two
    preconditioners differ by round-off only, so the solves agree far below the tolerance."""
    nx, ny = 240, 80
    gp = gpu_problem(CANTILEVER, nx, ny, precision=precision, rtol=1e-11)
    gp.set_penalization(3)
    gp.set_option("filter_pipe", 0)
    gp.set_option("filter_dct", 0)
    rho = torch.as_tensor(smoothed_density(nx, ny, 3), device="cuda:0")
    gp.set_option("band", 0)
    J0, w0, _ = gp.objective_and_gradient(rho)
    s0 = dict(gp.last_stats)
    gp.set_option("band", 1)
    J1, w1, _ = gp.objective_and_gradient(rho)
    print(precision, J0, J1, s0, gp.last_stats, flush=True)
    assert abs(J1 - J0) < 1e-9 * abs(J0)
    assert rel(w1, w0) < 1e-8


@pytest.mark.parametrize("nu_coarse,nu_filter,nu_fine_filter", [(1, 1, 1), (2, 2, 4), (5, 4, 2),
                                                                (8, 8, 8)])
def test_band_sweep_counts(nu_coarse, nu_filter, nu_fine_filter):
    """Halo widths follow the sweep counts (nu + 1 down, nu up): other counts than the
    defaults, up to the maximum, against the per-sweep launches."""
    nx, ny = 250, 90
    case = dict(CANTILEVER, fixed_sides=["Left", "Bottom"])
    gp = gpu_problem(case, nx, ny)
    gp.set_penalization(3)
    gp.set_option("filter_pipe", 0)
    gp.set_option("filter_dct", 0)
    gp.set_option("nu_coarse", nu_coarse)
    gp.set_option("nu_coarse_filter", nu_filter)
    gp.set_option("nu_fine_filter", nu_fine_filter)
    xi = torch.as_tensor(smoothed_density(nx, ny, 4, sigma=2.0), device="cuda:0")
    r, rf = vectors(nx, ny, 9)
    for prec, tol in (("fp64", 1e-12), ("fp32", 1e-5)):
        gp.set_option("band", 0)
        z_ref = gp.precondition(r, "elasticity", xi, precision=prec)
        f_ref = gp.precondition(rf, "filter", precision=prec)
        gp.set_option("band", 1)
        z = gp.precondition(r, "elasticity", xi, precision=prec)
        f = gp.precondition(rf, "filter", precision=prec)
        print(nu_coarse, nu_filter, nu_fine_filter, prec, rel(z, z_ref), rel(f, f_ref), flush=True)
        assert rel(z, z_ref) < tol
        assert rel(f, f_ref) < tol


@pytest.mark.parametrize("wide", [1, 2, 3])
@pytest.mark.parametrize("nx,ny", [(120, 40), (250, 90), (300, 100), (45, 15)])
def test_band_wide_equals_band(nx, ny, wide):
    """The 1024-thread one-node-per-thread band kernels (option band_wide 1) and the 512 and
    1024-thread kernels on the standard tiles (2, 3) on the elasticity's cell-weight levels
    against the 256-thread kernels: the same sweeps in the same order."""
    gp = gpu_problem(CANTILEVER, nx, ny)
    gp.set_penalization(3)
    xi = torch.as_tensor(smoothed_density(nx, ny, 3, sigma=2.0), device="cuda:0")
    g = torch.Generator(device="cpu").manual_seed(9)
    r = torch.randn(2 * (nx + 1) * (ny + 1), generator=g, dtype=torch.float64).cuda()
    for prec, tol in (("fp64", 1e-12), ("fp32", 1e-5)):
        gp.set_option("band_wide", 0)
        z0 = gp.precondition(r, "elasticity", xi, precision=prec)
        gp.set_option("band_wide", wide)
        z1 = gp.precondition(r, "elasticity", xi, precision=prec)
        err = float(torch.linalg.norm(z1 - z0) / torch.linalg.norm(z0))
        print(nx, ny, prec, err, flush=True)
        assert torch.isfinite(z1).all()
        assert err < tol
