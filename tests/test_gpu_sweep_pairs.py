"""The paired-sweep launches of the finest elasticity level (two block-Jacobi sweeps per launch:
zero start + sweep, optionally with the residual and its restriction fused behind them, and
prolongation + sweep + sweep) against the launch-per-sweep V-cycle.

Both compute the same sweeps in the same order, only without storing the intermediate iterate,
so the preconditioner output agrees to round-off: 1e-12 relative in FP64 and 1e-5 in FP32.
Cases cover nested and non-nested finest levels, every fixed-side pattern, both sources of the
cell factors (nodal xi and per-cell weights), chunk heights from 1 row up, and sweep counts
where the pair is followed by further sweeps or is absent (one sweep)."""

import pytest
import torch

from tests.cases import CANTILEVER, gpu_problem, smoothed_density

pytestmark = pytest.mark.gpu

SHAPES = [(60, 20), (120, 40), (45, 15), (250, 90), (1, 200), (300, 100)]
FIXED = [["Left"], ["Bottom"], ["Left", "Right", "Top", "Bottom"], ["Right", "Top"]]


def rel(a, b):
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


def residual_vector(nx, ny, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    return torch.randn(2 * (nx + 1) * (ny + 1), generator=g, dtype=torch.float64).cuda()


@pytest.mark.parametrize("nx,ny", SHAPES)
@pytest.mark.parametrize("fixed", FIXED, ids=["-".join(f) for f in FIXED])
def test_pairs_equal_per_sweep(nx, ny, fixed):
    case = dict(CANTILEVER, fixed_sides=fixed)
    gp = gpu_problem(case, nx, ny)
    gp.set_penalization(3)
    xi = torch.as_tensor(smoothed_density(nx, ny, 1, sigma=2.0), device="cuda:0")
    r = residual_vector(nx, ny, 5)
    for cw in (1, 2):  # bit 0: FP32 sweeps read per-cell weights; bit 1: FP64 sweeps do
        gp.set_option("fine_cell_weights", cw)
        for prec, tol in (("fp64", 1e-12), ("fp32", 1e-5)):
            gp.set_option("sweep_pairs", 0)
            gp.set_option("sweep_pairs_post", 0)
            z_ref = gp.precondition(r, "elasticity", xi, precision=prec)
            for rows, rr in ((16, 12), (1, 0), (5, 2), (12, 3), (8, 64)):
                gp.set_option("sweep_pairs", rows)
                gp.set_option("sweep_pairs_post", rows)
                gp.set_option("sweep_pair_restrict", rr)
                z = gp.precondition(r, "elasticity", xi, precision=prec)
                print(nx, ny, fixed, cw, prec, rows, rr, rel(z, z_ref), flush=True)
                assert torch.isfinite(z).all()
                assert rel(z, z_ref) < tol


@pytest.mark.parametrize("nu_pre,nu_post", [(1, 1), (3, 4), (1, 2), (2, 1)])
def test_pairs_sweep_counts(nu_pre, nu_post):
    nx, ny = 250, 90
    case = dict(CANTILEVER, fixed_sides=["Left", "Bottom"])
    gp = gpu_problem(case, nx, ny)
    gp.set_penalization(3)
    gp.set_option("nu_pre", nu_pre)
    gp.set_option("nu_post", nu_post)
    xi = torch.as_tensor(smoothed_density(nx, ny, 4, sigma=2.0), device="cuda:0")
    r = residual_vector(nx, ny, 9)
    for prec, tol in (("fp64", 1e-12), ("fp32", 1e-5)):
        gp.set_option("sweep_pairs", 0)
        gp.set_option("sweep_pairs_post", 0)
        z_ref = gp.precondition(r, "elasticity", xi, precision=prec)
        gp.set_option("sweep_pairs", 16)
        gp.set_option("sweep_pairs_post", 16)
        z = gp.precondition(r, "elasticity", xi, precision=prec)
        print(nu_pre, nu_post, prec, rel(z, z_ref), flush=True)
        assert rel(z, z_ref) < tol


def test_pairs_repeated_calls_identical():
    """Deterministic: no atomics, so two V-cycles give the same bits."""
    nx, ny = 300, 100
    gp = gpu_problem(CANTILEVER, nx, ny)
    gp.set_penalization(3)
    xi = torch.as_tensor(smoothed_density(nx, ny, 2, sigma=2.0), device="cuda:0")
    r = residual_vector(nx, ny, 6)
    for prec in ("fp64", "fp32"):
        a = gp.precondition(r, "elasticity", xi, precision=prec)
        b = gp.precondition(r, "elasticity", xi, precision=prec)
        assert torch.equal(a, b)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pairs_solve_matches(precision):
    """Objective and gradient with and without paired sweeps (PCG to 1e-11, where r.z comes
    from the paired launch's partial sums): the solves agree far below the tolerance."""
    nx, ny = 240, 80
    gp = gpu_problem(CANTILEVER, nx, ny, precision=precision, rtol=1e-11)
    gp.set_penalization(3)
    rho = torch.as_tensor(smoothed_density(nx, ny, 3), device="cuda:0")
    gp.set_option("sweep_pairs", 0)
    gp.set_option("sweep_pairs_post", 0)
    J0, w0, _ = gp.objective_and_gradient(rho)
    s0 = dict(gp.last_stats)
    gp.set_option("sweep_pairs", 16)
    gp.set_option("sweep_pairs_post", 16)
    J1, w1, _ = gp.objective_and_gradient(rho)
    print(precision, J0, J1, s0, gp.last_stats, flush=True)
    assert abs(J1 - J0) < 1e-9 * abs(J0)
    assert rel(w1, w0) < 1e-8


@pytest.mark.parametrize("nx,ny", [(120, 40), (250, 90), (300, 100)])
@pytest.mark.parametrize("levels,nu", [(2, 0), (2, 2), (3, 2)])
def test_paired_coarse_levels(nx, ny, levels, nu):
    """Paired kernels on the coarser element levels (option pair_levels) against one launch
    per sweep on the same levels (pair_rows_coarse 0): the same sweeps in the same order."""
    gp = gpu_problem(CANTILEVER, nx, ny)
    gp.set_penalization(3)
    xi = torch.as_tensor(smoothed_density(nx, ny, 2, sigma=2.0), device="cuda:0")
    r = residual_vector(nx, ny, 6)
    gp.set_option("pair_levels", levels)
    gp.set_option("pair_nu_coarse", nu)
    for prec, tol in (("fp64", 1e-12), ("fp32", 1e-5)):
        gp.set_option("pair_rows_coarse", 0)
        z_ref = gp.precondition(r, "elasticity", xi, precision=prec)
        for rows in (1, 4, 8, 64):
            gp.set_option("pair_rows_coarse", rows)
            z = gp.precondition(r, "elasticity", xi, precision=prec)
            print(nx, ny, levels, nu, prec, rows, rel(z, z_ref), flush=True)
            assert torch.isfinite(z).all()
            assert rel(z, z_ref) < tol


def test_paired_coarse_levels_solve():
    """The solve with paired level-1 kernels reaches the band kernels' objective and
    gradient (both to rtol 1e-10)."""
    nx, ny = 300, 100
    gp = gpu_problem(CANTILEVER, nx, ny, rtol=1e-10)
    gp.set_penalization(3)
    rho = torch.as_tensor(smoothed_density(nx, ny, 4), device="cuda:0")
    J0, w0, _ = gp.objective_and_gradient(rho)
    gp.set_option("pair_levels", 2)
    J1, w1, _ = gp.objective_and_gradient(rho)
    print(J0, J1, flush=True)
    assert abs(J1 - J0) < 1e-9 * abs(J0)
    assert rel(w1, w0) < 1e-8
