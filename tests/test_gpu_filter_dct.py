"""The filter's exact fast preconditioner: z = A'^-1 r for A' = eps^2 K + hx hy (W_y (x) W_x) (the
DCT-I in y folded into even and odd modes, tridiagonal solves in x by the partition method).  Checked by applying A' to z
in FP64 (torch, written out from its stencil) and comparing with r; and at the solve level
against the multigrid-preconditioned filter."""

import math

import pytest
import torch

from tests.cases import CANTILEVER, gpu_problem, smoothed_density

pytestmark = pytest.mark.gpu


def a_prime(z, nx, ny, hx, hy, eps):
    """eps^2 [(hy/hx) W_y T_x + (hx/hy) T_y W_x] z + hx hy W_y W_x z on the (ny+1, nx+1) grid."""
    Z = z.reshape(ny + 1, nx + 1)
    wx = torch.ones(nx + 1, dtype=torch.float64, device=z.device)
    wy = torch.ones(ny + 1, dtype=torch.float64, device=z.device)
    wx[0] = wx[-1] = 0.5
    wy[0] = wy[-1] = 0.5

    def lap(v, dim):
        out = 2 * v
        if dim == 1:
            out = out.clone()
            if v.shape[1] > 1:
                out[:, 1:] -= v[:, :-1]
                out[:, :-1] -= v[:, 1:]
            out[:, 0] = v[:, 0] - (v[:, 1] if v.shape[1] > 1 else 0)
            out[:, -1] = v[:, -1] - (v[:, -2] if v.shape[1] > 1 else 0)
        else:
            out = out.clone()
            out[1:, :] -= v[:-1, :]
            out[:-1, :] -= v[1:, :]
            out[0, :] = v[0, :] - v[1, :]
            out[-1, :] = v[-1, :] - v[-2, :]
        return out

    Y = (eps * eps * (hy / hx) * wy[:, None] * lap(Z, 1)
         + eps * eps * (hx / hy) * wx[None, :] * lap(Z, 0)
         + hx * hy * wy[:, None] * wx[None, :] * Z)
    return Y.reshape(-1)


# relative residual of A' z = r: FP32 round-off of the transforms and the reduction, and the
# same in FP64.  The residual bound is the round-off times the condition of A' along y, about
# 4 (eps / hy)^2 (the y Laplacian against the mass), so FP32 gets at least 6e-8 times that
# (1e-4 up to ny = 1000, 2.5e-3 at ny = 2048 with eps = 0.05)
TOL = {"fp32": 1e-4, "fp64": 1e-11}


def tolerance(precision, ny):
    if precision == "fp64":
        return TOL["fp64"]
    kappa = 4.0 * (CANTILEVER["filter_radius"] * ny / CANTILEVER["height"]) ** 2
    return max(TOL["fp32"], 6e-8 * kappa)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("nx,ny", [(60, 20), (120, 40), (250, 90), (20, 10), (300, 100), (15, 8),
                                   (1, 200), (3071, 4), (2047, 6), (45, 15), (63, 21), (140, 70),
                                   (3, 2048), (96, 175)])
def test_dct_solves_lumped_operator(nx, ny, precision):
    """Grids of more than 128 nodes (smaller filter grids are solved densely, exactly).  x rows
    of 2 nodes (one row per thread block), 3072 (the longest: 12 rows in each of the 256
    blocks) and 2048 (8 rows per block), and rows whose last block is short.  The y transform
    is an FFT of 2 ny points: radix plans 4-4-5 (ny = 20) .. with radix 2, 3 (ny = 90, 15),
    5 and 7 (ny = 21, 70, 175), odd ny, the longest column (ny = 2048), and odd column counts
    (nx + 1 odd: a workgroup's second column is outside the grid)."""
    gp = gpu_problem(CANTILEVER, nx, ny)
    gp.set_penalization(3)
    gp.set_option("filter_dct", 1)
    g = torch.Generator(device="cpu").manual_seed(11)
    r = torch.randn((nx + 1) * (ny + 1), generator=g, dtype=torch.float64).cuda()
    z = gp.precondition(r, "filter", precision=precision)
    hx, hy = CANTILEVER["width"] / nx, CANTILEVER["height"] / ny
    back = a_prime(z, nx, ny, hx, hy, CANTILEVER["filter_radius"])
    err = float(torch.linalg.norm(back - r) / torch.linalg.norm(r))
    print(nx, ny, precision, err, flush=True)
    assert torch.isfinite(z).all()
    assert err < tolerance(precision, ny)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("nx,ny", [(44, 22), (3072, 2), (8, 2100)])
def test_unsupported_grid_keeps_the_v_cycle(precision, nx, ny):
    """2 ny with a prime factor above 7, x rows longer than 3072 nodes, or columns longer than
    2049 nodes: the preconditioner stays the V-cycle (the same output with the option on or
    off)."""
    gp = gpu_problem(CANTILEVER, nx, ny)
    gp.set_penalization(3)
    r = torch.randn((nx + 1) * (ny + 1), dtype=torch.float64).cuda()
    gp.set_option("filter_dct", 0)
    a = gp.precondition(r, "filter", precision=precision)
    gp.set_option("filter_dct", 1)
    b = gp.precondition(r, "filter", precision=precision)
    assert torch.equal(a, b)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_dct_solve_matches_multigrid(precision):
    nx, ny = 240, 80
    gp = gpu_problem(CANTILEVER, nx, ny, precision=precision, rtol=1e-11)
    gp.set_penalization(3)
    rho = torch.as_tensor(smoothed_density(nx, ny, 3), device="cuda:0")
    gp.set_option("filter_dct", 0)
    J0, w0, _ = gp.objective_and_gradient(rho)
    gp.set_option("filter_dct", 1)
    J1, w1, _ = gp.objective_and_gradient(rho)
    print(precision, J0, J1, flush=True)
    assert abs(J1 - J0) < 1e-9 * abs(J0)
    assert float(torch.linalg.norm(w1 - w0) / torch.linalg.norm(w0)) < 1e-8
