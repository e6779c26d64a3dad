"""GPU tests of the multigrid preconditioner itself.

PCG converges to the right answer for any symmetric positive definite preconditioner, so the
solve parity tests cannot see a broken V-cycle.  These tests pin it directly:

* the fused kernels (residual + restriction, prolongation + smoothing, one-workgroup tail)
  give the same V-cycle as the unfused sequence, to round-off;
* the V-cycle is a symmetric positive definite operator (r.Ms = s.Mr, r.Mr > 0), as PCG needs;
* the FP32 V-cycle matches the FP64 one to single precision;
* the iteration count to a fixed tolerance does not grow with the mesh (multigrid optimality).
"""

import numpy as np
import pytest
import torch

from tests.cases import CANTILEVER, gpu_problem, smoothed_density

pytestmark = pytest.mark.gpu

# (nx, ny): even grids (nested levels then 9-point levels), an odd fine grid (9-point levels
# from the start) and a power-of-two grid (nested levels down to the dense coarsest)
SHAPES = [(60, 20), (120, 40), (45, 15), (64, 64)]


def rel(a, b):
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


def setup(nx, ny, seed=0):
    gp = gpu_problem(CANTILEVER, nx, ny)
    gp.set_penalization(3)
    gp.set_option("filter_dct", 0)  # these tests are of the filter's V-cycle
    xi = torch.as_tensor(smoothed_density(nx, ny, seed, sigma=2.0), device="cuda:0")
    g = torch.Generator(device="cpu").manual_seed(seed)
    r = torch.randn(2 * (nx + 1) * (ny + 1), generator=g, dtype=torch.float64).cuda()
    s = torch.randn(2 * (nx + 1) * (ny + 1), generator=g, dtype=torch.float64).cuda()
    rf = torch.randn((nx + 1) * (ny + 1), generator=g, dtype=torch.float64).cuda()
    return gp, xi, r, s, rf


@pytest.mark.parametrize("nx,ny", SHAPES)
def test_fused_equals_unfused(nx, ny):
    gp, xi, r, _, rf = setup(nx, ny)
    z_ref = gp.precondition(r, "elasticity", xi)
    f_ref = gp.precondition(rf, "filter")
    for fused, tail in [(0, 0), (1, 0), (0, 1)]:
        gp.set_option("fused", fused)
        gp.set_option("tail", tail)
        z = gp.precondition(r, "elasticity", xi)
        f = gp.precondition(rf, "filter")
        print(nx, ny, "levels", gp.levels, "fused", fused, "tail", tail, rel(z, z_ref),
              rel(f, f_ref), flush=True)
        assert rel(z, z_ref) < 1e-12
        assert rel(f, f_ref) < 1e-12


@pytest.mark.parametrize("nx,ny", SHAPES)
def test_vcycle_symmetric_positive(nx, ny):
    gp, xi, r, s, _ = setup(nx, ny, seed=1)
    mr = gp.precondition(r, "elasticity", xi)
    ms = gp.precondition(s, "elasticity", xi)
    a, b = float(torch.dot(s, mr)), float(torch.dot(r, ms))
    print(nx, ny, a, b, flush=True)
    assert abs(a - b) < 1e-10 * (abs(a) + abs(b))
    assert float(torch.dot(r, mr)) > 0 and float(torch.dot(s, ms)) > 0
    rf = torch.randn((nx + 1) * (ny + 1), dtype=torch.float64).cuda()
    sf = torch.randn((nx + 1) * (ny + 1), dtype=torch.float64).cuda()
    a = float(torch.dot(sf, gp.precondition(rf, "filter")))
    b = float(torch.dot(rf, gp.precondition(sf, "filter")))
    assert abs(a - b) < 1e-10 * (abs(a) + abs(b))


@pytest.mark.parametrize("nx,ny", SHAPES)
def test_fp32_vcycle_matches_fp64(nx, ny):
    gp, xi, r, _, rf = setup(nx, ny, seed=2)
    z64 = gp.precondition(r, "elasticity", xi, precision="fp64")
    z32 = gp.precondition(r, "elasticity", xi, precision="fp32")
    f64 = gp.precondition(rf, "filter", precision="fp64")
    f32 = gp.precondition(rf, "filter", precision="fp32")
    print(nx, ny, rel(z32, z64), rel(f32, f64), flush=True)
    assert rel(z32, z64) < 1e-4
    assert rel(f32, f64) < 1e-5


def test_iterations_mesh_independent():
    its = {}
    for N in (20, 40, 80):
        nx, ny = 3 * N, N
        gp = gpu_problem(CANTILEVER, nx, ny, rtol=1e-8)
        gp.set_penalization(3)
        xi = torch.as_tensor(smoothed_density(nx, ny, 0, sigma=5.0), device="cuda:0")
        gp.forward(xi)
        its[N] = gp.last_stats["iterations"]
    print(its, flush=True)
    assert max(its.values()) <= 30
    assert its[80] <= its[20] + 6
