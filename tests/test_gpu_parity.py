"""GPU parity: every libtmx.so entry point against the CPU oracle on the same inputs.

Tolerances: FP64 kernels agree with the oracle to round-off (1e-12 relative or tighter);
FP32 operator applications to 1e-5 relative; iterative solves are run to 1e-12 relative
residual and must match the oracle's direct (SuperLU) solve to 1e-8 relative displacement
error for FP64 PCG (with an FP64 or an FP32 V-cycle) and 1e-6 for FP32 inner solves inside
FP64 iterative refinement.
"""

import json

import numpy as np
import pytest
import torch

from tests.cases import (CANTILEVER, case_from_design, gpu_problem, oracle_problem,
                         smoothed_density, uniform_density)
from tests.conftest import design_path

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cuda(x):
    return torch.as_tensor(np.asarray(x), dtype=torch.float64, device="cuda:0")


def host(t):
    return t.detach().cpu().numpy()


# ------------------------------------------------------------------ load vector
@pytest.mark.parametrize("name,n", [("triangle", 10), ("cantilever", 20), ("short_cantilever", 10),
                                    ("bridge", 4)])
def test_load_vector_matches_oracle(name, n):
    design = json.load(open(design_path(name)))
    case = case_from_design(design)
    from oracle.fem import StructuredMesh as OM
    m = OM.from_design(case["width"], case["height"], n)
    orc = oracle_problem(case, m.nx, m.ny)
    gp = gpu_problem(case, m.nx, m.ny)
    F = host(gp.load_vector())
    print(name, "load nnz", np.count_nonzero(orc.load), flush=True)
    assert np.count_nonzero(orc.load) > 0
    assert rel(F, orc.load) < 1e-13
    # same support (membership of the body force / traction points is exact); the oracle's
    # quadrature leaves ~1e-16 relative residue where the exact weight is 0
    cut = 1e-12 * np.abs(orc.load).max()
    assert np.array_equal(np.abs(F) > cut, np.abs(orc.load) > cut)


# ------------------------------------------------------------------ operators
@pytest.mark.parametrize("p", [1, 2, 3, 5])
def test_matvec_fp64_and_fp32(p):
    nx, ny = 60, 20
    orc = oracle_problem(CANTILEVER, nx, ny)
    gp = gpu_problem(CANTILEVER, nx, ny)
    orc.set_penalization(p)
    gp.set_penalization(p)
    xi = smoothed_density(nx, ny, seed=p, sigma=2.0)
    rng = np.random.default_rng(100 + p)
    u = rng.standard_normal(2 * (nx + 1) * (ny + 1))
    y_ref = orc.masked_operator(xi) @ u
    y64 = host(gp.matvec(cuda(xi), cuda(u), "fp64"))
    y32 = host(gp.matvec(cuda(xi), cuda(u), "fp32"))
    print("p", p, "fp64", rel(y64, y_ref), "fp32", rel(y32, y_ref), flush=True)
    assert rel(y64, y_ref) < 1e-13
    assert rel(y32, y_ref) < 1e-5


def test_mass_and_integrate():
    nx, ny = 60, 20
    orc = oracle_problem(CANTILEVER, nx, ny)
    gp = gpu_problem(CANTILEVER, nx, ny)
    x = uniform_density((nx + 1) * (ny + 1), seed=3)
    assert rel(host(gp.filter.mass(cuda(x))), orc.mass @ x) < 1e-14
    assert abs(gp.integrate(cuda(x)) - orc.integrate(x)) < 1e-13 * abs(orc.integrate(x))


@pytest.mark.parametrize("eps", [0.0, 0.05])
def test_filter_apply_matches_oracle(eps):
    nx, ny = 60, 20
    case = dict(CANTILEVER, filter_radius=eps)
    orc = oracle_problem(case, nx, ny)
    gp = gpu_problem(case, nx, ny)
    rho = uniform_density((nx + 1) * (ny + 1), seed=7)
    xi = host(gp.filter.apply(cuda(rho)))
    ref = orc.filter_apply(rho)
    print("eps", eps, rel(xi, ref), flush=True)
    assert rel(xi, ref) < 1e-10
    if eps == 0.0:
        assert rel(xi, rho) < 1e-10


# ------------------------------------------------------------------ solves
@pytest.mark.parametrize("seed", [0, 1])
def test_cfg1_forward_objective_gradient_fp64(seed):
    """cfg1: 3x1 cantilever, N=20, E=1, nu=0.3, p=3, rho = 0.5 + 0.1 U(-1,1)."""
    nx, ny = 60, 20
    orc = oracle_problem(CANTILEVER, nx, ny)
    gp = gpu_problem(CANTILEVER, nx, ny, rtol=1e-13)
    orc.set_penalization(3)
    gp.set_penalization(3)
    rho = uniform_density((nx + 1) * (ny + 1), seed)
    J_ref = orc.calculate_objective(rho)
    w_ref = orc.calculate_objective_gradient()
    J = gp.calculate_objective(cuda(rho))
    w = host(gp.calculate_objective_gradient())
    u_err = rel(host(gp.u), orc.u)
    print("seed", seed, "J", J, J_ref, "u err", u_err, "w err", rel(w, w_ref), gp.last_stats,
          flush=True)
    assert u_err < 1e-8
    assert abs(J - J_ref) < 1e-9 * abs(J_ref)
    assert rel(w, w_ref) < 1e-8


@pytest.mark.parametrize("precision,tol", [("mixed", 1e-8), ("mixed-ir", 1e-6), ("mixed-ru", 1e-8)])
def test_cfg1_mixed_precision(precision, tol):
    nx, ny = 60, 20
    orc = oracle_problem(CANTILEVER, nx, ny)
    gp = gpu_problem(CANTILEVER, nx, ny, precision=precision, rtol=1e-12)
    orc.set_penalization(3)
    gp.set_penalization(3)
    rho = uniform_density((nx + 1) * (ny + 1), 0)
    J_ref = orc.calculate_objective(rho)
    J, w, mw = gp.objective_and_gradient(cuda(rho))
    w_ref = orc.calculate_objective_gradient()
    print(precision, "J", J, J_ref, "u err", rel(host(gp.u), orc.u), gp.last_stats, flush=True)
    assert rel(host(gp.u), orc.u) < tol
    assert abs(J - J_ref) < tol * abs(J_ref)
    assert rel(host(w), w_ref) < tol
    assert rel(host(mw), orc.mass @ w_ref) < tol


@pytest.mark.parametrize("p", [1, 3, 5])
def test_smoothed_density_high_contrast(p):
    """Smoothed random density in [1e-3, 1] (SIMP contrast down to m = 1e-6)."""
    nx, ny = 120, 40
    orc = oracle_problem(CANTILEVER, nx, ny)
    gp = gpu_problem(CANTILEVER, nx, ny, rtol=1e-13)
    orc.set_penalization(p)
    gp.set_penalization(p)
    xi = smoothed_density(nx, ny, seed=11, sigma=3.0)
    u_ref = orc.forward(xi)
    u = host(gp.forward(cuda(xi)))
    print("p", p, "u err", rel(u, u_ref), gp.last_stats, flush=True)
    assert rel(u, u_ref) < 1e-8
    um = host(gp.forward(cuda(xi), precision="mixed", rtol=1e-13))
    print("p", p, "mixed u err", rel(um, u_ref), gp.last_stats, flush=True)
    assert rel(um, u_ref) < 1e-8
    um = host(gp.forward(cuda(xi), precision="mixed-ir", rtol=1e-10))
    print("p", p, "mixed-ir u err", rel(um, u_ref), gp.last_stats, flush=True)
    assert rel(um, u_ref) < 1e-6
    um = host(gp.forward(cuda(xi), precision="mixed-ru", rtol=1e-12))
    print("p", p, "mixed-ru u err", rel(um, u_ref), gp.last_stats, flush=True)
    assert rel(um, u_ref) < 1e-8


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed-ir", "mixed-ru"])
def test_indefinite_stiffness_negative_filtered_density(precision):
    """A filtered design that dips below zero (the P1 Helmholtz filter has no maximum
    principle; EMD designs reach xi ~ -0.045) makes odd-penalty triangle factors negative and
    K(xi) indefinite.  The reference's LU solves it; here PCG breaks down and the solve falls
    back to MINRES preconditioned by the V-cycle of the factors floored at m."""
    nx, ny = 60, 20
    orc = oracle_problem(CANTILEVER, nx, ny)
    gp = gpu_problem(CANTILEVER, nx, ny, precision=precision, rtol=1e-12)
    orc.set_penalization(3)
    gp.set_penalization(3)
    # a disc of xi = -0.1 (181 triangles with factor -1e-3) in a [0.3, 1] design: K(xi) has
    # eigenvalues near -0.01 (checked with the oracle's assembled matrix)
    X, Y = np.meshgrid(np.linspace(0, 3, nx + 1), np.linspace(0, 1, ny + 1))
    hole = ((X - 2.0) ** 2 + (Y - 0.5) ** 2 < 0.3 ** 2).ravel()
    xi = np.where(hole, -0.1, 0.3 + 0.7 * smoothed_density(nx, ny, seed=4, sigma=2.0, lo=0.0))
    assert np.min(orc.element_weights(xi)) < 0
    u_ref = orc.forward(xi)
    u = host(gp.forward(cuda(xi)))
    st = gp.last_stats
    print(precision, "u err", rel(u, u_ref), st, flush=True)
    assert st["outer_iterations"] >= 2  # the MINRES fallback ran
    assert rel(u, u_ref) < 1e-8
    assert st["rel_residual"] < 1e-11


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed-ru"])
def test_indefinite_design_leaves_smoother_unchanged(precision):
    """The breakdown of an indefinite design retries its own solve with Jacobi smoothing and
    then MINRES; the next positive-definite design on the handle solves exactly as before it
    (same Chebyshev smoother, same iteration count, bitwise equal displacement)."""
    nx, ny = 60, 20
    gp = gpu_problem(CANTILEVER, nx, ny, precision=precision, rtol=1e-12)
    gp.set_penalization(3)
    X, Y = np.meshgrid(np.linspace(0, 3, nx + 1), np.linspace(0, 1, ny + 1))
    hole = ((X - 2.0) ** 2 + (Y - 0.5) ** 2 < 0.3 ** 2).ravel()
    good = 0.3 + 0.7 * smoothed_density(nx, ny, seed=4, sigma=2.0, lo=0.0)
    u_a = gp.forward(cuda(good)).clone()
    it_a = gp.last_stats["iterations"]
    gp.forward(cuda(np.where(hole, -0.1, good)))
    assert gp.last_stats["outer_iterations"] >= 2  # the fallback ran
    u_b = gp.forward(cuda(good))
    print(precision, it_a, gp.last_stats, flush=True)
    assert gp.last_stats["iterations"] == it_a and gp.last_stats["outer_iterations"] == 1
    assert torch.equal(u_a, u_b)
