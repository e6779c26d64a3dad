"""Differentiation: the hand-written adjoint behind torch.autograd against central finite
differences and the oracle gradient; plus the reference's error semantics on the GPU path.

dJ/drho = M w (w = filtered sensitivity, the reference's calculate_objective_gradient), so the
autograd gradient is checked along 32 seeded unit directions d against
(J(rho + h d) - J(rho - h d)) / 2h with h = 1e-6.  FD noise is ~ |J| * (solve tolerance) / h,
so the bound is relative to |grad|: 1e-6 for FP64 PCG (solves to 1e-13, h = 1e-6), with an
FP64 or an FP32 V-cycle.  FP32 inner solves inside FP64 iterative refinement stop at the FP64
residual floor (~1e-11 here), so that check uses rtol 1e-10, h = 1e-4 and a 1e-4 bound.
"""

import numpy as np
import pytest
import torch

from tests.cases import CANTILEVER, gpu_problem, oracle_problem, smoothed_density, uniform_density
from topomax_amd import compliance
from topomax_amd._lib import NotConvergedError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,rtol,h,tol", [("fp64", 1e-13, 1e-6, 1e-6),
                                                  ("mixed", 1e-13, 1e-6, 1e-6),
                                                  ("mixed-ir", 1e-10, 1e-4, 1e-4),
                                                  ("mixed-ru", 1e-10, 1e-4, 1e-4)])
def test_autograd_matches_central_differences_n400(precision, rtol, h, tol):
    N = 400
    nx, ny = 3 * N, N
    gp = gpu_problem(CANTILEVER, nx, ny, precision=precision, rtol=rtol)
    gp.set_option("filter_rtol", 1e-13)
    gp.set_penalization(3)
    rho = torch.as_tensor(smoothed_density(nx, ny, seed=3, sigma=5.0, lo=0.2), device="cuda:0")
    rho.requires_grad_(True)
    J = compliance(rho, gp)
    J.backward()
    g = rho.grad.detach()
    gnorm = float(torch.linalg.norm(g))
    gen = torch.Generator(device="cpu").manual_seed(1234)
    worst = 0.0
    with torch.no_grad():
        for k in range(32):
            d = torch.randn(rho.numel(), generator=gen, dtype=torch.float64).cuda()
            d /= torch.linalg.norm(d)
            jp, _, _ = gp.objective_and_gradient(rho + h * d)
            jm, _, _ = gp.objective_and_gradient(rho - h * d)
            fd = (jp - jm) / (2 * h)
            worst = max(worst, abs(fd - float(g @ d)) / gnorm)
    print(precision, "J", float(J), "|g|", gnorm, "worst FD error / |g|", worst, flush=True)
    assert worst < tol


def test_gradient_matches_oracle_and_second_call_reuses_state():
    nx, ny = 60, 20
    orc = oracle_problem(CANTILEVER, nx, ny)
    gp = gpu_problem(CANTILEVER, nx, ny, rtol=1e-13)
    orc.set_penalization(3)
    gp.set_penalization(3)
    rho = uniform_density((nx + 1) * (ny + 1), seed=5)
    orc.calculate_objective(rho)
    mw_ref = orc.mass @ orc.calculate_objective_gradient()
    r = torch.as_tensor(rho, device="cuda:0").requires_grad_(True)
    compliance(r, gp).backward()
    err = np.linalg.norm(r.grad.cpu().numpy() - mw_ref) / np.linalg.norm(mw_ref)
    print("autograd vs oracle M w", err, flush=True)
    assert err < 1e-8


def test_gradient_before_objective_raises():
    gp = gpu_problem(CANTILEVER, 12, 4)
    gp.set_penalization(3)
    with pytest.raises(ValueError, match="calculate_objective"):
        gp.calculate_objective_gradient()


def test_penalty_required_and_integer():
    gp = gpu_problem(CANTILEVER, 12, 4)
    rho = torch.full(((12 + 1) * (4 + 1),), 0.5, dtype=torch.float64, device="cuda:0")
    with pytest.raises(ValueError, match="penalization"):
        gp.calculate_objective(rho)
    with pytest.raises(ValueError):
        gp.set_penalization(2.5)


def test_nan_density_gives_nan_objective():
    """A non-finite design propagates to a NaN objective (the reference's direct solve does
    the same), so the optimiser's NaN exit rule (src/solver.py:268-271) still fires."""
    gp = gpu_problem(CANTILEVER, 30, 10)
    gp.set_penalization(3)
    rho = torch.full((31 * 11,), 0.5, dtype=torch.float64, device="cuda:0")
    rho[7] = float("nan")
    J = gp.calculate_objective(rho)
    assert np.isnan(J)
    J2, w, _ = gp.objective_and_gradient(rho)
    assert np.isnan(J2)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_handle_recovers_after_nan_design(precision):
    """After a non-finite design, a finite design on the same handle gives exactly what a
    fresh handle gives: no NaN survives in the solvers' buffers (even ny, so the filter solves
    run the DCT-I preconditioner, whose odd half carries a zero-padding row)."""
    nx, ny = 60, 20
    rho_ok = torch.as_tensor(smoothed_density(nx, ny, seed=5, sigma=2.0), device="cuda:0")
    bad = rho_ok.clone()
    bad[7] = float("nan")
    used = gpu_problem(CANTILEVER, nx, ny, precision=precision)
    used.set_penalization(3)
    J_bad, _, _ = used.objective_and_gradient(bad)
    assert np.isnan(J_bad)
    J1, w1, _ = used.objective_and_gradient(rho_ok)
    fresh = gpu_problem(CANTILEVER, nx, ny, precision=precision)
    fresh.set_penalization(3)
    J2, w2, _ = fresh.objective_and_gradient(rho_ok)
    print(precision, J1, J2, flush=True)
    assert np.isfinite(J1) and J1 == J2
    assert torch.equal(w1, w2)


def test_iteration_limit_reports_not_converged():
    gp = gpu_problem(CANTILEVER, 60, 20)
    gp.set_penalization(3)
    xi = torch.full((61 * 21,), 0.5, dtype=torch.float64, device="cuda:0")
    from topomax_amd.elasticity_problem import _ptr, solve_opts
    import ctypes
    from topomax_amd import _lib
    h = gp._h
    u = torch.empty(h.n_dofs, dtype=torch.float64, device="cuda:0")
    st = _lib.SolveStats()
    o = solve_opts(rtol=1e-14, max_iter=1)
    h.stream()
    rc = h.lib.tmx_elasticity_solve(h.ptr, _ptr(xi), _ptr(u), ctypes.byref(o), ctypes.byref(st))
    assert rc == _lib.TMX_ERR_NOT_CONVERGED and st.iterations == 1
    with pytest.raises(NotConvergedError):
        _lib.check(rc)
    # the handle stays usable afterwards
    u2 = gp.forward(xi)
    assert torch.isfinite(u2).all()
