"""torch.autograd wrapper of the compliance objective.

``compliance(rho, problem)`` returns J(rho) = F . u(F_filter(rho)) as a 0-d tensor whose
backward is the hand-written adjoint: the element sensitivity
-p rho^(p-1) (1-m) sigma0(u):eps(u) tested against the CG1 basis, chained through the
symmetric Helmholtz filter's adjoint solve (w = A^{-1} b) and the mass matrix
(dJ/drho = M w, since xi = A^{-1} M rho with A, M symmetric and compliance self-adjoint).
Reference: ``FEM_src/elasisity_problem.py:134-166``.
"""

from __future__ import annotations

import torch


class _Compliance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, problem, rtol, precision):
        J, w, mw = problem.objective_and_gradient(rho.detach(), rtol=rtol, precision=precision)
        ctx.save_for_backward(mw)
        return torch.tensor(J, dtype=torch.float64, device=rho.device)

    @staticmethod
    def backward(ctx, grad_out):
        (mw,) = ctx.saved_tensors
        return grad_out * mw, None, None, None


def compliance(rho: torch.Tensor, problem, *, rtol=None, precision=None) -> torch.Tensor:
    return _Compliance.apply(rho, problem, rtol, precision)
