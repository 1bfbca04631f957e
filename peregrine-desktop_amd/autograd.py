"""dL/dtheta of the value table as a torch.autograd.Function.

L(theta) = mean over fixed start states of the multilinear V at the start,
with V the FP64 table of the plain backup under drivetrain parameters theta
(30 entries: B row-major, A, q, diag S, diag D, c). The forward pass solves
the FP64 table on the device and runs the adjoint there (pg_solve_f64,
pg_adjoint through the C ABI); the backward pass returns the stored
gradient times the incoming one. No counterpart in the reference.
"""
from __future__ import annotations

import numpy as np
import torch

from .solver import HipSolver


class ValueLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta: torch.Tensor, prob, starts: np.ndarray, tol: float = 1e-12,
                max_sweeps: int = 20000):
        th = theta.detach().to(torch.float64).cpu().numpy()
        model = prob.model.with_theta(th)
        with HipSolver(prob.grid, model, prob.params, prob.ctl, prob.occ) as s:
            s.fill_seed(prob.seeds)
            its, linf = s.solve_f64(max_sweeps, tol)
            if not linf < tol:
                raise RuntimeError(f"FP64 table did not reach ||dV||_inf < {tol} in {its} sweeps ({linf:.2e})")
            L, grad, it = s.adjoint(starts)
            if it < 0:
                raise RuntimeError("the adjoint iteration did not converge")
        ctx.save_for_backward(torch.as_tensor(grad, dtype=theta.dtype, device=theta.device))
        return torch.tensor(L, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None, None, None, None


def value_loss(theta: torch.Tensor, prob, starts: np.ndarray, tol: float = 1e-12) -> torch.Tensor:
    return ValueLoss.apply(theta, prob, starts, tol)
