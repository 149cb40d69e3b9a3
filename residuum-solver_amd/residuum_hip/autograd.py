"""torch.autograd.Function around the adjoint entry point, for gradient
checks: loss(theta) with theta the per-element log-conductivity.

theta is a float64 host tensor.  forward uploads kappa = exp(theta) to the
library's device memory, runs the checkpointed Backward Euler solve and the
adjoint sweep through rsd_adjoint_grad_logk (one call returns the loss and
dL/dtheta) and keeps the gradient; backward scales it by grad_output.  The
GPU work runs on the library's own HIP stream, so torch's device runtime is
not involved.
"""
import numpy as np
import torch


class LogConductivityLoss(torch.autograd.Function):
    @staticmethod
    def forward(fctx, theta, problem, U0, t0, dt, n_steps, obs, t_obs, tol):
        th = theta.detach().cpu().numpy().astype(np.float64)
        ctx = problem.ctx
        d_kappa = ctx.from_numpy(np.exp(th))
        d_grad = ctx.empty(th.size)
        loss = problem.adjoint(d_kappa, U0, t0, dt, n_steps, obs, t_obs, tol=tol, want_grad=True, d_grad=d_grad)
        fctx.save_for_backward(torch.from_numpy(d_grad.numpy()))
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(fctx, g):
        (grad,) = fctx.saved_tensors
        return (g * grad,) + (None,) * 8


def loss_only(theta, problem, U0, t0, dt, n_steps, obs, t_obs, tol):
    """Loss without the adjoint sweep (for finite differences)."""
    th = theta.detach().cpu().numpy().astype(np.float64)
    d_kappa = problem.ctx.from_numpy(np.exp(th))
    return problem.adjoint(d_kappa, U0, t0, dt, n_steps, obs, t_obs, tol=tol, want_grad=False)
