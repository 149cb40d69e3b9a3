"""torch.autograd.Function around the adjoint entry point, for gradient
checks: loss(theta) with theta the per-element log-conductivity on the GPU.

forward runs the checkpointed Backward Euler solve and the adjoint sweep
through rsd_adjoint_grad_logk (one call returns the loss and dL/dtheta);
backward scales the stored gradient by the incoming grad_output.
"""
import torch

from . import dptr


class LogConductivityLoss(torch.autograd.Function):
    @staticmethod
    def forward(fctx, theta, problem, U0, t0, dt, n_steps, obs, t_obs, tol):
        assert theta.is_cuda and theta.dtype == torch.float64 and theta.is_contiguous()
        kappa = torch.exp(theta).contiguous()
        grad = torch.empty_like(theta)
        torch.cuda.synchronize()
        loss = problem.adjoint(kappa.data_ptr(), U0, t0, dt, n_steps, obs, t_obs, tol=tol, want_grad=True,
                               d_grad=grad.data_ptr())
        fctx.save_for_backward(grad)
        return torch.tensor(loss, dtype=torch.float64, device=theta.device)

    @staticmethod
    def backward(fctx, g):
        (grad,) = fctx.saved_tensors
        return (g * grad,) + (None,) * 8


def loss_only(theta, problem, U0, t0, dt, n_steps, obs, t_obs, tol):
    """Loss without the adjoint sweep (for finite differences)."""
    kappa = torch.exp(theta).contiguous()
    torch.cuda.synchronize()
    return problem.adjoint(kappa.data_ptr(), U0, t0, dt, n_steps, obs, t_obs, tol=tol, want_grad=False)
