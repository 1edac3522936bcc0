"""torch.autograd wrapper of the converged BA solve and its adjoint.

BASolve maps (observations [E, 2], per-edge weights [E]) to the refined
camera translations [C, 3] of a BAProblem: the forward pass runs
Levenberg-Marquardt to convergence from a fixed initial state; the backward
pass is the implicit-function-theorem adjoint of the C ABI (nsn_adjoint),
which reuses the Schur factorisation of the Gauss-Newton Hessian at the
solution.  torch only carries tensors here; all the work runs in
libnelson_amd.so.
"""
import numpy as np
import torch

from .problem import LevenbergMarquardt, LevenbergMarquardtSettings


def translation_loss_grad(t, g_t):
    """dL/d(camera) in the left-perturbation coordinates (omega, upsilon) of
    oplus (T <- exp(delta) T) for a loss with gradient g_t = dL/dt on the
    translations: dt/d(omega) = -[t]x and dt/d(upsilon) = I, so
    g_omega = t x g_t and g_upsilon = g_t."""
    t = np.asarray(t, np.float64).reshape(-1, 3)
    g_t = np.asarray(g_t, np.float64).reshape(-1, 3)
    return np.concatenate([np.cross(t, g_t), g_t], axis=1)


class BASolve(torch.autograd.Function):
    """translations = BASolve.apply(obs, weight, problem, cams0, pts0, settings, hessian_steps)"""

    @staticmethod
    def forward(ctx, obs, weight, problem, cams0, pts0, settings=None, hessian_steps=2):
        problem.set_observations(obs.detach().cpu().double().numpy(), weight.detach().cpu().double().numpy())
        problem.set_state(cams0, pts0)
        lm = LevenbergMarquardt(settings or LevenbergMarquardtSettings())
        reason = lm.solve(problem)
        if reason == "SolverFailure":
            raise RuntimeError("BASolve: the LM solve failed")
        t = problem.camera_translations()
        ctx.problem = problem
        ctx.hessian_steps = hessian_steps
        ctx.t = t
        ctx.dev, ctx.dtype = obs.device, obs.dtype
        return torch.as_tensor(t, dtype=obs.dtype, device=obs.device)

    @staticmethod
    def backward(ctx, grad_t):
        g = translation_loss_grad(ctx.t, grad_t.detach().cpu().double().numpy())
        dobs, dw = ctx.problem.adjoint(g_cams=g, hessian_steps=ctx.hessian_steps)
        go = torch.as_tensor(dobs, dtype=ctx.dtype, device=ctx.dev)
        gw = torch.as_tensor(dw, dtype=ctx.dtype, device=ctx.dev)
        return go, gw, None, None, None, None, None
