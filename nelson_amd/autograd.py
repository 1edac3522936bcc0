"""torch.autograd wrapper of the converged BA solve and its adjoint.

BASolve maps (observations [E, 2], per-edge weights [E]) to the refined
camera translations [C, 3] of a BAProblem: the forward pass runs
Levenberg-Marquardt to convergence from a fixed initial state; the backward
pass is the implicit-function-theorem adjoint of the C ABI (nsn_adjoint),
which reuses the Schur factorisation of the Gauss-Newton Hessian at the
solution.  torch only carries tensors here; all the work runs in
libnelson_amd.so.
"""
import warnings

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
    """translations = BASolve.apply(obs, weight, problem, cams0, pts0, settings, hessian_steps)

    The adjoint assumes a converged solve (gradient b(x*) = 0): a solver
    failure raises, and a solve that stops on an iteration limit warns.  The
    solution and the observations of this forward are kept and put back into
    the problem before its backward, so another forward on the same problem in
    between (e.g. a finite-difference check) does not change the gradients."""

    @staticmethod
    def forward(ctx, obs, weight, problem, cams0, pts0, settings=None, hessian_steps=2):
        o = obs.detach().cpu().double().numpy().copy()
        w = weight.detach().cpu().double().numpy().copy()
        problem.set_observations(o, w)
        problem.set_state(cams0, pts0)
        lm = LevenbergMarquardt(settings or LevenbergMarquardtSettings())
        reason = lm.solve(problem)
        if reason == "SolverFailure":
            raise RuntimeError("BASolve: the LM solve failed")
        if reason in ("MaxIterationReached", "MaxSubIterationReached"):
            warnings.warn(f"BASolve: LM stopped with {reason}; the adjoint assumes a zero gradient at the "
                          "solution", RuntimeWarning)
        t = problem.camera_translations()
        ctx.problem = problem
        ctx.hessian_steps = hessian_steps
        ctx.t = t
        ctx.solution = problem.get_state_rt()
        ctx.obs, ctx.weight = o, w
        ctx.dev, ctx.dtype = obs.device, obs.dtype
        return torch.as_tensor(t, dtype=obs.dtype, device=obs.device)

    @staticmethod
    def backward(ctx, grad_t):
        p = ctx.problem
        p.set_observations(ctx.obs, ctx.weight)
        p.set_state_rt(*ctx.solution)
        g = translation_loss_grad(ctx.t, grad_t.detach().cpu().double().numpy())
        dobs, dw = p.adjoint(g_cams=g, hessian_steps=ctx.hessian_steps)
        go = torch.as_tensor(dobs, dtype=ctx.dtype, device=ctx.dev)
        gw = torch.as_tensor(dw, dtype=ctx.dtype, device=ctx.dev)
        return go, gw, None, None, None, None, None
