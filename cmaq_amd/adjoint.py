"""PyTorch autograd over the gfx950 ROS3 solver and its discrete adjoint.

``integrate(solver, y0, rki, chemstep, nsteps)`` advances every cell through
``nsteps`` chemistry sync steps of ``chemstep`` minutes with the rate
constants ``rki`` held fixed (CCTM recomputes them each sync step from the
met inputs, rbdriver.F:521; here they are an input so that gradients with
respect to them can be taken).  Its backward pass is the exact discrete
adjoint of the integrator's accepted steps (step sizes held fixed), computed
on the GPU by cmaq_ros3_adjoint_device: gradients with respect to the
initial concentrations and to the rate constants (for log rate constants,
take ``rki = logk.exp()``).  Forward-mode AD (``torch.autograd.forward_ad``,
``torch.func.jvp``) is the tangent-linear counterpart: each sync step is
rerun from its stored state by cmaq_ros3_tlm_device, which takes the same
steps and carries the tangent through them.

Memory: the forward keeps the state at the start of every sync step and the
per-cell accepted-step counts; stage vectors are recomputed within each sync
step by the adjoint (replay into per-step tapes, then the reverse sweep).
"""

import torch

from .ros3 import Ros3Solver


class Ros3Integrate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y0, rki, solver, chemstep, nsteps):
        if not isinstance(solver, Ros3Solver):
            raise TypeError("solver must be a Ros3Solver")
        n = y0.shape[-1]
        y = y0.detach().clone().contiguous()
        rk = rki.detach().contiguous()
        states, counts = [], []
        for _ in range(nsteps):
            states.append(y.clone())
            na = torch.empty(n, dtype=torch.int32, device=y.device)
            # a cell that cannot be integrated (DT below DTMIN) keeps its last
            # accepted state and nacc = -1; its gradients come out NaN
            solver.solve_device(chemstep, rk, y, na, raise_on_fail=False)
            counts.append(na)
        ctx.solver, ctx.chemstep = solver, chemstep
        ctx.save_for_backward(rk, *states, *counts)
        ctx.nsteps = nsteps
        ctx.fwd_states = (rk, states)       # for jvp
        return y

    @staticmethod
    def jvp(ctx, dy0, drki, *_):
        rk, states = ctx.fwd_states
        dy = dy0.detach().clone().contiguous() if dy0 is not None else torch.zeros_like(states[0])
        dk = drki.detach().contiguous() if drki is not None else None
        for s in range(ctx.nsteps):
            y = states[s].clone()
            ctx.solver.tlm_device(ctx.chemstep, rk, y, dy, dk, raise_on_fail=False)
        return dy

    @staticmethod
    def backward(ctx, grad_y):
        saved = ctx.saved_tensors
        k = ctx.nsteps
        rk, states, counts = saved[0], saved[1:1 + k], saved[1 + k:]
        lam = grad_y.detach().clone().contiguous()
        grad_rki = torch.zeros_like(rk) if ctx.needs_input_grad[1] else None
        for s in range(k - 1, -1, -1):
            ctx.solver.adjoint_device(ctx.chemstep, rk, states[s], counts[s], lam, grad_rki)
        return (lam if ctx.needs_input_grad[0] else None), grad_rki, None, None, None


def integrate(solver, y0, rki, chemstep=5.0, nsteps=1):
    """y_end after nsteps sync steps; differentiable in y0 and rki (reverse
    and forward mode)."""
    return Ros3Integrate.apply(y0, rki, solver, chemstep, nsteps)
