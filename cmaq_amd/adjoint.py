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
per-cell accepted-step counts.  With ``tape="stored"`` (the default) it also
stores every accepted step (Y, K1, K2, K3, h) in one arena of up to
``tape_fraction`` of the free device memory (cmaq_ros3_solve_tape_device),
and the backward runs only the reverse sweeps (cmaq_ros3_reverse_device).  If
the arena fills up, or with ``tape="replay"``, the stage vectors are instead
recomputed within each sync step by cmaq_ros3_adjoint_device (replay into
per-step tapes, then the reverse sweep).  The arena is released, and a large
one returned to the device, when the backward pass has run.
"""

import torch

from .ros3 import Ros3Solver


class Ros3Integrate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y0, rki, solver, chemstep, nsteps, tape="stored", tape_fraction=0.7, tape_records=None):
        if not isinstance(solver, Ros3Solver):
            raise TypeError("solver must be a Ros3Solver")
        if tape not in ("stored", "replay"):
            raise ValueError("tape must be 'stored' or 'replay'")
        n = y0.shape[-1]
        y = y0.detach().clone().contiguous()
        rk = rki.detach().contiguous()
        arena = None
        if tape == "stored" and (ctx.needs_input_grad[0] or ctx.needs_input_grad[1]) and n > 0:
            free, _ = torch.cuda.mem_get_info(y.device)
            rs = solver.query("tape_record_doubles")
            cap = min(int(free * tape_fraction) // (8 * rs), n * nsteps * 64)
            if tape_records is not None:
                cap = min(cap, int(tape_records))
            # skip the attempt when even the fewest possible steps cannot fit
            if n * nsteps * min_accepted_steps(chemstep) <= cap:
                arena = solver.new_tape(cap, device=y.device)
        states, counts, heads = [], [], []
        for _ in range(nsteps):
            states.append(y.clone())
            na = torch.empty(n, dtype=torch.int32, device=y.device)
            # a cell that cannot be integrated (DT below DTMIN) keeps its last
            # accepted state and nacc = -1; its gradients come out NaN
            if arena is not None:
                head = torch.empty(n, dtype=torch.int64, device=y.device)
                if solver.solve_tape_device(chemstep, rk, y, na, *arena, head, raise_on_fail=False):
                    heads.append(head)
                else:       # arena full: the step is complete, the backward replays
                    arena, heads = None, []
                    _release(True)
            else:
                solver.solve_device(chemstep, rk, y, na, raise_on_fail=False)
            counts.append(na)
        ctx.arena, ctx.heads = arena, heads
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
        arena, heads = ctx.arena, ctx.heads
        if arena is not None and len(heads) == k:
            for s in range(k - 1, -1, -1):
                ctx.solver.reverse_device(ctx.chemstep, rk, counts[s], *arena, heads[s], lam, grad_rki)
        else:
            for s in range(k - 1, -1, -1):
                ctx.solver.adjoint_device(ctx.chemstep, rk, states[s], counts[s], lam, grad_rki)
        if arena is not None:
            big = arena[0].numel() * 8 > (8 << 30)
            ctx.arena, ctx.heads = None, []
            del arena, heads
            _release(big)
        return (lam if ctx.needs_input_grad[0] else None), grad_rki, None, None, None, None, None, None


def min_accepted_steps(chemstep, dtstart=0.05, dtmax=10.0, facmax=10.0):
    """Fewest accepted steps any cell can take in one sync step: the step
    starts at DTSTART, may not grow on the first accepted step and grows at
    most FACMAX-fold after that, up to DTMAX, and the last step ends exactly
    at the sync step (rbsolver.F:104-113, 364-369, 509-515).  Four for a
    5-minute sync step (0.05, 0.05, 0.5, 4.4)."""
    t, h, k, first = 0.0, dtstart, 0, True
    while t < chemstep:
        h = min(h, chemstep - t)
        t += h
        k += 1
        h = min(h * (1.0 if first else facmax), dtmax)
        first = False
    return k


def _release(big):
    """Return a large tape arena's memory to the device once it is freed."""
    if big:
        torch.cuda.empty_cache()


def integrate(solver, y0, rki, chemstep=5.0, nsteps=1, tape="stored", tape_fraction=0.7, tape_records=None):
    """y_end after nsteps sync steps; differentiable in y0 and rki (reverse
    and forward mode).  tape: "stored" keeps every accepted step for the
    backward pass in an arena of up to tape_fraction of the free device
    memory (and at most tape_records records, if given), else it falls back
    to replay; "replay" recomputes the stages per sync step.  The stored
    tape is not attempted when ncells x nsteps x min_accepted_steps(chemstep)
    records cannot fit."""
    return Ros3Integrate.apply(y0, rki, solver, chemstep, nsteps, tape, tape_fraction, tape_records)
