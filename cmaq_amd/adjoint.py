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
per-step tapes, then the reverse sweep).  The solver keeps the arena after the
backward pass and reuses it for its next integration (sized from this one's
record count); ``release_tape(solver)`` returns it to the device.
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
            arena = _take_arena(solver, y.device, n, nsteps, chemstep, tape_fraction, tape_records)
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
                    _give_back(solver, arena, overflowed=True)
                    arena, heads = None, []
            else:
                solver.solve_device(chemstep, rk, y, na, raise_on_fail=False)
            counts.append(na)
        if arena is not None:     # sizes the next integration's arena
            solver._tape_used = (int(arena[1].item()), n, nsteps)
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
            ctx.arena, ctx.heads = None, []
            _give_back(ctx.solver, arena)
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


def _take_arena(solver, device, n, nsteps, chemstep, tape_fraction, tape_records):
    """A stored-tape arena (tape, count) with its count at zero, or None.

    The solver keeps one arena between integrations and reuses it, so that an
    optimisation loop neither allocates nor returns ~100 GB per call.  Its
    size comes from the previous integration's record count (plus 25 % and one
    record per cell and sync step) when that had the same cells and sync
    steps, else from the old upper bound (tape_fraction of the free memory, at
    most 64 records per cell and sync step).  An arena still held by an
    integration whose backward has not run is not shared: the call then gets
    a fresh one."""
    rs = solver.query("tape_record_doubles")
    used = getattr(solver, "_tape_used", None)
    want = None
    if used is not None and used[1:] == (n, nsteps):
        want = int(used[0] * 1.25) + n * nsteps
    held = getattr(solver, "_arena", None)
    if held is not None and not getattr(solver, "_arena_busy", False):
        cap = held[0].numel() // rs
        if (want is None or want <= cap <= 2 * want) and (tape_records is None or cap <= int(tape_records)) \
                and cap >= n * nsteps * min_accepted_steps(chemstep):
            held[1].zero_()
            solver._arena_busy = True
            return held
        solver._arena = None            # too small or too large: replace it
        del held
        torch.cuda.empty_cache()
    free, _ = torch.cuda.mem_get_info(device)
    cap = min(int(free * tape_fraction) // (8 * rs), n * nsteps * 64)
    if want is not None:
        cap = min(cap, want)
    if tape_records is not None:
        cap = min(cap, int(tape_records))
    # skip the attempt when even the fewest possible steps cannot fit
    if n * nsteps * min_accepted_steps(chemstep) > cap:
        return None
    arena = solver.new_tape(cap, device=device)
    if getattr(solver, "_arena_busy", False):
        return arena                    # private to this integration
    solver._arena, solver._arena_busy = arena, True
    return arena


def _give_back(solver, arena, overflowed=False):
    """End of an integration's use of its arena: the solver's own arena is
    kept for the next call (an overflowed one is dropped); a private one is
    freed."""
    if overflowed:
        solver._tape_used = None        # the next arena comes from the upper bound
    if getattr(solver, "_arena", None) is arena:
        solver._arena_busy = False
        if overflowed:
            solver._arena = None
            torch.cuda.empty_cache()


def release_tape(solver):
    """Return the solver's kept stored-tape arena to the device."""
    solver._arena, solver._arena_busy = None, False
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
