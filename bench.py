"""Throughput of the gfx950 ROS3 gas-chemistry solver on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload regional|clean|mixed]

A "step" is one 5-minute chemistry sync step (CALC_RCONST + RBSOLVER) over
every cell of the workload, with inputs already resident in HBM.  The
default workload is the regional CB6r5 case: 459 x 299 x 35 = 4,803,435
cells, layer-major, J scaled by exp(-k/20) per layer.  The metric is
attempted ROS3 cell-steps per second (accepted steps are reported too).

The JSON line also carries
  roofline     - algorithmic HBM bytes of the solver kernel (32 bytes per LU
                 entry per attempted step, i.e. the factored matrix written once
                 and read by the three stage solves, plus the per-cell input/
                 output of one sync step) over its HIP-event-timed duration;
  cpu_baseline - the CPU oracle (a C restatement of the reference ROS3, built
                 with the reference's gcc flags) timed on a bounded sample of
                 the same cells on the host cores (one process per core, at
                 most 16), each with the LORDERING stiffness pre-pass and
                 stock 50-cell blocks in stiffness order; also the 4096-cell
                 plain reference case in full (config0_full);
  sync_loop    - the uneven-stiffness workload: 1,048,576 mixed-regime cells
                 (40 % clean day, 30 % night, 30 % near-source), 12 sync steps
                 of 5 min replayed from one captured hipGraph (rates -> solve
                 -> checkpoint per step);
  fwd_adj_ratio - forward + discrete adjoint over forward alone, on the
                 differentiation workload: 262,144 mixed-regime cells, 12
                 sync steps of 5 min, loss = sum(O3 + 0.1 NO2) at the end,
                 gradients with respect to all initial concentrations and
                 all rate constants, every accepted step stored in the
                 forward (details under "adjoint"; the mode that recomputes
                 the stages per sync step under "adjoint.replay_mode");
  adjoint_larger_mechanism - the same for CRACMM2 on 2,097,152 mixed cells
                 with VOC/SVOC species, stages recomputed per sync step;
  smvgear      - CMAQ's other sparse solver, SMVGEAR (variable-order Gear
                 BDF), on 1,048,576 regional cells, one 5-min sync step:
                 successful Gear cell-steps/s, its time against ROS3 on the
                 same cells, and the oracle's SMVGEAR on a sample of them on
                 the host cores.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E, GB/s
PROFILES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")
FP64_PEAK_TFS = 78.6       # MI355X FP64 vector, TFLOP/s


def _args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--mech", default="cb6r5_ae7_aq")
    p.add_argument("--workload", default="regional", choices=["regional", "clean", "mixed"])
    p.add_argument("--ncells", type=int, default=0, help="override the cell count (smaller runs)")
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline")
    p.add_argument("--cpu-workers", type=int, default=0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--adj-cells", type=int, default=262144)
    p.add_argument("--adj-sync-steps", type=int, default=12)
    p.add_argument("--adj-mech", default=None, help="mechanism of the adjoint leg (default --mech)")
    p.add_argument("--adj-svoc", action="store_true",
                   help="adjoint leg: VOC/SVOC species log-uniform on [1e-5, 5] ppb (the CRACMM2 workload)")
    p.add_argument("--no-adj", action="store_true")
    p.add_argument("--adj-tape", default="stored", choices=["stored", "replay"],
                   help="adjoint leg's headline mode: every accepted step stored, or stages recomputed")
    p.add_argument("--loop-cells", type=int, default=1048576)
    p.add_argument("--loop-sync-steps", type=int, default=12)
    p.add_argument("--no-loop", action="store_true")
    p.add_argument("--no-chem", action="store_true", help="skip the CHEM gather/scatter leg")
    p.add_argument("--big-adj-mech", default="cracmm2")
    p.add_argument("--big-adj-cells", type=int, default=2097152)
    p.add_argument("--gear-cells", type=int, default=1048576)
    p.add_argument("--no-gear", action="store_true", help="skip the SMVGEAR leg")
    p.add_argument("--gear-tail-after", type=int, default=None,
                   help="attempted Gear steps before a cell moves to the one-cell-per-wavefront kernel "
                        "(cmaq_gear_set_tail; default: the library's)")
    p.add_argument("--no-big-adj", action="store_true",
                   help="skip the larger-mechanism forward + adjoint leg (CRACMM2, replay mode)")
    return p.parse_args()


def _cpu_worker(job):
    """Times the oracle on one slice of sample cells (runs in a child process)
    as one MPI rank of the reference runs its cells: the LORDERING pre-pass
    (one CALC_RCONST + RBFEVAL per block for the stiffness estimate ERRMX2,
    rbsolver.F:346-360), the cells sorted by it (rbdriver.F:683-724), then
    CALC_RCONST + RBSOLVER in 50-cell blocks of that order."""
    mech_name, cells, seed, workload, nlay_cells, blk = job
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from cmaq_amd.inputs import make_cells
    from cmaq_amd.ros3 import load_mechanism
    from oracle.oracle import Oracle
    mech = load_mechanism(mech_name)
    o = Oracle(mech)
    d = make_cells(mech, np.asarray(cells), seed=seed, workload=workload, nlay_cells=nlay_cells)
    t0 = time.perf_counter()
    err = o.stiffness(d["temp"], d["pres"], d["h2o"], d["rj"], d["y"], 1.0e-3, 1.0e-7, blksize=blk)
    order = np.argsort(err, kind="stable")
    t_pre = time.perf_counter() - t0
    r = o.run(5.0, d["temp"][order], d["pres"][order], d["h2o"][order], d["rj"][:, order], d["y"][:, order],
              1.0e-3, 1.0e-7, blksize=blk)
    dt = time.perf_counter() - t0
    # cell-steps: every cell of a block takes every step of the block
    return dt, int((r["nacc"] + r["nrej"]).sum()), int(r["nacc"].sum()), r["rc"], t_pre


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_run(mech_name, workers, jobs):
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    t0 = time.perf_counter()
    with ctx.Pool(workers) as pool:
        res = pool.map(_cpu_worker, jobs)
    return res, time.perf_counter() - t0


def cpu_baseline(args, mech, ncells, nlay_cells):
    # one process per core, as CMAQ runs one MPI rank per core; at most the
    # 16 cores of one GPU's share of the host
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    workers = args.cpu_workers or min(16, avail)
    # calibrate: ~2.5 ms per cell-sync-step per core for CB6r5; pick a sample that
    # fits the budget, spread evenly over the domain
    per_cell_s = 2.5e-3 * (1.0 if mech["nrxns"] < 400 else 3.0)
    per_worker = max(50, int(args.cpu_seconds / per_cell_s))
    per_worker -= per_worker % 50
    total = per_worker * workers
    idx = np.linspace(0, ncells - 1, total).astype(np.int64)
    jobs = [(args.mech, idx[w::workers].tolist(), args.seed, args.workload, nlay_cells, 50)
            for w in range(workers)]
    res, wall = _cpu_run(args.mech, workers, jobs)
    att = sum(r[1] for r in res)
    acc = sum(r[2] for r in res)
    busy = max(r[0] for r in res)
    out = {"value": att / busy, "unit": "attempted cell-steps/s", "cores": workers, "kind": "port",
           "accepted_per_s": acc / busy,
           "cpu_model": _cpu_model(), "cpus_available": avail,
           "sample": f"{total} cells evenly spaced over the {ncells}-cell {args.workload} workload, "
                     f"one 5-min step, {workers} processes on disjoint cells (one per core, as CMAQ's MPI "
                     f"ranks), each with the LORDERING stiffness pre-pass and 50-cell blocks in stiffness "
                     f"order (the reference's performance configuration); slowest process {busy:.1f} s "
                     f"(pre-pass {max(r[4] for r in res):.2f} s), wall {wall:.1f} s",
           "failed_blocks": int(sum(1 for r in res if r[3] != 0))}
    # the plain reference case in full: 4096 clean cells, one 5-min step
    cells0 = np.arange(4096)
    jobs0 = [(args.mech, cells0[w::workers].tolist(), 0, "clean", None, 50) for w in range(workers)]
    res0, wall0 = _cpu_run(args.mech, workers, jobs0)
    busy0 = max(r[0] for r in res0)
    out["config0_full"] = {"cells": 4096, "attempted_cell_steps_per_s": sum(r[1] for r in res0) / busy0,
                           "seconds": busy0, "wall_s": wall0,
                           "failed_blocks": int(sum(1 for r in res0 if r[3] != 0))}
    return out


def measured_traffic_per_step(mech_name, variant):
    """HBM bytes per attempted cell-step of the solver kernel, from the
    calibrated rocprofv3 FETCH_SIZE/WRITE_SIZE summary under profiles/, or
    None when no summary covers this mechanism and kernel variant."""
    import glob
    import re
    key = re.escape(f"[{mech_name}/{variant}]")
    for path in sorted(glob.glob(os.path.join(PROFILES, "*_pmc.md"))):
        with open(path) as f:
            m = re.search(r"traffic_bytes_per_attempted_cell_step" + key + r"\s*=\s*([0-9.eE+]+)", f.read())
        if m:
            return float(m.group(1)), os.path.basename(path)
    return None, None


def row_stream_ceiling(waves_per_cu):
    """Read throughput of the solver's access stream (scattered 512-byte lane
    rows from per-wave 2.75 MB regions) measured by cmaq_amd/hbm_rows at the
    solver's resident waves per CU (4 or 8); None if the probe is absent."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cmaq_amd", "hbm_rows")
    if not os.path.exists(exe):
        return None
    try:
        r = subprocess.run([exe, "4000"], capture_output=True, text=True, timeout=120)
    except subprocess.TimeoutExpired:
        return None
    best = None
    for line in r.stdout.splitlines():
        f = line.split()
        if len(f) == 8 and f[1] == str(waves_per_cu) and f[3] == "8":
            best = float(f[7])
    return best


def loop_leg(args, mech, solver, dev):
    """12 sync steps of the mixed workload through the captured graph."""
    import torch
    from cmaq_amd.inputs import make_cells_torch
    n, k = args.loop_cells, args.loop_sync_steps
    d = make_cells_torch(mech, n, seed=args.seed + 2, workload="mixed", device=dev)
    y = d["y"].clone()
    na = torch.empty((k, n), dtype=torch.int32, device=dev)
    nj = torch.empty((k, n), dtype=torch.int32, device=dev)
    args_ = (5.0, k, d["temp"], d["pres"], d["h2o"], d["rj"], y, None, na, nj)
    solver.run_steps_device(*args_, raise_on_fail=False)       # captures the graph
    y.copy_(d["y"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.run_steps_device(*args_, raise_on_fail=False)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    dev_ms = solver.last_steps_timing()
    ok = (na >= 0).all(dim=0)
    att = int((na.clamp(min=0) + nj).sum())
    per_cell = (na.clamp(min=0) + nj).sum(dim=0)[ok].to(torch.float64)
    q = torch.quantile(per_cell, torch.tensor([0.01, 0.5, 0.99], dtype=torch.float64, device=dev)).tolist()
    return {"workload": f"{args.mech} ROS3, mixed regimes (40% clean day, 30% night, 30% near-source with "
                        f"NO/NO2 log-uniform on [10,500] ppb and VOC x10), {n} cells, {k} sync steps of "
                        f"5 min, one hipGraph per sync step (rates -> solve -> checkpoint)",
            "ncells": n, "sync_steps": k, "ms": wall * 1e3, "device_ms": dev_ms,
            "attempted_cell_steps_per_s": att / wall,
            "attempted_steps_per_cell_p01_p50_p99": q,
            "cells_failing_some_step": int((~ok).sum())}


def adjoint_leg(args, mech, solver, dev):
    """Times 12 sync steps forward, then forward + adjoint through autograd."""
    import torch
    from cmaq_amd.adjoint import integrate, release_tape
    from cmaq_amd.inputs import make_cells_torch
    from cmaq_amd.ros3 import Ros3Solver, load_mechanism
    if args.adj_mech and args.adj_mech != args.mech:
        mech = load_mechanism(args.adj_mech)
        solver = Ros3Solver(mech)
    n, k = args.adj_cells, args.adj_sync_steps
    d = make_cells_torch(mech, n, seed=args.seed + 3, workload="mixed", device=dev,
                         svoc_ppb=(1e-5, 5.0) if args.adj_svoc else None)
    rki = torch.empty((mech["nrxns"], n), dtype=torch.float64, device=dev)
    solver.rates_device(d["temp"], d["pres"], d["h2o"], d["rj"], rki)
    names = mech["chemistry_spc"]
    io3, ino2 = names.index("O3"), names.index("NO2")
    y = torch.empty_like(d["y"])
    na = torch.empty(n, dtype=torch.int32, device=dev)

    def fwd():
        y.copy_(d["y"])
        att = 0
        for _ in range(k):
            solver.solve_device(5.0, rki, y, na, raise_on_fail=False)
            att += int(na.clamp(min=0).sum())
        return att

    def fwd_adj(tape):
        y0 = d["y"].clone().requires_grad_(True)
        kk = rki.clone().requires_grad_(True)
        yend = integrate(solver, y0, kk, 5.0, k, tape=tape)
        loss = (yend[io3] + 0.1 * yend[ino2]).sum()
        loss.backward()
        return y0.grad, kk.grad

    fwd()                                   # warm-up (allocations, code load)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    accepted = fwd()
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    print(f"[bench] adjoint leg: forward {t_fwd * 1e3:.0f} ms", file=sys.stderr, flush=True)
    # the replay mode (only sync-step states kept, stages recomputed per sync
    # step), then the stored tape (every accepted step kept, reverse sweeps
    # only).  Each mode releases its large tape when it returns; a large
    # allocation right after such a release can wait seconds for the device
    # to reclaim the memory.  In this order the CB6r5 workload shows no such
    # wait; the reverse order adds it to the replay mode.
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fwd_adj("replay")
    torch.cuda.synchronize()
    t_replay = time.perf_counter() - t0
    rep, rev, chunks = solver.last_adjoint_timing()
    print(f"[bench] adjoint leg: forward+adjoint with replay {t_replay * 1e3:.0f} ms", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gy, gk = fwd_adj(args.adj_tape)
    torch.cuda.synchronize()
    t_both = time.perf_counter() - t0
    rep_s, rev_s, chunks_s = solver.last_adjoint_timing()
    stored = chunks_s == 0
    print(f"[bench] adjoint leg: forward+adjoint with {'stored tape' if stored else 'replay'} {t_both * 1e3:.0f} ms",
          file=sys.stderr, flush=True)
    # the same call again, as in an optimisation loop: the solver's arena is
    # reused (sized from the first call's record count), nothing is allocated
    del gy, gk
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gy, gk = fwd_adj(args.adj_tape)
    torch.cuda.synchronize()
    t_again = time.perf_counter() - t0
    print(f"[bench] adjoint leg: second call {t_again * 1e3:.0f} ms", file=sys.stderr, flush=True)
    bad = int((~torch.isfinite(gy)).any(dim=0).sum())
    # one sync step with one tangent-linear direction (d ln O3 of every cell
    # and d ln k of all reactions at once) against the plain solve
    y.copy_(d["y"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.solve_device(5.0, rki, y, na, raise_on_fail=False)
    torch.cuda.synchronize()
    t_one = time.perf_counter() - t0
    y.copy_(d["y"])
    dy = torch.zeros_like(y)
    dy[io3] = d["y"][io3]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.tlm_device(5.0, rki, y, dy, rki, nacc=na, raise_on_fail=False)
    torch.cuda.synchronize()
    t_tlm = time.perf_counter() - t0
    y.copy_(d["y"])
    dy4 = torch.zeros((4,) + tuple(y.shape), dtype=y.dtype, device=dev)
    for p, nm in enumerate(("O3", "NO2", "FORM", "ISOP")):
        if nm in names:
            dy4[p, names.index(nm)] = d["y"][names.index(nm)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.tlm_device(5.0, rki, y, dy4, nacc=na, raise_on_fail=False)
    torch.cuda.synchronize()
    t_tlm4 = time.perf_counter() - t0
    release_tape(solver)        # the kept arena would starve the next legs
    print(f"[bench] tangent-linear: one sync step {t_one * 1e3:.0f} ms, with one direction {t_tlm * 1e3:.0f} ms, "
          f"with four {t_tlm4 * 1e3:.0f} ms", file=sys.stderr, flush=True)
    return t_both / t_fwd, {
        "workload": f"{mech['mechname'].lower()} ROS3, mixed regimes{' + VOC/SVOC' if args.adj_svoc else ''}, "
                    f"{n} cells, {k} sync steps of 5 min; "
                    f"loss = sum over cells of final O3 + 0.1 NO2; gradients w.r.t. all initial "
                    f"concentrations and all rate constants; "
                    + ("every accepted step stored, reverse sweeps only" if stored
                       else "sync-step states stored, stages recomputed"),
        "ncells": n, "sync_steps": k, "variant": solver.variant, "fwd_ms": t_fwd * 1e3, "fwd_adj_ms": t_both * 1e3,
        "second_call_fwd_adj_ms": t_again * 1e3, "second_call_fwd_adj_over_fwd": t_again / t_fwd,
        "tape": "stored" if stored else "replay",
        "accepted_cell_steps": accepted,
        "last_sync_step_reverse_ms": rev_s,
        "replay_mode": {"fwd_adj_ms": t_replay * 1e3, "fwd_adj_over_fwd": t_replay / t_fwd,
                        "last_sync_step_replay_ms": rep, "last_sync_step_reverse_ms": rev, "tape_chunks": chunks},
        "cells_without_gradient": bad,
        "tlm_one_direction_over_fwd": t_tlm / t_one, "tlm_sync_step_ms": t_tlm * 1e3,
        "tlm_four_directions_over_fwd": t_tlm4 / t_one,
        "fwd_sync_step_ms": t_one * 1e3}


def big_adjoint_leg(args, dev):
    """The larger-mechanism workload: CRACMM2 forward plus discrete adjoint
    over 12 sync steps of 5 min on 2,097,152 mixed cells with the VOC/SVOC
    species log-uniform on [1e-5, 5] ppb; only sync-step states are kept and
    the stages are recomputed within each sync step (the tape of every
    accepted step would need about 1.5 TB)."""
    import torch
    from cmaq_amd.adjoint import integrate
    from cmaq_amd.inputs import make_cells_torch
    from cmaq_amd.ros3 import Ros3Solver, load_mechanism
    mech = load_mechanism(args.big_adj_mech)
    solver = Ros3Solver(mech)
    n, k = args.big_adj_cells, args.adj_sync_steps
    d = make_cells_torch(mech, n, seed=args.seed + 4, workload="mixed", device=dev, svoc_ppb=(1e-5, 5.0))
    rki = torch.empty((mech["nrxns"], n), dtype=torch.float64, device=dev)
    solver.rates_device(d["temp"], d["pres"], d["h2o"], d["rj"], rki)
    del d["temp"], d["pres"], d["h2o"], d["rj"]
    names = mech["chemistry_spc"]
    io3, ino2 = names.index("O3"), names.index("NO2")
    y = torch.empty_like(d["y"])
    na = torch.empty(n, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    y.copy_(d["y"])
    accepted = 0
    for _ in range(k):
        solver.solve_device(5.0, rki, y, na, raise_on_fail=False)
        accepted += int(na.clamp(min=0).sum())
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    print(f"[bench] larger-mechanism leg: forward {t_fwd * 1e3:.0f} ms", file=sys.stderr, flush=True)
    del y
    y0 = d["y"].clone().requires_grad_(True)
    kk = rki.requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    yend = integrate(solver, y0, kk, 5.0, k, tape="replay")
    loss = (yend[io3] + 0.1 * yend[ino2]).sum()
    loss.backward()
    torch.cuda.synchronize()
    t_both = time.perf_counter() - t0
    rep, rev, chunks = solver.last_adjoint_timing()
    bad = int((~torch.isfinite(y0.grad)).any(dim=0).sum())
    print(f"[bench] larger-mechanism leg: forward+adjoint {t_both * 1e3:.0f} ms", file=sys.stderr, flush=True)
    return {"workload": f"{mech['mechname'].lower()} ROS3, mixed regimes + VOC/SVOC log-uniform on [1e-5, 5] "
                        f"ppb, {n} cells, {k} sync steps of 5 min; loss = sum over cells of final O3 + 0.1 NO2; "
                        f"gradients w.r.t. all initial concentrations and all rate constants; sync-step states "
                        f"stored, stages recomputed per sync step",
            "ncells": n, "sync_steps": k, "variant": solver.variant, "fwd_ms": t_fwd * 1e3,
            "fwd_adj_ms": t_both * 1e3, "fwd_adj_over_fwd": t_both / t_fwd, "accepted_cell_steps": accepted,
            "last_sync_step_replay_ms": rep, "last_sync_step_reverse_ms": rev, "tape_chunks": chunks,
            "cells_without_gradient": bad}


def _gear_cpu_worker(job):
    """Times the oracle's SMVGEAR on one slice of cells with per-cell control
    (blocks of one cell, the GPU kernel's computation; runs in a child
    process)."""
    mech_name, cells, seed, workload, nlay_cells = job
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from cmaq_amd.inputs import make_cells
    from cmaq_amd.ros3 import load_mechanism
    from oracle.oracle import Oracle
    mech = load_mechanism(mech_name)
    o = Oracle(mech)
    d = make_cells(mech, np.asarray(cells), seed=seed, workload=workload, nlay_cells=nlay_cells)
    t0 = time.perf_counter()
    r = o.gear_run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"], blksize=1)
    dt = time.perf_counter() - t0
    return dt, int(np.clip(r["nsteps"], 0, None).sum()), r["rc"]


def gear_leg(args, mech, dev, workers):
    """SMVGEAR (cmaq_gear_run_device) and ROS3 on the same regional cells, one
    5-min sync step each, and the oracle's SMVGEAR on a bounded sample of them
    with per-cell control on the host cores."""
    import torch
    from cmaq_amd.inputs import make_cells_torch
    from cmaq_amd.ros3 import Ros3Solver
    n = args.gear_cells
    nlay_cells = max(1, n // 35)
    d = make_cells_torch(mech, n, seed=args.seed + 6, workload="regional", nlay_cells=nlay_cells, device=dev)
    s = Ros3Solver(mech)
    if args.gear_tail_after is not None:
        s.set_gear_tail(args.gear_tail_after, 0)
    y = torch.empty_like(d["y"])
    ns_ = torch.zeros(n, dtype=torch.int32, device=dev)
    nf = torch.zeros(n, dtype=torch.int32, device=dev)
    best = None
    for k in range(2):
        y.copy_(d["y"])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.gear_run_device(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], y, ns_, nf, raise_on_fail=False)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1000.0
        best = ms if best is None else min(best, ms)
        print(f"[bench] SMVGEAR call {k + 1}/2: {ms:.0f} ms, {s.query('gear_tail_cells_last')} long cells finished "
              f"one per wavefront", file=sys.stderr, flush=True)
    gear_ms = s.gear_last_timing()[1]
    steps = ns_.clamp(min=0).to(torch.int64)
    nsteps_total = int(steps.sum())
    attempts_max = int((ns_.clamp(min=0) + nf).max())
    q = torch.quantile(steps.to(torch.float64), torch.tensor([0.01, 0.5, 0.99], dtype=torch.float64, device=dev))
    y.copy_(d["y"])
    na = torch.zeros(n, dtype=torch.int32, device=dev)
    nj = torch.zeros(n, dtype=torch.int32, device=dev)
    s.run_device(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], y, na, nj, raise_on_fail=False)
    torch.cuda.synchronize()
    ros3_ms = s.last_timing()[1]
    out = {"workload": f"{args.mech} SMVGEAR, regional, {n} cells, one 5-min sync step, GEAR_RTOL 1e-3, "
                       f"GEAR_MAX/MIN_ATOL 1e-7/1e-9 ppm",
           "ncells": n, "ms": best, "gear_kernel_ms": gear_ms,
           "successful_cell_steps_per_s": nsteps_total / (gear_ms / 1000.0),
           "steps_per_cell_p01_p50_p99": [float(v) for v in q.cpu()],
           "failed_cells": int((ns_ < 0).sum()), "attempted_steps_longest_cell": attempts_max,
           "long_cells_one_per_wavefront": int(s.query("gear_tail_cells_last")),
           "ros3_kernel_ms_same_cells": ros3_ms, "gear_over_ros3_time": gear_ms / ros3_ms}
    s.close()
    if not args.no_cpu:
        per_worker = 600
        idx = np.linspace(0, n - 1, per_worker * workers).astype(np.int64)
        jobs = [(args.mech, idx[w::workers].tolist(), args.seed + 6, "regional", nlay_cells) for w in range(workers)]
        import multiprocessing as mp
        t0 = time.perf_counter()
        with mp.get_context("spawn").Pool(workers) as pool:
            res = pool.map(_gear_cpu_worker, jobs)
        busy = max(r[0] for r in res)
        out["cpu_baseline"] = {"value": sum(r[1] for r in res) / busy, "unit": "successful cell-steps/s",
                               "cores": workers, "kind": "port",
                               "sample": f"{per_worker * workers} cells evenly spaced over the {n} cells, "
                                         f"{workers} processes, per-cell control; slowest process {busy:.1f} s",
                               "failed_blocks": int(sum(1 for r in res if r[2] != 0))}
    return out


def chem_io_leg(args, mech, solver, d, dims, dev):
    """CHEM-level entry points on the headline grid: the CGRID gather and
    scatter kernels (HBM-bound copies with unit conversion) and one whole
    cmaq_ros3_chem_device call."""
    import torch
    ncol, nrow, nlay = dims
    n = ncol * nrow * nlay
    ns, nph, nhet = mech["numb_mech_spc"], mech["nphotab"], mech["nhetero"]
    mwair, mwwat, rgas = 28.9628, 18.0153, 8.314510
    g = torch.Generator(device="cpu").manual_seed(args.seed + 5)
    nspcsd = ns + 64
    idx = (torch.randperm(nspcsd, generator=g)[:ns] + 1).to(torch.int32)
    conv = np.array([1 if c else 0 for c in mech["convert_conc"]], np.int32)
    mw = np.array(mech["species_molwt"], np.float64)
    solver.set_cgrid_map(nspcsd, idx.numpy(), conv, mw)
    pres_pa = d["pres"] * 101325.0
    dens = pres_pa * mwair * 1.0e-3 / (rgas * d["temp"])
    to_ug = torch.as_tensor(mw, device=dev)[:, None] * dens[None, :] / (1.0e-3 * mwair)
    cgrid = torch.full((nspcsd, n), -1.0, dtype=torch.float32, device=dev)
    convd = torch.as_tensor(conv, device=dev)[:, None] == 1
    cgrid[idx.to(dev).long() - 1] = torch.where(convd, d["y"] * to_ug, d["y"]).float()
    del to_ug, convd
    grid = {"ncols": ncol, "nrows": nrow, "nlays": nlay, "cgrid": cgrid,
            "ta": d["temp"].float(), "dens": dens.float(), "pres": pres_pa.float(),
            "qv": (d["h2o"] * 1.0e-6 * mwwat / mwair).float(),
            "seawater_zone": torch.rand(ncol * nrow, generator=g).to(dev),
            "seaice": (0.2 * torch.rand(ncol * nrow, generator=g)).to(dev),
            "rj": d["rj"].float(),
            "khetero": (1e-3 * torch.rand((n, nhet + 1), generator=g, dtype=torch.float64)).to(dev)}
    del dens, pres_pa
    blk = solver.gather_device(grid)
    torch.cuda.synchronize()
    reps = 5
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record()
    for _ in range(reps):
        solver.gather_device(grid, out=blk)
    ev[1].record()
    for _ in range(reps):
        solver.scatter_device(grid, blk["y"])
    ev[2].record()
    torch.cuda.synchronize()
    g_ms = ev[0].elapsed_time(ev[1]) / reps
    s_ms = ev[1].elapsed_time(ev[2]) / reps
    del blk
    # algorithmic bytes per cell: every input read once, every output written once
    b_gather = 4 * 4 + 4 * ns + 4 * nph + 8 * nhet + 8.0 / nlay + 8 * (4 + ns + nph + nhet)
    b_scatter = 4 + 8 * ns + 4 * ns
    gb_g = b_gather * n / (g_ms / 1e3) / 1e9
    gb_s = b_scatter * n / (s_ms / 1e3) / 1e9
    na = torch.zeros(n, dtype=torch.int32, device=dev)
    nj = torch.zeros(n, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.chem_device(grid, 5.0, nacc=na, nrej=nj, raise_on_fail=False)
    torch.cuda.synchronize()
    chem_s = time.perf_counter() - t0
    att = int((na.clamp(min=0) + nj).sum())
    return {"grid": f"{ncol}x{nrow}x{nlay}", "ncells": n, "nspcsd": nspcsd,
            "gather_kernel_ms": g_ms, "scatter_kernel_ms": s_ms,
            "gather_bytes_per_cell": b_gather, "scatter_bytes_per_cell": b_scatter,
            "gather_GBps": gb_g, "scatter_GBps": gb_s,
            "gather_frac_of_hbm_peak": gb_g / HBM_PEAK_GBS, "scatter_frac_of_hbm_peak": gb_s / HBM_PEAK_GBS,
            "chem_device_ms": chem_s * 1e3, "chem_device_attempted_cell_steps_per_s": att / chem_s,
            "chem_device_failed_cells": int((na < 0).sum())}


def main():
    args = _args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the ROS3 path shards by cell range with no inter-GPU "
              f"traffic, and runs on more than one GPU are not part of this build (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from cmaq_amd.inputs import make_cells_torch
    from cmaq_amd.ros3 import Ros3Solver, load_mechanism

    mech = load_mechanism(args.mech)
    ncol, nrow, nlay = 459, 299, 35
    nlay_cells = ncol * nrow
    if args.ncells:
        ncells = args.ncells
        nlay_cells = max(1, ncells // nlay)
    else:
        ncells = ncol * nrow * nlay if args.workload == "regional" else 1048576
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    print(f"[bench] generating {ncells} cells ({args.workload}, {args.mech})", file=sys.stderr, flush=True)
    d = make_cells_torch(mech, ncells, seed=args.seed, workload=args.workload, nlay_cells=nlay_cells, device=dev)
    y0 = d["y"]
    y = torch.empty_like(y0)
    na = torch.zeros(ncells, dtype=torch.int32, device=dev)
    nj = torch.zeros(ncells, dtype=torch.int32, device=dev)
    sunlit = (d["rj"] > 0).any(dim=0)
    solver = Ros3Solver(mech)
    stream = torch.cuda.current_stream().cuda_stream

    def step():
        y.copy_(y0)
        # cells the reference cannot integrate either (DT below DTMIN) are
        # counted, not fatal: n_accepted = -1 marks them
        solver.run_device(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], y, na, nj, stream=stream,
                          raise_on_fail=False)

    for w in range(args.warmup):
        step()
        torch.cuda.synchronize()
        print(f"[bench] warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    solve_ms, rates_ms = [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        step()
        r_ms, s_ms = solver.last_timing()
        rates_ms.append(r_ms)
        solve_ms.append(s_ms)
        print(f"[bench] step {k + 1}/{args.steps}: solve kernel {s_ms:.1f} ms, rates kernel {r_ms:.1f} ms",
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps

    failed = int((na < 0).sum())
    nacc_ok = na.clamp(min=0).to(torch.int64)
    att = nacc_ok + nj.to(torch.int64)
    attempted = int(att.sum())
    accepted = int(nacc_ok.sum())
    nnz_day = solver.query("nnz_day")
    nnz_night = solver.query("nnz_night")
    ns, nphot, nhet = mech["numb_mech_spc"], mech["nphotab"], mech["nhetero"]
    # per cell per sync step I/O: y in + out, T, P, H2O, J, het, seawater
    b_io = 8 * (2 * ns + 3 + nphot + nhet + 1)
    b_step = torch.where(sunlit, 32 * nnz_day, 32 * nnz_night).to(torch.int64)
    alg_bytes = float((att * b_step).sum()) + float(ncells * b_io)
    solve_s = float(np.mean(solve_ms)) / 1000.0
    achieved = alg_bytes / solve_s / 1e9
    per_step, traffic_src = measured_traffic_per_step(args.mech, solver.variant)
    traffic = per_step * attempted if per_step else None
    from cmaq_amd.codegen.programs import flops_per_step
    from cmaq_amd.codegen.symbolic import analyse
    f_step = flops_per_step(analyse(mech))
    fp64_tfs = f_step * attempted / solve_s / 1e12
    value = attempted / (ms_per_step / 1000.0)

    out = {
        "metric": "CB6r5 ROS3 cell-steps/s on 1 GPU vs CPU; fwd+adjoint / fwd time ratio",
        "value": value,
        "unit": "attempted cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.mech} ROS3 forward, {args.workload}, {ncells} cells, "
                               f"one 5-min sync step, RTOL 1e-3, ATOL 1e-7 ppm",
                   "ncells": ncells, "sync_steps": 1, "chemstep_min": 5.0},
        "accepted_cell_steps_per_s": accepted / (ms_per_step / 1000.0),
        "attempted_steps_per_cell": attempted / ncells,
        "failed_cells": failed,
        "solve_kernel_ms": float(np.mean(solve_ms)),
        "rates_kernel_ms": float(np.mean(rates_ms)),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src, "kernel": "solve_kernel", "variant": solver.variant,
                     "bytes_per_launch": alg_bytes},
        "fp64": {"achieved": fp64_tfs, "peak": FP64_PEAK_TFS, "unit": "TFLOP/s",
                 "frac": fp64_tfs / FP64_PEAK_TFS, "flops_per_attempted_step": f_step},
        "fwd_adj_ratio": None,
    }
    if not args.no_chem and args.workload == "regional" and not args.ncells:
        print("[bench] CHEM gather/scatter leg", file=sys.stderr, flush=True)
        del y
        torch.cuda.empty_cache()
        out["chem_io"] = chem_io_leg(args, mech, solver, d, (ncol, nrow, nlay), dev)
        y = None
        torch.cuda.empty_cache()
    if not (args.no_adj and args.no_loop):
        del y, y0, d, na, nj, sunlit
        torch.cuda.empty_cache()
    if not args.no_loop:
        print("[bench] sync-loop leg", file=sys.stderr, flush=True)
        out["sync_loop"] = loop_leg(args, mech, solver, dev)
        torch.cuda.empty_cache()
    if not args.no_adj:
        print("[bench] adjoint leg", file=sys.stderr, flush=True)
        ratio, info = adjoint_leg(args, mech, solver, dev)
        out["fwd_adj_ratio"] = ratio
        out["adjoint"] = info
        torch.cuda.empty_cache()
    if not args.no_big_adj:
        print("[bench] larger-mechanism adjoint leg", file=sys.stderr, flush=True)
        out["adjoint_larger_mechanism"] = big_adjoint_leg(args, dev)
        torch.cuda.empty_cache()
    if not args.no_gear:
        print("[bench] SMVGEAR leg", file=sys.stderr, flush=True)
        try:
            avail = len(os.sched_getaffinity(0))
        except AttributeError:
            avail = os.cpu_count() or 1
        out["smvgear"] = gear_leg(args, mech, dev, args.cpu_workers or min(16, avail))
        torch.cuda.empty_cache()
    wpc = max(1, solver.query("solve_waves") // max(1, torch.cuda.get_device_properties(dev).multi_processor_count))
    ceiling = row_stream_ceiling(wpc)
    if ceiling:
        out["roofline"]["row_stream_ceiling"] = ceiling
        out["roofline"]["row_stream_waves_per_cu"] = wpc
        if traffic:
            out["roofline"]["traffic_frac_of_row_stream_ceiling"] = traffic / solve_s / 1e9 / ceiling
    if not args.no_cpu:
        print("[bench] timing the CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args, mech, ncells, nlay_cells)
        # the reference publishes no figure for this metric, so vs_baseline
        # stays null; this is the ratio to the CPU baseline timed here
        out["gpu_over_cpu_baseline"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
