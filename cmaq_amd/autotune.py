"""Autotuner for the generated ROS3 kernels.

The knobs of a kernel variant are the register budget (doubles of a lane's
workspace held in VGPRs, which fixes the occupancy), the load lookahead, the
LU schedule (column order, or 2-D tiles visited row- or column-block-wise,
with stage 1's forward substitution fused) and an optional LDS tier; at run
time come the waves-per-CU cap of the persistent grid and the refill
threshold of its cell queue.  One cell per lane is fixed: a cell's working
set (27 KB for CB6r5) is far above a lane's 2 KB of registers.

Three steps, per mechanism:

  propose   on the CPU, every (budget, lookahead, LU schedule, LDS) point of a
            grid is register-allocated and the schedules ranked by the memory
            operations of one attempted step; the best schedules of each
            (budget, lookahead) become candidates;
  build     the candidates are compiled into their own library
            (libcmaq_ros3_tune.so, loaded with CMAQ_ROS3_LIB);
  tune      on the GPU, every candidate x refill threshold x waves cap is
            timed on the regional workload and on the mixed-regime one, and
            the variant with the best geometric mean is written to
            tuning.json as "knobs"; the next build generates it as the
            mechanism's default (build.variants_of), whether or not the
            hand-kept table lists it.

    python -m cmaq_amd.autotune propose --mech cb6r5_ae7_aq
    python -m cmaq_amd.autotune build   --mech cb6r5_ae7_aq
    python -m cmaq_amd.autotune tune    --mech cb6r5_ae7_aq --out cmaq_amd/tuning.json   # GPU
    python -m cmaq_amd.autotune time    --mech cb6r5_ae7_aq   # the built library's variants (GPU)
"""

import argparse
import json
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
BUDGETS = (150, 170)
LOOKAHEADS = (64, 96, 128)
SCHEDULES = ("column", "ctile8x8", "ctile10x10", "ctile12x10", "ctile14x10", "ctile14x14", "tile10x10",
             "tile12x12")


def step_memory_ops(sym, budget, lookahead, order, lds=0):
    """Workspace loads + stores of one attempted step of the forward kernel:
    two f evaluations, Jacobian + LU with stage 1's forward substitution, and
    the three stage solves, as regalloc schedules them."""
    from .codegen import programs as P
    from .codegen.regalloc import allocate
    day = sym.day
    piv = frozenset(("LU", day.diag[j]) for j in range(sym.ns))
    every = lambda s: True
    total = 0
    cnt = allocate(P.feval_ops(sym, "scatter"), {"F"}, budget, lookahead, every, lds)[1]
    total += 2 * (cnt.get("load", 0) + cnt.get("store", 0))
    cnt = allocate(P.jac_lu_ops(sym, order, True), {"LU", "VD", "K1"}, budget, lookahead, every, lds, dead=piv)[1]
    total += cnt.get("load", 0) + cnt.get("store", 0)
    for stage in (1, 2, 3):
        ops = P.solve_ops(sym, stage, backward_only=(stage == 1))
        cnt = allocate(ops, {f"K{stage}"} | ({"YP"} if stage == 1 else set()), budget, lookahead, every, lds)[1]
        total += cnt.get("load", 0) + cnt.get("store", 0)
    return total


def propose(mech_name, budgets=BUDGETS, lookaheads=LOOKAHEADS, schedules=SCHEDULES, per_point=2, lds=(0, 32),
            log=print):
    """Candidate knobs [(budget, lookahead, order)] and the static table."""
    from .build import load_mech
    from .codegen.symbolic import analyse
    sym = analyse(load_mech(mech_name))
    table, cands = [], []
    for b in budgets:
        for la in lookaheads:
            ranked = []
            for sch in schedules:
                ops = step_memory_ops(sym, b, la, sch)
                ranked.append((ops, sch))
                table.append({"budget": b, "lookahead": la, "schedule": sch, "lds": 0, "memory_ops_per_step": ops})
                log(f"[propose] {mech_name} b{b} l{la} {sch}: {ops} memory operations per step")
            ranked.sort()
            for ops, sch in ranked[:per_point]:
                cands.append((b, la, sch + "+k1"))
    # the LDS tier on the statically best point
    best = min(table, key=lambda r: r["memory_ops_per_step"])
    for n in lds:
        if n:
            ops = step_memory_ops(sym, best["budget"], best["lookahead"], best["schedule"], n)
            table.append(dict(best, lds=n, memory_ops_per_step=ops))
            cands.append((best["budget"], best["lookahead"], best["schedule"] + f"+k1+lds{n}"))
    return cands, table


def tune(mech_name, ncells, workloads=("regional", "mixed"), seed=1, refills=(1, 2), log=print, svoc_ppb=None,
         waves_per_cu=(0,), variants=None):
    """Times every built variant (or `variants`) x refill x waves cap on each
    workload; returns the table entry (best geometric mean of the rates)."""
    import torch
    from .inputs import make_cells_torch
    from .ros3 import Ros3Solver, load_mechanism
    mech = load_mechanism(mech_name)
    dev = torch.device("cuda:0")
    base = Ros3Solver(mech)
    names = [v for v in base.variants() if not variants or v in variants]
    base.close()
    rates = {}
    results = []
    for wl in workloads:
        d = make_cells_torch(mech, ncells, seed=seed, workload=wl, nlay_cells=max(1, ncells // 35),
                             svoc_ppb=svoc_ppb, device=dev)
        y = torch.empty_like(d["y"])
        na = torch.zeros(ncells, dtype=torch.int32, device=dev)
        nj = torch.zeros(ncells, dtype=torch.int32, device=dev)
        ref_y = None
        for var, rf, wpc in [(v, r, w) for v in names for r in refills for w in waves_per_cu]:
            s = Ros3Solver(mech, refill_min=rf, waves_per_cu=wpc)
            s.set_variant(var)
            best = None
            for _ in range(2):
                y.copy_(d["y"])
                s.run_device(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], y, na, nj, raise_on_fail=False)
                ms = s.last_timing()[1]
                best = ms if best is None else min(best, ms)
            # every candidate must give the same result, bit for bit
            if ref_y is None:
                ref_y = y.clone()
            identical = bool(torch.equal(y, ref_y))
            att = int((na.clamp(min=0) + nj).sum())
            rate = att / (best / 1000.0)
            rates.setdefault((var, rf, wpc), []).append(rate)
            results.append({"variant": var, "refill_min": rf, "waves_per_cu": wpc, "workload": wl, "solve_ms": best,
                            "cell_steps_per_s": rate, "bit_identical": identical})
            log(f"[autotune] {mech_name} {wl} {var} refill {rf} waves/CU {wpc or 'max'}: {best:.1f} ms, "
                f"{rate / 1e6:.2f} M cell-steps/s{'' if identical else ', RESULT DIFFERS'}")
            s.close()
            if not identical:
                raise RuntimeError(f"{var} does not reproduce the first variant's result")
        del d, y
        torch.cuda.empty_cache()
    score = {k: math.exp(sum(math.log(r) for r in v) / len(v)) for k, v in rates.items() if len(v) == len(workloads)}
    (var, rf, wpc), sc = max(score.items(), key=lambda kv: kv[1])
    return {"variant": var, "refill_min": rf, "waves_per_cu": wpc, "ncells": ncells, "workload": list(workloads),
            "geomean_cell_steps_per_s": sc, "results": results}


def _knobs_of(variant, cands):
    from .build import variant_name
    for b, la, order in cands:
        if variant_name(b, la, order) == variant:
            return [b, la, order]
    return None


def main():
    p = argparse.ArgumentParser()
    p.add_argument("step", choices=["propose", "build", "tune", "time"])
    p.add_argument("--mech", action="append", default=None)
    p.add_argument("--ncells", type=int, default=1048576)
    p.add_argument("--out", default=os.path.join(HERE, "tuning.json"))
    p.add_argument("--workload", nargs="*", default=["regional", "mixed"])
    p.add_argument("--svoc", action="store_true", help="VOC/SVOC species log-uniform on [1e-5, 5] ppb")
    p.add_argument("--refill", type=int, nargs="*", default=[1, 2])
    p.add_argument("--variant", nargs="*", default=None, help="time only these built variants")
    p.add_argument("--waves-per-cu", type=int, nargs="*", default=[0],
                   help="resident-wave caps to sweep (0 = as many as the registers allow)")
    a = p.parse_args()
    mechs = a.mech or ["cb6r5_ae7_aq"]
    log = lambda s: print(s, file=sys.stderr, flush=True)
    cand_file = os.path.join(HERE, "csrc", "gen_tune", "candidates.json")
    if a.step == "propose":
        out = {}
        for m in mechs:
            cands, table = propose(m, log=log)
            out[m] = {"candidates": cands, "static": table}
        os.makedirs(os.path.dirname(cand_file), exist_ok=True)
        with open(cand_file, "w") as f:
            json.dump(out, f, indent=1)
        print(json.dumps({m: v["candidates"] for m, v in out.items()}))
        return
    if a.step == "build":
        from . import build as B
        with open(cand_file) as f:
            props = json.load(f)
        variants = {m: [tuple(c) for c in props[m]["candidates"]] for m in mechs}
        # the current default joins as the reference of the comparison
        for m in mechs:
            dflt = next(v for v in B.variants_of(m) if B.variant_name(*v) == B.tuned(m)[0])
            if dflt not in variants[m]:
                variants[m].insert(0, dflt)
        B.build(mechs, variants=variants, lib=B.TUNE_LIB, gen=B.TUNE_GEN, obj=os.path.join(B.TUNE_GEN, "obj"))
        return
    table = {}
    if os.path.exists(a.out):
        with open(a.out) as f:
            table = json.load(f)
    if a.step == "time":
        out = {m: tune(m, a.ncells, workloads=tuple(a.workload), refills=tuple(a.refill),
                       waves_per_cu=tuple(a.waves_per_cu), variants=a.variant, log=log,
                       svoc_ppb=(1e-5, 5.0) if a.svoc else None) for m in mechs}
        print(json.dumps({m: {k: v for k, v in t.items() if k != "results"} for m, t in out.items()}))
        return
    # tune: the candidate library's variants
    from . import build as B
    with open(cand_file) as f:
        props = json.load(f)
    for m in mechs:
        entry = tune(m, a.ncells, workloads=tuple(a.workload), refills=tuple(a.refill),
                     waves_per_cu=tuple(a.waves_per_cu), variants=a.variant, log=log,
                     svoc_ppb=(1e-5, 5.0) if a.svoc else None)
        cands = [tuple(c) for c in props[m]["candidates"]] + list(B.variants_of(m))
        entry["knobs"] = _knobs_of(entry["variant"], cands)
        entry["static"] = props[m]["static"]
        table[m] = entry
        with open(a.out, "w") as f:
            json.dump(table, f, indent=1)
    print(json.dumps({m: {k: v for k, v in t.items() if k not in ("results", "static")} for m, t in table.items()}))


if __name__ == "__main__":
    main()
