"""Autotuner for the generated ROS3 kernels.

For one mechanism, times every built code-generation variant (register budget
and load lookahead, which fix the occupancy) and a few refill thresholds of
the persistent kernel on a synthetic workload resident in HBM, and records
the fastest in a JSON table (cmaq_amd/tuning.json); the build marks that
variant as the default.

    python -m cmaq_amd.autotune --mech cb6r5_ae7_aq --ncells 1048576 --out tuning.json
"""

import argparse
import json
import os
import sys


def tune(mech_name, ncells, workload="regional", seed=1, refills=(1, 2, 4), log=print, svoc_ppb=None,
         waves_per_cu=(0,), variants=None):
    import torch
    from .inputs import make_cells_torch
    from .ros3 import Ros3Solver, load_mechanism
    mech = load_mechanism(mech_name)
    dev = torch.device("cuda:0")
    d = make_cells_torch(mech, ncells, seed=seed, workload=workload,
                         nlay_cells=max(1, ncells // 35), svoc_ppb=svoc_ppb, device=dev)
    y = torch.empty_like(d["y"])
    na = torch.zeros(ncells, dtype=torch.int32, device=dev)
    nj = torch.zeros(ncells, dtype=torch.int32, device=dev)
    results = []
    base = Ros3Solver(mech)
    names = [v for v in base.variants() if not variants or v in variants]
    for var, rf, wpc in [(v, r, w) for v in names for r in refills for w in waves_per_cu]:
        s = Ros3Solver(mech, refill_min=rf, waves_per_cu=wpc)
        s.set_variant(var)
        best = None
        for _ in range(2):
            y.copy_(d["y"])
            s.run_device(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], y, na, nj, raise_on_fail=False)
            ms = s.last_timing()[1]
            best = ms if best is None else min(best, ms)
        att = int((na.clamp(min=0) + nj).sum())
        rate = att / (best / 1000.0)
        results.append({"variant": var, "refill_min": rf, "waves_per_cu": wpc, "solve_ms": best,
                        "cell_steps_per_s": rate})
        log(f"[autotune] {mech_name} {var} refill {rf} waves/CU {wpc or 'max'}: {best:.1f} ms, "
            f"{rate / 1e6:.2f} M cell-steps/s")
        s.close()
    best = max(results, key=lambda r: r["cell_steps_per_s"])
    return {"variant": best["variant"], "refill_min": best["refill_min"], "waves_per_cu": best["waves_per_cu"],
            "ncells": ncells,
            "workload": workload, "results": results}


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--mech", action="append", default=None)
    p.add_argument("--ncells", type=int, default=1048576)
    p.add_argument("--out", default="tuning.json")
    p.add_argument("--workload", default="regional", choices=["regional", "clean", "mixed"])
    p.add_argument("--svoc", action="store_true", help="VOC/SVOC species log-uniform on [1e-5, 5] ppb")
    p.add_argument("--refill", type=int, nargs="*", default=[1, 2, 4], help="refill thresholds to sweep")
    p.add_argument("--variant", nargs="*", default=None, help="time only these built variants")
    p.add_argument("--waves-per-cu", type=int, nargs="*", default=[0],
                   help="resident-wave caps to sweep (0 = as many as the registers allow)")
    a = p.parse_args()
    table = {}
    if os.path.exists(a.out):
        with open(a.out) as f:
            table = json.load(f)
    for m in a.mech or ["cb6r5_ae7_aq"]:
        table[m] = tune(m, a.ncells, workload=a.workload, svoc_ppb=(1e-5, 5.0) if a.svoc else None,
                        refills=tuple(a.refill), waves_per_cu=tuple(a.waves_per_cu), variants=a.variant,
                        log=lambda s: print(s, file=sys.stderr, flush=True))
        with open(a.out, "w") as f:
            json.dump(table, f, indent=1)
    print(json.dumps({m: {k: v for k, v in t.items() if k != "results"} for m, t in table.items()}))


if __name__ == "__main__":
    main()
