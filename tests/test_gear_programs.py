"""SMVGEAR's generated operation lists (cmaq_amd/codegen/programs.py:
gear_subfun_ops, gear_pderiv_lu_ops, msolve_ops), evaluated in Python double
precision, against the oracle's grsubfun / grpderiv + grdecomp / grbacksub
restatement: bit for bit, for every LU schedule, with species at and below
ZBOUND in the state (the CEFF and GLOSS clamps)."""
import numpy as np
import pytest

from cmaq_amd.codegen import programs as P
from cmaq_amd.codegen.symbolic import analyse
from cmaq_amd.inputs import make_cells
from progeval import compile_ops


@pytest.mark.parametrize("order", ["column", "ctile10x10", "tile10x10", "ctile14x10"])
def test_gear_programs_bit_exact(cb6r5, oracle_cb6r5, order):
    o = oracle_cb6r5
    sym = analyse(cb6r5)
    day = sym.day
    subfun = compile_ops(P.gear_subfun_ops(sym))
    pdlu = compile_ops(P.gear_pderiv_lu_ops(sym, order), ["r1delt"])
    back = compile_ops(P.msolve_ops(sym, "GL"))
    inew = o.inew2old() - 1
    d = make_cells(cb6r5, np.arange(3), seed=7, workload="mixed")
    rki, _ = o.rates(d["temp"], d["pres"], d["h2o"], d["rj"])
    for c in range(3):
        cn = np.maximum(d["y"][inew, c], 1e-30)
        cn[[3, 40, 120]] = [0.0, 1e-30, -1e-12]        # CEFF / clamp cases
        W = {"RK": list(map(float, rki[:, c])), "CN": list(map(float, cn)), "GL": [0.0] * sym.ns,
             "CF": [0.0 if v <= 1e-30 else float(v) for v in cn],       # CEFF (grpderiv.F:121-129)
             "RATE": [0.0] * sym.nr, "LU": [0.0] * day.nnz, "VD": [0.0] * sym.ns}
        subfun(W)
        g_ref = o.subfun(1, rki[:, c:c + 1], cn[:, None])[:, 0]
        assert np.array_equal(np.array(W["GL"]), g_ref)
        r1delt = -0.43 * (c + 1)
        pdlu(W, r1delt)
        cc2 = o.pderiv(1, rki[:, c:c + 1], cn[:, None], r1delt)
        cc2, vd = o.decomp(1, cc2)
        # the generated LU keeps every entry except the pivots, which only
        # feed VD (the solves read VD)
        piv = set(day.diag)
        lu = np.array(W["LU"])
        keep = np.array([s not in piv for s in range(day.nnz)])
        assert np.array_equal(lu[keep], cc2[1:day.nnz + 1, 0][keep])
        assert np.array_equal(np.array(W["VD"]), vd[:, 0])
        rhs = np.sin(np.arange(sym.ns) + c) * 1e-3
        W["GL"] = list(map(float, rhs))
        back(W)
        assert np.array_equal(np.array(W["GL"]), o.solve(1, cc2, vd, rhs[:, None])[:, 0])


def test_handoff_record_holds_every_state_variable():
    """A cell moved from gear_kernel to gear_tail_kernel carries its control
    state in GearLaneState, filled and read through the GEAR_LANE_* X-macro
    lists (cmaq_amd/csrc/ros3_gear.h).  Every variable gear_kernel declares
    for a lane's state machine is in those lists, apart from the queue's
    bookkeeping (which cell, whether the lane is idle, the queue drained,
    the hand-off refused) and the wave's PDERIV wait counter."""
    import os
    import re
    csrc = os.path.join(os.path.dirname(os.path.abspath(__file__)), os.pardir, "cmaq_amd", "csrc")
    header = open(os.path.join(csrc, "ros3_gear.h")).read()
    listed = set()
    for lst in ("GEAR_LANE_INTS", "GEAR_LANE_BOOLS", "GEAR_LANE_DBLS"):
        m = re.search(r"#define " + lst + r"\(X\)((?:[^\n]*\\\n)*[^\n]*)", header)
        listed |= set(re.findall(r"X\((\w+)\)", m.group(1)))
    kernel = open(os.path.join(csrc, "ros3_gear_kernel.inc")).read()
    body = kernel[kernel.index("void gear_kernel(GearParams P)"):]
    body = body[body.index("const int64_t n = P.ncells;"):body.index("for (;;)")]
    declared = set()
    for decl in re.findall(r"^\s*(?:bool|int|int64_t|double)\s+([^;]+);", body, re.M):
        for part in decl.split(","):
            declared.add(part.split("=")[0].strip())
    bookkeeping = {"cell", "need", "drained", "stay", "pd_age"}
    assert declared - bookkeeping <= listed, sorted(declared - bookkeeping - listed)
    assert len(listed) >= 37 and listed <= declared, sorted(listed - declared)
