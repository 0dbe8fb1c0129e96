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
