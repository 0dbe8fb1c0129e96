"""The register-allocated form of the generated functions: the allocator's
loads, stores and register operations, executed in Python against a
workspace whose outputs start poisoned (NaN), must leave in memory exactly
what the plain operation list computes -- for every slot a later function
reads -- and must not write back the factored pivots (they are dead: the
solves read their reciprocals VD)."""
import math
import re

import numpy as np
import pytest

from cmaq_amd.codegen import programs as P
from cmaq_amd.codegen.regalloc import allocate
from cmaq_amd.codegen.symbolic import analyse
from cmaq_amd.inputs import make_cells
from progeval import CONST, compile_ops

_REG = re.compile(r"\br(\d+)\b")


def _run_items(items, W, params):
    lines = ["def _f(W, R, " + ", ".join(params) + "):"]
    for it in items:
        if it[0] == "load":
            lines.append(f"    R[{it[1]}] = W['{it[2][0]}'][{it[2][1]}]")
        elif it[0] == "store":
            lines.append(f"    W['{it[2][0]}'][{it[2][1]}] = R[{it[1]}]")
        else:
            lines.append("    " + _REG.sub(lambda m: f"R[{m.group(1)}]", it[1]).rstrip(";"))
    ns = dict(CONST)
    exec(compile("\n".join(lines) + "\n", "<items>", "exec"), ns)
    return ns["_f"]


@pytest.mark.parametrize("order", ["ctile10x10", "tile10x10", "column"])
def test_allocated_jac_lu_leaves_the_program_result(cb6r5, oracle_cb6r5, order):
    sym = analyse(cb6r5)
    o = oracle_cb6r5
    day = sym.day
    d = make_cells(cb6r5, np.arange(1), seed=2, workload="mixed")
    rki, _ = o.rates(d["temp"], d["pres"], d["h2o"], d["rj"])
    inew = o.inew2old() - 1
    y = list(d["y"][inew, 0])
    rk = list(rki[:, 0])
    f = list(np.asarray(o.feval(1, rki, d["y"][inew])).ravel())
    gdtinv = 1.0 / (0.05 * CONST["GAM"])
    ops = P.jac_lu_ops(sym, order, True)
    ref = {"RK": list(rk), "Y": list(y), "F": list(f), "LU": [0.0] * day.nnz, "VD": [0.0] * sym.ns,
           "K1": [0.0] * sym.ns}
    compile_ops(ops, ["gdtinv"])(ref, gdtinv=gdtinv)
    piv = frozenset(("LU", day.diag[j]) for j in range(sym.ns))
    items, cnt = allocate(ops, {"LU", "VD", "K1"}, 170, 96, lambda s: True, dead=piv)
    nan = math.nan
    W = {"RK": list(rk), "Y": list(y), "F": list(f), "LU": [nan] * day.nnz, "VD": [nan] * sym.ns,
         "K1": [nan] * sym.ns}
    _run_items(items, W, ["gdtinv"])(W, [nan] * 170, gdtinv)
    diag = {day.diag[j] for j in range(sym.ns)}
    for s in range(day.nnz):
        if s not in diag:
            assert W["LU"][s] == ref["LU"][s]
    assert W["VD"] == ref["VD"] and W["K1"] == ref["K1"]
    # a pivot is stored only when it leaves the registers before its last
    # read (the reciprocal), never as a final write-back
    _, cnt_all = allocate(ops, {"LU", "VD", "K1"}, 170, 96, lambda s: True)
    piv_stores = sum(1 for it in items if it[0] == "store" and it[2] in piv)
    assert cnt["store"] + cnt["load"] < cnt_all["store"] + cnt_all["load"] and piv_stores < sym.ns // 4
    print(order, "stores", cnt_all["store"], "->", cnt["store"], "pivot stores left", piv_stores)
