"""The oracle and the product's symbolic analysis on every distinct chemmech
mechanism of the CMAQ tree beyond CB6r5 and CRACMM2 (CCTM/src/MECHS/*).
cb6r5_ae7_aqkmt2 and cb6r5hap_ae7_aq carry the same RXNS_DATA and
RXNS_FUNC modules as cb6r5_ae7_aq, saprc07tic_ae7i_aqkmt2 the same as
saprc07tic_ae7i_aq, so they share those mechanisms' tables and kernels.

Each check is independent of the oracle's code, as in test_oracle.py:
rate constants bit for bit against the reference's generated CALC_RCONST
(tests/golden/make_rate_golden.py), the stored LU pattern against a
brute-force boolean elimination, f against a dense stoichiometric product,
J against central differences of f, LU and solves against numpy.
"""
import collections
import json
import os

import numpy as np
import pytest

from cmaq_amd.inputs import make_cells

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
GAM = 0.43586652150845899941601945119356

# N, R (NUMB_MECH_SPC, NRXNS in RXNS_DATA_MODULE.F90) and the KTYPE histogram
TABLES = {
    "cb6r3_ae7_aq": (147, 343, {-1: 17, 0: 41, 1: 96, 2: 2, 3: 134, 4: 3, 6: 28, 8: 1, 9: 3, 10: 17, 12: 1}),
    "cb6r5m_ae7_aq": (182, 453, {-1: 39, 0: 68, 1: 107, 2: 2, 3: 172, 4: 3, 6: 34, 8: 1, 9: 3, 10: 24}),
    "cracmm1_aq": (178, 508, {-1: 7, 0: 39, 1: 201, 2: 1, 3: 237, 4: 2, 5: 4, 8: 1, 9: 4, 10: 11, 12: 1}),
    "cracmm1amore_aq": (181, 516, {-1: 7, 0: 40, 1: 203, 2: 1, 3: 242, 4: 2, 5: 4, 8: 1, 9: 4, 10: 11,
                                   12: 1}),
    "racm2_ae6_aq": (168, 411, {-1: 5, 0: 33, 1: 131, 2: 1, 3: 218, 4: 2, 5: 4, 8: 1, 9: 4, 10: 11, 12: 1}),
    "saprc07tc_ae6_aq": (198, 752, {-1: 11, 0: 55, 1: 122, 2: 4, 3: 112, 4: 10, 6: 414, 8: 1, 9: 3, 10: 19,
                                    12: 1}),
    "saprc07tic_ae7i_aq": (230, 929, {-1: 24, 0: 65, 1: 158, 2: 4, 3: 162, 4: 10, 6: 482, 8: 1, 9: 3,
                                      10: 19, 12: 1}),
}
NAMES = list(TABLES)


@pytest.fixture(scope="module", params=NAMES)
def mo(request):
    from cmaq_amd.ros3 import load_mechanism
    from oracle.oracle import Oracle
    mech = load_mechanism(request.param)
    return request.param, mech, Oracle(mech)


def _dense(o, ncsp, cc2col):
    rows, cols = o.pattern(ncsp)
    a = np.zeros((o.ischan, o.ischan))
    a[rows - 1, cols - 1] = cc2col[1:len(rows) + 1]
    return a


def _rates_per_cell(o, d):
    """CALC_RCONST per cell (the oracle decides day/night per call)."""
    n = d["temp"].size
    out = None
    for c in range(n):
        r, _ = o.rates(d["temp"][c:c + 1], d["pres"][c:c + 1], d["h2o"][c:c + 1], d["rj"][:, c:c + 1])
        if out is None:
            out = np.empty((r.shape[0], n))
        out[:, c] = r[:, 0]
    return out


def test_tables(mo):
    name, mech, _ = mo
    ns, nr, hist = TABLES[name]
    assert mech["numb_mech_spc"] == ns and mech["nrxns"] == nr
    assert collections.Counter(mech["ktype"]) == hist
    assert len(mech["irr"]) == nr and all(len(r) == mech["mxprd"] + 3 for r in mech["irr"])


def test_rates_bit_exact_vs_generated_code(mo):
    name, mech, o = mo
    g = json.load(open(os.path.join(GOLDEN, f"rates_{name}.json")))
    exp = np.array([[float.fromhex(v) for v in row] for row in g["rki_hex"]])
    rj, het, sea = np.array(g["rj"]), np.array(g["het"]), np.array(g["seawater"])
    for c in range(g["ncells"]):
        r, _ = o.rates(np.array(g["temp"][c:c + 1]), np.array(g["pres_atm"][c:c + 1]),
                       np.array(g["h2o_ppm"][c:c + 1]), rj[:, c:c + 1], het[:, c:c + 1], sea[c:c + 1])
        assert np.array_equal(r[:, 0], exp[:, c]), f"{name} cell {c}"
    lit = rj.max(axis=0) > 0
    assert (lit & (sea > 0.001)).any() and (~lit & (sea > 0.001)).any() and (sea == 0).any()


def test_symbolic_product_matches_oracle(mo):
    from cmaq_amd.codegen.symbolic import analyse
    name, mech, o = mo
    s = analyse(mech)
    assert s.ischan == o.ischan == mech["numb_mech_spc"]
    assert list(o.inew2old() - 1) == s.inew2old
    for ncsp, p in ((1, s.day), (2, s.night)):
        assert list(o.nkuserat(ncsp) - 1) == p.active
        rows, cols = o.pattern(ncsp)
        assert list(rows - 1) == p.rows and list(cols - 1) == p.cols
        ij, ik, kj = o.lu_ops(ncsp)
        assert [(a - 1, b - 1, c - 1) for a, b, c in zip(ij, ik, kj)] == p.lu_ops


def test_fill_is_dense_elimination(mo):
    name, mech, o = mo
    rows, cols = o.pattern(1)
    n = o.ischan
    iold2new = np.argsort(o.inew2old() - 1)
    a = np.eye(n, dtype=bool)
    for r in o.nkuserat(1) - 1:
        irr = mech["irr"][r]
        re = [iold2new[s - 1] for s in irr[:mech["nreact"][r]]]
        pr = [iold2new[s - 1] for s in irr[3:3 + mech["nprdct"][r]]]
        for c in re:
            a[re, c] = True
            a[pr, c] = True
    for k in range(n):
        lower = np.nonzero(a[k + 1:, k])[0] + k + 1
        upper = np.nonzero(a[k, k + 1:])[0] + k + 1
        a[np.ix_(lower, upper)] = True
    assert set(zip(rows - 1, cols - 1)) == set(zip(*np.nonzero(a)))


def test_feval_jacobian_lu(mo):
    """f = S r (dense), J = central differences of f, LU and solve = numpy."""
    name, mech, o = mo
    d = make_cells(mech, np.arange(2), seed=31)
    rki = _rates_per_cell(o, d)
    inew2old = o.inew2old() - 1
    iold2new = np.argsort(inew2old)
    y = np.ascontiguousarray(d["y"][inew2old])
    n = o.ischan
    # f
    f = np.zeros_like(y)
    for r in o.nkuserat(1) - 1:
        irr = mech["irr"][r]
        re = [iold2new[s - 1] for s in irr[:mech["nreact"][r]]]
        rate = rki[r].copy()
        for s in re:
            rate = rate * y[s]
        for s in re:
            f[s] -= rate
        for k in range(mech["nprdct"][r]):
            f[iold2new[irr[3 + k] - 1]] += mech["sc"][r][k] * rate
    got = o.feval(1, rki, y)
    assert np.abs(got - f).max() <= 1e-12 * np.abs(f).max()
    # J, one cell
    yc, rc = y[:, :1], rki[:, :1]
    jac = _dense(o, 1, o.jacob(1, rc, yc)[:, 0])
    fd = np.zeros((n, n))
    for j in range(n):
        # f is at most quadratic in each y_j, so the difference is exact for any h
        h = 1e-2 + yc[j, 0]
        yp, ym = yc.copy(), yc.copy()
        yp[j] += h
        ym[j] -= h
        fd[:, j] = (o.feval(1, rc, yp)[:, 0] - o.feval(1, rc, ym)[:, 0]) / (2 * h)
    tol = 1e-9 * np.abs(jac).max(axis=1, keepdims=True) + 1e-9 * np.abs(fd)
    assert (np.abs(jac - fd) <= tol + 1e-30).all()
    assert (np.abs(fd)[jac == 0] <= 1e-6 * np.abs(jac).max()).all()
    # LU + solve on M = J - I/(gamma h)
    hstep = 0.01
    cc2 = o.jacob(1, rki, y)
    rows, cols = o.pattern(1)
    cc2[np.nonzero(rows == cols)[0] + 1] -= 1.0 / (GAM * hstep)
    lu, vdiag = o.decomp(1, cc2)
    rhs = np.random.default_rng(0).normal(size=(n, 2))
    x = o.solve(1, lu, vdiag, rhs)
    for c in range(2):
        m = _dense(o, 1, cc2[:, c])
        want = np.linalg.solve(m, rhs[:, c])
        assert np.abs(x[:, c] - want).max() <= 1e-10 * np.abs(want).max()


def test_clean_cells_integrate(mo):
    """Every clean cell integrates one 5-min sync step with the reference's
    tolerances (the workload generator starts sub-DTMIN species near zero)."""
    name, mech, o = mo
    d = make_cells(mech, np.arange(128), seed=0)
    r = o.run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"], 1e-3, 1e-7)
    assert r["rc"] == 0 and (r["nacc"] > 0).all()
    assert np.isfinite(r["y"]).all() and (r["y"] >= 1e-30).all()


def test_forward_programs_bit_exact(mo):
    """The generated operation lists of the schedule the library is built
    with reproduce the oracle's per-cell run bit for bit (same per-entry
    rounding order), as test_programs.py checks for CB6r5."""
    from cmaq_amd.build import VARIANTS
    from cmaq_amd.codegen.symbolic import analyse
    from progeval import Programs
    name, mech, o = mo
    order = VARIANTS[name][0][2].split("+")
    pg = Programs(analyse(mech), order[0], "k1" in order[1:])
    inew = o.inew2old() - 1
    for seed in (1, 4):
        d = make_cells(mech, np.arange(1), seed=seed, workload="mixed")
        rki = _rates_per_cell(o, d)[:, 0].copy()
        y0 = d["y"][:, 0].copy()
        h, yend = o.steps_cell(5.0, rki, y0)
        W = pg.workspace(rki, y0[inew])
        for hn in h:
            pg.stages(W, hn)
            W["Y"] = pg.ynew(W)
        assert np.array_equal(np.array(W["Y"]), yend[inew]), f"{name} seed {seed}"
