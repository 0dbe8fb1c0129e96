"""The synthetic-input rule for short-lived species, and the failure dump.

cmaq_amd.inputs starts short-lived species at 1e-6 ppb.  These tests
re-derive that list from the rate constants (first-order loss above
FAST_LOSS_PER_MIN anywhere in the sampled T/P range), show on the oracle what
the rule removes (with ROR drawn log-uniform, config-0 cell 1841 fails with
PAR's error term dominating ERR, the state rbsolver.F:532-583 would print),
and check that the failure dump the oracle writes is the failing step's.
"""
import numpy as np
import pytest

import cmaq_amd.inputs as I
from cmaq_amd.build import MECHS
from cmaq_amd.ros3 import load_mechanism

D1, D2, D3 = 0.5, -0.29079558716805469821718236208017e+01, 0.22354069897811569627360909276199e+00
B1, B2, B3 = 1.0, 0.61697947043828245592553615689730e+01, -0.42772256543218573326238373806514e+00


@pytest.mark.parametrize("name", sorted(MECHS))
def test_fast_species_list_from_rate_constants(name):
    from oracle.oracle import Oracle
    m = load_mechanism(name)
    o = Oracle(m)
    spc = m["chemistry_spc"]
    fast = set()
    for T in (250.0, 280.0, 310.0):
        for P in (500.0 / 1013.25, 1013.0 / 1013.25):
            rki, _ = o.rates(np.array([T]), np.array([P]), np.array([1e4]), I.clear_sky_j(m)[:, None])
            freq = np.zeros(len(spc))
            for r in range(m["nrxns"]):
                if m["nreact"][r] == 1:
                    freq[m["irr"][r][0] - 1] += rki[r, 0]
            fast |= {spc[i] for i in np.flatnonzero(freq > I.FAST_LOSS_PER_MIN)}
    listed = I.sub_dtmin_species(m)
    extra = {"ISON"} if m["mechname"] == "CRACMM1AMORE_AQ" else set()
    assert fast == listed - extra, (name, sorted(fast), sorted(listed))
    assert "O1D" in fast


def _one_cell(m, cell, drawn):
    """Config-0 inputs of one cell with the species in `drawn` taken from the
    plain log-uniform draw instead of 1e-6 ppb."""
    d = I.make_cells(m, np.array([cell]), seed=0)
    for nm in drawn:
        i = m["chemistry_spc"].index(nm)
        d["y"][i] = I._logu(I.uniform(0, I.F_SPC + i, np.array([cell])), 1.0e-6, 1.0e2) * 1.0e-3
    return d


def test_drawn_ror_fails_through_par_and_dump_is_the_failing_step(cb6r5, oracle_cb6r5):
    o, m = oracle_cb6r5, cb6r5
    spc = m["chemistry_spc"]
    inew = o.inew2old() - 1
    d = _one_cell(m, 1841, ["ROR"])
    rki, _ = o.rates(d["temp"], d["pres"], d["h2o"], d["rj"])
    y0 = d["y"][inew]
    rt, at = np.full(o.ns, 1e-3), np.full(o.ns, 1e-7)
    rc, y, na, nj, dump = o.rbsolver(1, 5.0, rki, y0, rt, at, dump=True)
    assert rc == -1 and dump is not None
    ns = o.ns
    err, Y, YP = dump[0], dump[1:1 + ns], dump[1 + ns:1 + 2 * ns]
    K1, K2, K3 = (dump[1 + (2 + q) * ns:1 + (3 + q) * ns] for q in range(3))
    # the dump is the rejected step: Y the state the solver returns, YP and
    # ERR recomputed from the stages as rbsolver.F:425-441 does
    assert np.array_equal(Y, y[:, 0])
    assert np.array_equal(YP, ((Y + B1 * K1) + B2 * K2) + B3 * K3)
    e = (D1 * K1 + D2 * K2 + D3 * K3) / (at + rt * np.abs(YP))
    acc = 0.0
    for v in e:
        acc = acc + v * v
    assert acc == err
    share = e * e / err
    top = int(np.argmax(share))
    assert spc[inew[top]] == "PAR" and share[top] > 0.9
    assert YP[top] < 0.0                       # ROR's decay drove PAR negative
    # the same cell with ROR at 1e-6 ppb integrates
    d2 = I.make_cells(m, np.array([1841]), seed=0)
    r2 = o.run(5.0, d2["temp"], d2["pres"], d2["h2o"], d2["rj"], d2["y"], 1e-3, 1e-7)
    assert r2["rc"] == 0 and r2["nacc"][0] > 0


def test_config0_has_no_failed_cell(cb6r5, oracle_cb6r5):
    """4096 config-0 cells (seed 0) all integrate under the input rule."""
    d = I.make_cells(cb6r5, np.arange(4096), seed=0)
    r = oracle_cb6r5.run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"], 1e-3, 1e-7)
    assert r["rc"] == 0 and (r["nacc"] > 0).all()
