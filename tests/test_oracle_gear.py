"""Pins the SMVGEAR restatement of the CPU oracle (oracle/gear_oracle.c)
before it checks the GPU Gear path.  The checks are independent of the
oracle's own code:

* ASET: the Nordsieck coefficients of the fixed-step BDF of order q are the
  coefficients of prod_{i=1..q} (1 + x / i), normalised so that l_1 = 1;
* PERTST columns 1 and 2 (the error-test constants of orders q and q+1) are
  (q+1) / l_0(q) and (q+2) / l_0(q), to the four digits the table holds;
* the predictor matrix is I + r1delt * J with J the (separately pinned) ROS3
  Jacobian restatement, and SUBFUN is RBFEVAL plus the ZBOUND clamp;
* OPTIMAL_ATOL_PPM picks GEAR_MAX_ATOL when every species is above it and
  GEAR_MIN_ATOL when every species is between the floor and it, and never
  lowers the tolerance when concentrations grow;
* a hand-derived step sequence with zero rates (first step from the
  LSODES formula, two steps before the first order/step test, growth by
  RDELMAX = 10, the last step clipped to the sync step);
* A -> B decay: exact conservation of A + B (BDF keeps linear invariants) and
  the error of A within the tolerance the controller targets;
* whole CB6r5 cells against scipy's Radau at tight tolerance.
"""
import numpy as np
import pytest

from cmaq_amd.inputs import make_cells

from test_oracle import _decay_mech, _dense, _state


def test_aset_is_bdf_nordsieck_vector():
    from oracle.oracle import Oracle, lib
    aset, pertst, enqq, conp = (np.zeros(n) for n in (42, 21, 21, 14))
    p = lambda a: a.ctypes.data_as(__import__("ctypes").POINTER(__import__("ctypes").c_double))
    lib().om_gear_coefficients(p(aset), p(pertst), p(enqq), p(conp))
    aset, pertst = aset.reshape(6, 7), pertst.reshape(7, 3)
    for q in range(1, 7):
        poly = np.array([1.0])
        for i in range(1, q + 1):
            poly = np.convolve(poly, [1.0, 1.0 / i])     # ascending powers of x
        lvec = poly / poly[1]
        assert np.allclose(aset[q - 1, :q + 1], lvec, rtol=1e-15, atol=0), (q, aset[q - 1], lvec)
        assert np.all(aset[q - 1, q + 1:] == 0.0)
        l0 = lvec[0]
        # squared in the table (grsprse.F:687-689)
        assert abs(np.sqrt(pertst[q - 1, 0]) - (q + 1) / l0) <= 5e-4 * (q + 1) / l0
        if q <= 5:
            assert abs(np.sqrt(pertst[q - 1, 1]) - (q + 2) / l0) <= 5e-4 * (q + 2) / l0
    # ENQQ1..3 = 0.5 / q, 0.5 / (q+1), 0.5 / (q+2) in REAL(4)
    enqq = enqq.reshape(7, 3)
    for q in range(1, 8):
        assert list(enqq[q - 1]) == [float(np.float32(0.5) / np.float32(q + k)) for k in range(3)]


def test_pderiv_is_identity_plus_scaled_jacobian(cb6r5, oracle_cb6r5):
    o = oracle_cb6r5
    d, rki, y = _state(cb6r5, o, seed=3, n=3)
    r1delt = -0.37
    for ncsp in (1, 2):
        p = o.pderiv(ncsp, rki, y, r1delt)
        j = o.jacob(ncsp, rki, y)
        for c in range(y.shape[1]):
            want = np.eye(o.ischan) + r1delt * _dense(o, ncsp, j[:, c])
            got = _dense(o, ncsp, p[:, c])
            assert np.allclose(got, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())


def test_subfun_is_feval_with_zbound_clamp(cb6r5, oracle_cb6r5):
    o = oracle_cb6r5
    d, rki, y = _state(cb6r5, o, seed=4, n=3)
    g = o.subfun(1, rki, y)
    assert np.array_equal(g, o.feval(1, rki, y))
    # a species at zero keeps no loss: its negative rate of change is cut to 0
    y2 = y.copy()
    f = o.feval(1, rki, y2)
    i = int(np.flatnonzero(f[:, 0] < 0)[0])
    y2[i, 0] = 0.0
    f2, g2 = o.feval(1, rki, y2), o.subfun(1, rki, y2)
    assert g2[i, 0] == (f2[i, 0] if f2[i, 0] >= 0 else 0.0)
    mask = np.ones_like(g2, bool)
    mask[i, 0] = False
    assert np.array_equal(g2[mask], f2[mask])


def test_optimal_atol_levels(cb6r5, oracle_cb6r5):
    o = oracle_cb6r5
    ns = o.ns
    hi, lo = float(np.float32(1e-7)), float(np.float32(1e-9))
    y = np.full((ns, 1), 1e-3)
    assert o.optimal_atol(y)[0] == pytest.approx(hi, rel=1e-12)
    y = np.full((ns, 1), 5e-11)          # above the 1e-12 floor, below every level
    assert o.optimal_atol(y)[0] == pytest.approx(lo, rel=1e-12)
    rng = np.random.default_rng(0)
    y = 10.0 ** rng.uniform(-14, -5, size=(ns, 64))
    a1 = o.optimal_atol(y)
    a2 = o.optimal_atol(y * 10.0)
    assert np.all(a2 >= a1) and np.all((a1 >= lo * (1 - 1e-12)) & (a1 <= hi * (1 + 1e-12)))


def test_zero_rate_step_sequence():
    """f = 0, night (HMAX 5 min, H0FAC 1): DELT0 = sqrt(EPSLOW / ABST2) =
    sqrt(1e-3 * 25) = 0.158 min; two steps at DELT0 (IDOUB = 2), then the
    order/step test allows RDELMAX = 10: 1.58 min twice (IDOUB = 2 again),
    then the last 1.52 min clipped to the sync step -- 5 steps."""
    from oracle.oracle import Oracle
    o = Oracle(_decay_mech(0.0))
    r = o.gear_run(5.0, np.full(1, 298.0), np.full(1, 1.0), np.full(1, 1e4), np.zeros((0, 1)),
                   np.array([[1.0], [0.0]]), rki=np.zeros((1, 1)))
    assert r["rc"] == 0 and r["nsteps"][0] == 5 and r["nfail"][0] == 0
    assert r["y"][0, 0] == 1.0


@pytest.mark.parametrize("k_per_min", [0.05, 1.0, 30.0])
def test_decay_conserves_and_meets_tolerance(k_per_min):
    from oracle.oracle import Oracle
    o = Oracle(_decay_mech(k_per_min / 60.0))
    y0 = np.array([[0.05], [1e-4]])
    r = o.gear_run(5.0, np.full(1, 298.0), np.full(1, 1.0), np.full(1, 1e4), np.zeros((0, 1)), y0,
                   rki=np.full((1, 1), k_per_min))
    assert r["rc"] == 0
    a, b = r["y"][:, 0]
    total = y0.sum()
    exact = 0.05 * np.exp(-k_per_min * 5.0)
    if k_per_min > 5.0:
        # A decays through zero; the clamp of CONC <= CONCMIN to ZBOUND with its
        # derivative columns zeroed (grsmvgear.F:815-827) ends the invariant
        assert a <= 1e-7 and abs(b - total) <= 1e-2 * total
        return
    assert abs((a + b) - total) <= 1e-14 * total         # linear invariant kept
    # local error per step held near RTOL * (|c| + ATOL); the global error of
    # a few dozen steps stays within a few percent of that scale
    assert abs(a - exact) <= 2e-2 * exact + 1e-7, (a, exact)


def test_integration_vs_radau(cb6r5, oracle_cb6r5):
    """Whole SMVGEAR integration of CB6r5 cells against scipy's Radau."""
    from scipy.integrate import solve_ivp
    o = oracle_cb6r5
    d, rki, y = _state(cb6r5, o, seed=25, n=3)
    for c in range(3):
        sun = bool((d["rj"][:, c] > 0).any())
        ncsp = 1 if sun else 2
        k = rki[:, c:c + 1]
        rc, yg, st = o.smvgear(ncsp, 5.0, k, y[:, c:c + 1], sun)
        assert rc == 0 and st["nsteps"] > 0
        f = lambda t, v: o.feval(ncsp, k, v[:, None])[:, 0]
        jac = lambda t, v: _dense(o, ncsp, o.jacob(ncsp, k, v[:, None])[:, 0])
        ref = solve_ivp(f, (0.0, 5.0), np.maximum(y[:, c], 1e-30), method="Radau", rtol=1e-11, atol=1e-18,
                        jac=jac)
        assert ref.success
        yref = ref.y[:, -1]
        big = yref > 1e-5
        rel = np.abs(yg[big, 0] - yref[big]) / yref[big]
        assert rel.max() < 3e-2, rel.max()


def test_block_of_one_equals_single_cell_calls(cb6r5, oracle_cb6r5):
    """Per-cell control (blocks of one) equals one SMVGEAR call per cell, and
    50-cell blocks run the reference's block-shared control."""
    o = oracle_cb6r5
    d = make_cells(cb6r5, np.arange(60), seed=2)
    r1 = o.gear_run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"], blksize=1)
    assert r1["rc"] == 0
    rk, _ = o.rates(d["temp"][:1], d["pres"][:1], d["h2o"][:1], d["rj"][:, :1])
    inew = o.inew2old() - 1
    rc, yc, st = o.smvgear(1 if (d["rj"][:, 0] > 0).any() else 2, 5.0, rk,
                           np.maximum(d["y"][inew, :1], 1e-30), bool((d["rj"][:, 0] > 0).any()))
    assert rc == 0 and st["nsteps"] == r1["nsteps"][0]
    assert np.array_equal(yc[:, 0], r1["y"][inew, 0])
    r50 = o.gear_run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"], blksize=50)
    assert r50["rc"] == 0
    assert len(set(r50["nsteps"][:30].tolist())) == 1      # one step sequence per block


def test_night_pattern_equals_day_pattern(cb6r5, oracle_cb6r5):
    """With J = 0 the day pattern's extra entries carry exact zeros, so a
    night cell integrated on the day pattern (as the GPU does for every cell)
    equals the reference's night pattern (NCSP = 2) bit for bit."""
    o = oracle_cb6r5
    d, rki, y = _state(cb6r5, o, seed=12, n=16)
    night = np.flatnonzero(~(d["rj"] > 0).any(axis=0))
    assert night.size > 0
    for c in night[:3]:
        r1 = o.smvgear(1, 5.0, rki[:, c:c + 1], np.maximum(y[:, c:c + 1], 1e-30), False)
        r2 = o.smvgear(2, 5.0, rki[:, c:c + 1], np.maximum(y[:, c:c + 1], 1e-30), False)
        assert r1[0] == r2[0] == 0
        assert np.array_equal(r1[1], r2[1]) and r1[2] == r2[2]


def gear_bar(y_ref, y, steps_ref, steps, rtol=1.0e-3, atol=1.0e-9, rates_differ=False):
    """SMVGEAR parity bar for results that may differ in the last bit of a
    pow() of the step control or, with rates_differ, of the rate constants.

    Same rate constants: cells whose step count matches are within
    0.1 * (RTOL |c| + ATOL); the rest -- cells where the controller's discrete
    choices (order, Jacobian reuse, corrector convergence) went another way
    after a last-bit difference -- within 100 * (RTOL |c| + ATOL); and the
    step counts match in >= 99 % of cells (test_reference_sensitivity_to_pow_ulps).

    Rate constants that differ in the last bit (another libm's exp and pow in
    CALC_RCONST): the step counts match in >= 98 % of cells, no more than 2 %
    of cells differ by more than 0.1 * (RTOL |c| + ATOL) and none by more than
    1000 x -- the reference differs from itself by up to ~450 x, also in cells
    whose step count does not change, when its rate constants move by one ulp
    (test_reference_sensitivity_to_rate_ulps).  Returns (ok, summary)."""
    q = np.abs(y - y_ref) / (rtol * np.abs(y_ref) + atol)
    same = steps == steps_ref
    close = (q <= 0.1).all(axis=0)
    if rates_differ:
        ok = same.mean() >= 0.98 and (~close).mean() <= 0.02 and bool((q <= 1000.0).all())
    else:
        ok = same.mean() >= 0.99 and bool(close[same].all()) and bool((q <= 100.0).all())
    return ok, {"step_match": float(same.mean()), "max_q_matching": float(q[:, same].max(initial=0.0)),
                "max_q": float(q.max(initial=0.0)), "beyond_0.1": float((~close).mean())}


def test_reference_sensitivity_to_rate_ulps(cb6r5, oracle_cb6r5):
    """The reference against itself when every rate constant moves by at most
    one ulp (a random direction per constant; the difference between two
    libms' exp and pow in CALC_RCONST): a percent of cells take another step
    sequence or another order sequence of the same length, and differ by up
    to a few hundred (RTOL |c| + ATOL).  The bar for a GPU run with its own
    rate constants (gear_bar with rates_differ) is set from this."""
    o = oracle_cb6r5
    n = 1024
    d = make_cells(cb6r5, np.arange(n), seed=5, workload="clean")
    rki = np.zeros((cb6r5["nrxns"], n))
    for c in range(n):
        rki[:, c:c + 1], _ = o.rates(d["temp"][c:c + 1], d["pres"][c:c + 1], d["h2o"][c:c + 1], d["rj"][:, c:c + 1])
    r0 = o.gear_run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"], rki=rki)
    mv = np.random.default_rng(0).integers(-1, 2, size=rki.shape)
    rk = np.where(mv > 0, np.nextafter(rki, np.inf), np.where(mv < 0, np.nextafter(rki, -np.inf), rki))
    r1 = o.gear_run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"], rki=rk)
    ok = (r0["nsteps"] >= 0) & (r1["nsteps"] >= 0)
    good, s = gear_bar(r0["y"][:, ok], r1["y"][:, ok], r0["nsteps"][ok], r1["nsteps"][ok], rates_differ=True)
    print(s)
    assert good, s
    assert s["step_match"] < 1.0 and s["max_q"] > 100.0          # beyond the same-rates bar


def test_reference_sensitivity_to_pow_ulps(cb6r5, oracle_cb6r5):
    """SMVGEAR's discrete control makes the algorithm itself sensitive to the
    last bit of its pow() calls (the step/order selection, grsmvgear.F:
    880-905): moving every pow result by one ulp -- the difference between
    two libms -- changes the step count of a few tenths of a percent of
    cells, the stiffest ones, and their trace species by up to ~65
    (RTOL |c| + ATOL); every other cell stays within rounding.  The bar
    against libm's pow (gear_bar) is set from this.  Between glibc's pow and
    a correctly rounded one the difference is far smaller: glibc misrounds
    only a few calls in ten thousand."""
    from oracle.oracle import lib
    o = oracle_cb6r5
    d = make_cells(cb6r5, np.arange(1024), seed=11)
    r0 = o.gear_run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"])
    worst = 0.0
    for ulps in (1, -1):
        try:
            lib().om_gear_set_pow_nudge(ulps)
            r1 = o.gear_run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"])
        finally:
            lib().om_gear_set_pow_nudge(0)
        same = r0["nsteps"] == r1["nsteps"]
        assert 0.98 <= same.mean() < 1.0                   # the perturbation matters, in few cells
        assert np.median(r0["nsteps"][~same]) > 3 * np.median(r0["nsteps"])   # the stiff ones
        ok, s = gear_bar(r0["y"], r1["y"], r0["nsteps"], r1["nsteps"])
        print(ulps, s)
        assert ok, s
        assert s["max_q_matching"] < 1e-3                  # matching cells: rounding only
        worst = max(worst, s["max_q"])
    assert worst > 1.0                                      # beyond the plain 0.1 bar
    # libm's pow against a correctly rounded one (quad-precision powq, the
    # device's step control): glibc's pow misrounds a few calls in ten
    # thousand, which moves a few percent of cells by rounding only
    try:
        lib().om_gear_set_pow_exact(1)
        r2 = o.gear_run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"])
    finally:
        lib().om_gear_set_pow_exact(0)
    ok, s = gear_bar(r0["y"], r2["y"], r0["nsteps"], r2["nsteps"])
    identical = np.all(r0["y"] == r2["y"], axis=0).mean()
    print("correctly rounded", s, identical)
    assert ok and s["step_match"] >= 0.995 and 0.9 <= identical < 1.0, (s, identical)
