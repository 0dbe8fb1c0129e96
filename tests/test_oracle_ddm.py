"""The oracle's restatement of the CMAQ-DDM-3D chemistry step, and how the
discrete adjoint compares with it.

DDM-3D (ddm3d/DDM3D_CHEM.F:155-511) advances first-order sensitivities by one
Crank-Nicolson step per sync step with the Jacobian at the time-averaged
concentration and REAL(4) matrices.  It is pinned here by the closed forms of
that step on the decay A -> B and the chain A -> B -> C, and by the same step in
float64 with numpy on a CB6r5 cell.  Against the exact discrete adjoint it agrees to a few
1e-3 where the chemistry is slow over the sync step (night); in sunlit cells
its one-step CN approximation of fast photochemistry differs at O(0.1-1) in the
fast species, so the comparison there is reported, not bounded.
"""
import numpy as np

from cmaq_amd.inputs import make_cells


def test_ddm_linear_decay_closed_form():
    """A -> B, rate k, one sync step h: with J = [[-k, 0], [k, 0]],
    S_A = (1 - hk/2)/(1 + hk/2) C0 and S_B = hk C0/(1 + hk/2) for the initial
    condition of A; S_A = -hk Yave_A/(1 + hk/2), S_B = -S_A for ln k."""
    from oracle.oracle import Oracle
    from test_oracle import _decay_mech
    k, h, c0 = 0.05, 5.0, 0.7
    o = Oracle(_decay_mech(k))
    y0 = np.array([c0, 0.0])
    rki = np.array([k])
    steps, _ = o.steps_cell(h, rki, y0)
    # time-averaged concentration (trapezoid over accepted steps)
    y, yave = y0.copy(), np.zeros(2)
    for hn in steps:
        yn = o.replay_cell(rki, y, [hn])
        yave += 0.5 * (y + yn) * hn
        y = yn
    yave /= h
    s_init, s_rate, _ = o.ddm_cell(h, rki, y0)
    z = h * k / 2.0
    np.testing.assert_allclose(s_init[0], [(1 - z) / (1 + z) * c0, 2 * z * c0 / (1 + z)], rtol=2e-6)
    np.testing.assert_allclose(s_rate[0], [-2 * z * yave[0] / (1 + z), 2 * z * yave[0] / (1 + z)], rtol=2e-6)


def test_ddm_vs_adjoint_at_night(cb6r5, oracle_cb6r5):
    """A night cell (J = 0) with slow chemistry: the seed-contracted DDM-3D
    sensitivities to ln C0 match the discrete adjoint's to 1e-2 of the largest
    entry, over 1 and 12 sync steps.  (Night cells with fast NO3/N2O5 cycling
    agree only in the dominant entry; see tests/test_gpu_adjoint.py.)"""
    o = oracle_cb6r5
    names = cb6r5["chemistry_spc"]
    w = np.zeros(len(names))
    w[names.index("O3")] = 1.0
    w[names.index("NO2")] = 0.1
    d = make_cells(cb6r5, np.arange(1), seed=3)
    assert not d["rj"].any()
    rki, _ = o.rates(d["temp"], d["pres"], d["h2o"], d["rj"])
    k, y0 = rki[:, 0], d["y"][:, 0].copy()
    for nsteps in (1, 12):
        ys, hs, y = [y0], [], y0.copy()
        for _ in range(nsteps):
            h, y = o.steps_cell(5.0, k, y)
            hs.append(h)
            ys.append(y)
        lam = w.copy()
        for s in range(nsteps - 1, -1, -1):
            lam, _ = o.adjoint_cell(k, ys[s], hs[s], lam)
        s_adj = lam * y0
        s_ddm, _, _ = o.ddm_cell(5.0, k, y0, nsteps=nsteps, w=w)
        assert np.abs(s_adj - s_ddm).max() <= 1e-2 * np.abs(s_adj).max(), nsteps


def _chain_mech(k1, k2):
    """A -> B -> C, first order, k in 1/s (KTYPE 1, IORDER 1: RKI = k * SFACT)."""
    from test_oracle import _decay_mech
    m = _decay_mech(k1)
    m.update({"mechname": "CHAIN", "numb_mech_spc": 3, "nrxns": 2, "irr": [[1, 0, 0, 2], [2, 0, 0, 3]],
              "sc": [[1.0], [1.0]], "nreact": [1, 1], "nprdct": [1, 1], "ktype": [1, 1], "irxbits": [0, 0],
              "iorder": [1, 1], "rtdat": [[k1, 0.0, 0.0], [k2, 0.0, 0.0]], "chemistry_spc": ["A", "B", "C"]})
    return m


def test_ddm_linear_chain_closed_form():
    """A -> B -> C (rates k1, k2), one sync step h, a = h k1 / 2, b = h k2 / 2:
    for the initial condition of A the CN step gives
    S = C0 ((1-a)/(1+a), 2a/((1+a)(1+b)), 2ab/((1+a)(1+b))); for ln k2,
    S = (0, -2b Yave_B/(1+b), 2b Yave_B/(1+b))."""
    from oracle.oracle import Oracle
    k1, k2, h, c0 = 0.01, 0.004, 5.0, 0.9
    o = Oracle(_chain_mech(k1, k2))
    y0 = np.array([c0, 0.02, 0.0])
    rki = np.array([k1, k2]) * 60.0
    steps, _ = o.steps_cell(h, rki, y0)
    y, yave = y0.copy(), np.zeros(3)
    for hn in steps:
        yn = o.replay_cell(rki, y, [hn])
        yave += 0.5 * (y + yn) * hn
        y = yn
    yave /= h
    s_init, s_rate, _ = o.ddm_cell(h, rki, y0)
    a, b = h * rki[0] / 2.0, h * rki[1] / 2.0
    want = c0 * np.array([(1 - a) / (1 + a), 2 * a / ((1 + a) * (1 + b)), 2 * a * b / ((1 + a) * (1 + b))])
    np.testing.assert_allclose(s_init[0], want, rtol=3e-6)
    np.testing.assert_allclose(s_init[1], y0[1] * np.array([0.0, (1 - b) / (1 + b), 2 * b / (1 + b)]),
                               rtol=3e-6, atol=1e-12)
    wb = 2 * b * yave[1] / (1 + b)
    np.testing.assert_allclose(s_rate[1], [0.0, -wb, wb], rtol=3e-6, atol=1e-12)


def test_ddm_step_vs_dense_float64_crank_nicolson(cb6r5, oracle_cb6r5):
    """One DDM-3D step of a CB6r5 night cell against the same Crank-Nicolson
    step in float64 with numpy: the Jacobian at YAVE from the FD-pinned sparse
    Jacobian, (I - h/2 J) S = (I + h/2 J) S0 + P solved by numpy.  The REAL(4)
    LINPACK result agrees to 1e-3 of the largest entry (a wrong pivot, sign or
    transposed matrix would be O(1) off)."""
    o, m = oracle_cb6r5, cb6r5
    d = make_cells(m, np.arange(1), seed=3)
    rki, _ = o.rates(d["temp"], d["pres"], d["h2o"], d["rj"])
    k, y0, h = rki[:, 0], d["y"][:, 0].copy(), 5.0
    steps, _ = o.steps_cell(h, k, y0)
    y, yave = y0.copy(), np.zeros_like(y0)
    for hn in steps:
        yn = o.replay_cell(k, y, [hn])
        yave += 0.5 * (y + np.maximum(yn, 1e-30)) * hn
        y = yn
    yave /= h
    inew = o.inew2old() - 1
    ns = len(y0)
    cc2 = np.asarray(o.jacob(1, rki, yave[inew][:, None])).reshape(-1)
    rows, cols = o.pattern(1)
    jn = np.zeros((ns, ns))
    for s_ in range(len(rows)):
        jn[rows[s_] - 1, cols[s_] - 1] = cc2[s_ + 1]
    jac = np.empty_like(jn)                    # mechanism order
    jac[np.ix_(inew, inew)] = jn
    M, N = np.eye(ns) - 0.5 * h * jac, np.eye(ns) + 0.5 * h * jac
    s_init, s_rate, _ = o.ddm_cell(h, k, y0)
    names = m["chemistry_spc"]
    for j in (names.index("O3"), names.index("NO2"), names.index("N2O5")):
        s0 = np.zeros(ns)
        s0[j] = y0[j]
        ref = np.linalg.solve(M, N @ s0)
        assert np.abs(s_init[j] - ref).max() <= 1e-3 * np.abs(ref).max(), names[j]
    for r in range(0, m["nrxns"], 37):
        if m["nreact"][r] == 0 or k[r] == 0.0:
            continue
        p = np.zeros(ns)
        re = [m["irr"][r][q] - 1 for q in range(m["nreact"][r])]
        rate = h * k[r] * np.prod(yave[re])
        for a in re:
            p[a] -= rate
        for q in range(m["nprdct"][r]):
            p[m["irr"][r][3 + q] - 1] += m["sc"][r][q] * rate
        ref = np.linalg.solve(M, p)
        if np.abs(ref).max() > 0:
            assert np.abs(s_rate[r] - ref).max() <= 1e-3 * np.abs(ref).max(), r


def test_ddm_step_alone_equals_the_cell_step(cb6r5, oracle_cb6r5):
    """om_ddm_step from the trapezoid YAVE of the oracle's own accepted steps
    reproduces om_ddm_cell bit for bit (a day cell; initial-condition and rate
    parameters), so the step alone is what the cell step pins."""
    o, m = oracle_cb6r5, cb6r5
    d = make_cells(m, np.arange(8), seed=5)
    c = int(np.flatnonzero(~d["night"])[0])
    rki, _ = o.rates(d["temp"], d["pres"], d["h2o"], d["rj"])
    k, y0, h = rki[:, c], d["y"][:, c].copy(), 5.0
    steps, _ = o.steps_cell(h, k, y0)
    y, yave = y0.copy(), np.zeros_like(y0)
    for hn in steps:
        yn = np.maximum(o.replay_cell(k, y, [hn]), 1e-30)
        yave = yave + 0.5 * (y + yn) * hn
        y = yn
    s_init, s_rate, _ = o.ddm_cell(h, k, y0)
    ns, nr = len(y0), len(k)
    sens = np.zeros((ns + nr, ns), np.float32)
    sens[np.arange(ns), np.arange(ns)] = y0.astype(np.float32)
    prate = [0] * ns + list(range(1, nr + 1))
    out = o.ddm_step(h, k, yave / h, sens, prate)
    assert np.array_equal(out[:ns], s_init.astype(np.float32))
    assert np.array_equal(out[ns:], s_rate.astype(np.float32))
