"""The oracle's restatement of the CMAQ-DDM-3D chemistry step, and how the
discrete adjoint compares with it.

DDM-3D (ddm3d/DDM3D_CHEM.F:155-511) advances first-order sensitivities by one
Crank-Nicolson step per sync step with the Jacobian at the time-averaged
concentration and REAL(4) matrices.  It is pinned here by its closed form on
the linear decay A -> B.  Against the exact discrete adjoint it agrees to a few
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
