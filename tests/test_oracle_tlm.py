"""The oracle's tangent-linear (forward-mode) sensitivity of the fixed-step
ROS3 map (om_tlm_cell; no reference counterpart, an exact alternative to
DDM-3D's Crank-Nicolson step).

Pinned by: central finite differences (h = 1e-5 relative) of the replayed
map along initial-condition and rate-constant directions; the duality with
the separately pinned discrete adjoint, w . TLM(dy0, dk) = lam0 . dy0 +
gk . dk, for random directions; and linearity in the direction.
"""
import numpy as np
import pytest

from cmaq_amd.inputs import make_cells

FD_REL = 1.0e-5


def _cell(mech, o, seed, workload="clean"):
    d = make_cells(mech, np.arange(1), seed=seed, workload=workload)
    rki, _ = o.rates(d["temp"], d["pres"], d["h2o"], d["rj"])
    return rki[:, 0].copy(), d["y"][:, 0].copy()


@pytest.mark.parametrize("seed,workload", [(3, "clean"), (7, "mixed")])
def test_tlm_vs_central_fd(cb6r5, oracle_cb6r5, seed, workload):
    o = oracle_cb6r5
    rki, y0 = _cell(cb6r5, o, seed, workload)
    h, _ = o.steps_cell(5.0, rki, y0)
    names = cb6r5["chemistry_spc"]
    rng = np.random.default_rng(seed)
    # semi-normalised directions, as DDM-3D's INIT and RATE parameters
    dirs = [(np.eye(len(y0))[names.index("NO2")] * y0[names.index("NO2")], None),
            (np.eye(len(y0))[names.index("FORM")] * y0[names.index("FORM")], None),
            (np.zeros(len(y0)), rki * (rng.uniform(size=rki.size) < 0.2)),
            (y0 * rng.uniform(-1, 1, y0.size), rki * rng.uniform(-1, 1, rki.size))]
    yend = o.replay_cell(rki, y0, h)
    for dy0, dk in dirs:
        t = o.tlm_cell(rki, y0, h, dy0, dk)
        e = 1e-5
        kp = rki + e * (dk if dk is not None else 0.0)
        km = rki - e * (dk if dk is not None else 0.0)
        fd = (o.replay_cell(kp, y0 + e * dy0, h) - o.replay_cell(km, y0 - e * dy0, h)) / (2 * e)
        # the difference quotient's rounding floor is ~1e-16 |y_end| / e
        bar = FD_REL * np.abs(fd).max() + 1e-10 * np.abs(yend).max()
        assert np.abs(t - fd).max() <= bar, (np.abs(t - fd).max(), np.abs(fd).max())


def test_tlm_adjoint_duality(cb6r5, oracle_cb6r5):
    o = oracle_cb6r5
    rng = np.random.default_rng(1)
    for seed in (2, 5):
        rki, y0 = _cell(cb6r5, o, seed, "mixed")
        h, _ = o.steps_cell(5.0, rki, y0)
        w = rng.normal(size=y0.size)
        lam0, gk = o.adjoint_cell(rki, y0, h, w)
        for _ in range(3):
            dy0, dk = y0 * rng.normal(size=y0.size), rki * rng.normal(size=rki.size)
            lhs = w @ o.tlm_cell(rki, y0, h, dy0, dk)
            rhs = lam0 @ dy0 + gk @ dk
            scale = np.abs(w).max() * np.abs(o.tlm_cell(rki, y0, h, dy0, dk)).sum()
            assert abs(lhs - rhs) <= 1e-10 * scale, (lhs, rhs)


def test_tlm_is_linear(cb6r5, oracle_cb6r5):
    o = oracle_cb6r5
    rki, y0 = _cell(cb6r5, o, 9, "mixed")
    h, _ = o.steps_cell(5.0, rki, y0)
    rng = np.random.default_rng(2)
    d1, d2 = y0 * rng.normal(size=y0.size), y0 * rng.normal(size=y0.size)
    k1, k2 = rki * rng.normal(size=rki.size), rki * rng.normal(size=rki.size)
    a = o.tlm_cell(rki, y0, h, 2.0 * d1 - 3.0 * d2, 2.0 * k1 - 3.0 * k2)
    b = 2.0 * o.tlm_cell(rki, y0, h, d1, k1) - 3.0 * o.tlm_cell(rki, y0, h, d2, k2)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
