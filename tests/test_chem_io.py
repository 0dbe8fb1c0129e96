"""Pins the oracle's CHEM gather and scatter (oracle/ros3_oracle.c om_gather,
om_scatter; rbdriver.F:460-509, 577-595) by the physics the driver's REAL(4)
factors encode, evaluated here in float64 from the same float32 inputs:

* aerosol species: mixing ratio [ppm] = mass concentration [ug/m3] * 1e-9 /
  air density [kg/m3] * MWAIR / MOLWT * 1e6;
* pressure [atm] = Pa / 101325; water vapour [ppm] = QV * 1e6 * MWAIR / MWWAT;
* sea-water fraction = max(0, open+surf zone - sea ice), per column;
* J-values and heterogeneous rates are copied (KHETERO with its own leading
  dimension); gas species are copied;
* every value is clamped to CONCMIN = 1e-30 (REAL(4)).
The REAL(4) results sit within a few float32 ulps of these; a dropped factor
or a wrong exponent fails them by orders of magnitude.  The scatter is
checked as the inverse of the gather and against the same physics.
"""
import numpy as np
import pytest

from cmaq_amd.inputs import make_grid

MWAIR, MWWAT = 28.9628, 18.0153
ULP4 = 4 * 2.0 ** -24          # four float32 ulps, relative


@pytest.fixture(scope="module", params=["cb6r5_ae7_aq", "saprc07tic_ae7i_aq"])
def case(request):
    from cmaq_amd.ros3 import load_mechanism
    m = load_mechanism(request.param)
    g, idx, conv, mw, d = make_grid(m, 9, 7, 4, seed=3)
    # a few values below CONCMIN, zero and negative
    g["cgrid"][idx[0] - 1, :5] = np.array([0.0, -1.0, 1e-35, 1e-31, 2e-30], np.float32)
    return m, g, idx, conv, mw, d


def test_gather_matches_unit_physics(case):
    from oracle.oracle import chem_gather
    m, g, idx, conv, mw, d = case
    b = chem_gather(m, g, idx, conv, mw)
    n = g["ta"].size
    conc = g["cgrid"][idx - 1].astype(np.float64)
    dens = g["dens"].astype(np.float64)[None, :]
    want = np.where(conv[:, None] == 1, conc * 1e-9 / dens * MWAIR / mw[:, None] * 1e6, conc)
    want = np.maximum(want, np.float64(np.float32(1e-30)))
    assert np.all(np.abs(b["y"] - want) <= ULP4 * np.abs(want))
    assert np.all(b["y"][conv == 0] == np.maximum(conc[conv == 0], np.float32(1e-30)))
    assert b["y"][0, 0] == b["y"][0, 1] == b["y"][0, 2] == np.float64(np.float32(1e-30))
    assert np.array_equal(b["temp"], g["ta"].astype(np.float64))
    pa = g["pres"].astype(np.float64) / 101325.0
    assert np.all(np.abs(b["pres_atm"] - pa) <= ULP4 * pa)
    h2o = g["qv"].astype(np.float64) * 1e6 * MWAIR / MWWAT
    assert np.all(np.abs(b["h2o"] - h2o) <= ULP4 * h2o)
    nlay = g["nlays"]
    sea2 = np.maximum(0.0, g["seawater_zone"].astype(np.float64) - g["seaice"].astype(np.float64))
    assert np.all(np.abs(b["sea"] - np.tile(sea2, nlay)) <= 1e-7)
    assert (b["sea"] > 0).any() and (b["sea"] == 0).any()
    assert np.array_equal(b["rj"], g["rj"].astype(np.float64))
    assert np.array_equal(b["het"][:m["nhetero"]], g["khetero"][:, :m["nhetero"]].T)
    assert b["y"].shape == (m["numb_mech_spc"], n)


def test_scatter_inverts_gather(case):
    from oracle.oracle import chem_gather, chem_scatter
    m, g, idx, conv, mw, d = case
    b = chem_gather(m, g, idx, conv, mw)
    cg = chem_scatter(m, g, idx, conv, mw, b["y"])
    before = np.maximum(g["cgrid"][idx - 1], np.float32(1e-30)).astype(np.float64)
    after = cg[idx - 1].astype(np.float64)
    assert np.array_equal(after[conv == 0], before[conv == 0])
    assert np.all(np.abs(after - before) <= ULP4 * before)
    others = np.setdiff1d(np.arange(g["cgrid"].shape[0]), idx - 1)
    assert np.array_equal(cg[others], g["cgrid"][others])


def test_scatter_matches_unit_physics(case):
    from oracle.oracle import chem_scatter
    m, g, idx, conv, mw, d = case
    y = d["y"] * 1.7
    cg = chem_scatter(m, g, idx, conv, mw, y)
    dens = g["dens"].astype(np.float64)[None, :]
    want = np.where(conv[:, None] == 1, y * 1e-6 * mw[:, None] / MWAIR * dens * 1e9, y)
    got = cg[idx - 1].astype(np.float64)
    assert np.all(np.abs(got - want) <= ULP4 * np.abs(want))
