"""CHEM-level entry points on the GPU: the CGRID gather and scatter bit for bit
against the oracle's restatement of rbdriver.F:460-509 / 577-595, and one
whole CHEM call (gather, CALC_RCONST + RBSOLVER, scatter) against the oracle
chain on the same grid with the solver's parity bar: every species within
0.1 * (RTOL*|c| + ATOL) (plus float32 storage), equal accepted-step counts in
>= 99.9 % of cells, the same failed cells.
"""
import numpy as np
import pytest

from cmaq_amd.inputs import make_grid

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1.0e-3, 1.0e-7
ULP4 = 4 * 2.0 ** -24


def _dev_grid(g):
    import torch
    out = dict(g)
    for k in ("cgrid", "ta", "dens", "pres", "qv", "seawater_zone", "seaice", "rj", "khetero"):
        if g.get(k) is not None:
            out[k] = torch.as_tensor(np.ascontiguousarray(g[k]), device="cuda")
    return out


@pytest.fixture(scope="module", params=["cb6r5_ae7_aq", "saprc07tic_ae7i_aq"])
def setup(request):
    from cmaq_amd.ros3 import Ros3Solver, load_mechanism
    from oracle.oracle import Oracle
    m = load_mechanism(request.param)
    print(f"mechanism {request.param}", flush=True)
    g, idx, conv, mw, d = make_grid(m, 37, 23, 5, seed=7)
    g["cgrid"][idx[0] - 1, :5] = np.array([0.0, -1.0, 1e-35, 1e-31, 2e-30], np.float32)
    s = Ros3Solver(m, RTOL, ATOL)
    s.set_cgrid_map(g["nspcsd"], idx, conv, mw)
    return request.param, m, Oracle(m), s, g, idx, conv, mw


def test_gather_bit_exact(setup):
    from oracle.oracle import chem_gather
    name, m, o, s, g, idx, conv, mw = setup
    ref = chem_gather(m, g, idx, conv, mw)
    got = s.gather_device(_dev_grid(g))
    for k in ("temp", "pres_atm", "h2o", "sea", "rj", "y"):
        assert np.array_equal(got[k].cpu().numpy(), ref[k]), (name, k)
    assert np.array_equal(got["het"].cpu().numpy(), ref["het"][:m["nhetero"]]), name


def test_scatter_bit_exact(setup):
    import torch
    from oracle.oracle import chem_gather, chem_scatter
    name, m, o, s, g, idx, conv, mw = setup
    y = chem_gather(m, g, idx, conv, mw)["y"] * 1.3
    ref = chem_scatter(m, g, idx, conv, mw, y)
    dg = _dev_grid(g)
    s.scatter_device(dg, torch.as_tensor(y, device="cuda"))
    torch.cuda.synchronize()
    assert np.array_equal(dg["cgrid"].cpu().numpy(), ref), name


def test_chem_call_matches_oracle_chain(setup):
    import torch
    from oracle.oracle import chem_gather, chem_scatter
    name, m, o, s, g, idx, conv, mw = setup
    b = chem_gather(m, g, idx, conv, mw)
    r = o.run(5.0, b["temp"], b["pres_atm"], b["h2o"], b["rj"], b["y"], RTOL, ATOL,
              het=b["het"][:m["nhetero"]], seawater=b["sea"])
    ref = chem_scatter(m, g, idx, conv, mw, r["y"])
    dg = _dev_grid(g)
    n = g["ta"].size
    na = torch.zeros(n, dtype=torch.int32, device="cuda")
    s.chem_device(dg, 5.0, nacc=na, raise_on_fail=False)
    got = dg["cgrid"].cpu().numpy()
    nag = na.cpu().numpy()
    fail = r["nacc"] < 0
    assert np.array_equal(fail, nag < 0)
    ok = ~fail
    assert (nag[ok] == r["nacc"][ok]).mean() >= 0.999
    # compare as mixing ratios: the oracle's gather of both grids
    yg = chem_gather(m, dict(g, cgrid=got), idx, conv, mw)["y"]
    yr = chem_gather(m, dict(g, cgrid=ref), idx, conv, mw)["y"]
    tol = 0.1 * (RTOL * np.abs(yr) + ATOL) + ULP4 * np.abs(yr)
    bad = int((np.abs(yg - yr) > tol)[:, ok].sum())
    print(f"{name} CHEM on a 37x23x5 grid: failed {int(fail.sum())}, species outside tol {bad}, "
          f"step match {(nag[ok] == r['nacc'][ok]).mean():.5f}")
    assert bad == 0
    others = np.setdiff1d(np.arange(g["cgrid"].shape[0]), idx - 1)
    assert np.array_equal(got[others], g["cgrid"][others])


def test_host_chem_equals_device_chem(setup):
    import torch
    name, m, o, s, g, idx, conv, mw = setup
    gh = dict(g, cgrid=g["cgrid"].copy())
    na_h, _ = s.chem(gh, 5.0, raise_on_fail=False)
    dg = _dev_grid(g)
    na = torch.zeros(g["ta"].size, dtype=torch.int32, device="cuda")
    s.chem_device(dg, 5.0, nacc=na, raise_on_fail=False)
    assert np.array_equal(gh["cgrid"], dg["cgrid"].cpu().numpy())
    assert np.array_equal(na_h, na.cpu().numpy())


def test_map_errors(setup):
    from cmaq_amd.ros3 import Ros3Error, Ros3Solver
    name, m, o, s, g, idx, conv, mw = setup
    fresh = Ros3Solver(m, RTOL, ATOL)
    with pytest.raises(Ros3Error) as e:
        fresh.gather_device(_dev_grid(g))
    assert e.value.code == -1 and "set_cgrid_map" in str(e.value)
    bad = idx.copy()
    bad[1] = bad[0]
    with pytest.raises(Ros3Error) as e:
        fresh.set_cgrid_map(g["nspcsd"], bad, conv, mw)
    assert e.value.code == -1
    with pytest.raises(Ros3Error):
        fresh.set_cgrid_map(int(idx.max()) - 1, idx, conv, mw)


def test_chem_after_steps_loop_on_one_context(setup):
    """One context used as bench.py uses it: the CGRID map is set, the graph
    sync loop runs on a cell count below the grid's (which once released the
    species map and the CHEM buffers), then a CHEM call.  The CHEM result and
    the loop's must equal those of fresh contexts bit for bit."""
    import torch
    from cmaq_amd.inputs import make_cells
    from cmaq_amd.ros3 import Ros3Solver
    name, m, o, s, g, idx, conv, mw = setup
    ctx = Ros3Solver(m, RTOL, ATOL)
    ctx.set_cgrid_map(g["nspcsd"], idx, conv, mw)
    n = 777
    d = make_cells(m, np.arange(n), seed=11)
    t = {k: torch.as_tensor(d[k], device="cuda") for k in ("temp", "pres", "h2o", "rj", "y")}

    def loop(solver):
        y = t["y"].clone()
        na = torch.zeros(2, n, dtype=torch.int32, device="cuda")
        solver.run_steps_device(5.0, 2, t["temp"], t["pres"], t["h2o"], t["rj"], y, nacc=na,
                                raise_on_fail=False)
        torch.cuda.synchronize()
        return y.cpu().numpy(), na.cpu().numpy()

    def chem(solver):
        dg = _dev_grid(g)
        na = torch.zeros(g["ta"].size, dtype=torch.int32, device="cuda")
        solver.chem_device(dg, 5.0, nacc=na, raise_on_fail=False)
        return dg["cgrid"].cpu().numpy(), na.cpu().numpy()

    y1, na1 = loop(ctx)
    c1, nc1 = chem(ctx)
    y2, na2 = loop(ctx)
    fresh = Ros3Solver(m, RTOL, ATOL)
    fresh.set_cgrid_map(g["nspcsd"], idx, conv, mw)
    c0, nc0 = chem(fresh)
    y0, na0 = loop(Ros3Solver(m, RTOL, ATOL))
    ctx.close()
    assert np.array_equal(c1, c0) and np.array_equal(nc1, nc0), name
    assert np.array_equal(y1, y0) and np.array_equal(na1, na0), name
    assert np.array_equal(y2, y0) and np.array_equal(na2, na0), name
