"""GPU parity for every other distinct chemmech mechanism of the CMAQ tree
(CB6r3, CB6r5m, CRACMM1, CRACMM1-AMORE, RACM2, SAPRC07TC, SAPRC07TIC), with
the same bars as test_gpu_parity.py and test_gpu_adjoint.py: rate constants
to a few ulps, every species within 0.1 * (RTOL*|c| + ATOL) of the oracle,
accepted-step counts equal in >= 99.9 % of cells, and the discrete adjoint
within 1e-8 of the oracle's (relative to the cell's largest entry).
"""
import numpy as np
import pytest

from cmaq_amd.inputs import make_cells

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1.0e-3, 1.0e-7
NAMES = ["cb6r3_ae7_aq", "cb6r5m_ae7_aq", "cracmm1_aq", "cracmm1amore_aq", "racm2_ae6_aq",
         "saprc07tc_ae6_aq", "saprc07tic_ae7i_aq"]


def _dev(a, dtype=None):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda", dtype=dtype)


@pytest.fixture(scope="module", params=NAMES)
def mos(request):
    from cmaq_amd.ros3 import Ros3Solver, load_mechanism
    from oracle.oracle import Oracle
    mech = load_mechanism(request.param)
    print(f"mechanism {request.param}", flush=True)
    return request.param, mech, Oracle(mech), Ros3Solver(mech, RTOL, ATOL)


def _het_inputs(mech, n, seed):
    rng = np.random.default_rng(seed)
    het = np.exp(rng.uniform(np.log(1e-6), np.log(1e-1), (max(mech["nhetero"], 1), n)))
    sea = np.where(rng.uniform(size=n) < 0.5, rng.uniform(0.0, 1.0, n), 0.0)
    return np.ascontiguousarray(het), sea


def test_rates_match_oracle(mos):
    import torch
    name, mech, o, s = mos
    n = 64
    d = make_cells(mech, np.arange(n), seed=3)
    het, sea = _het_inputs(mech, n, 3)
    rki = torch.zeros((mech["nrxns"], n), dtype=torch.float64, device="cuda")
    s.rates_device(_dev(d["temp"]), _dev(d["pres"]), _dev(d["h2o"]), _dev(d["rj"]), rki,
                   het=_dev(het), seawater=_dev(sea))
    torch.cuda.synchronize()
    g = rki.cpu().numpy()
    for c in range(n):
        r, _ = o.rates(d["temp"][c:c + 1], d["pres"][c:c + 1], d["h2o"][c:c + 1], d["rj"][:, c:c + 1],
                       het[:, c:c + 1], sea[c:c + 1])
        np.testing.assert_allclose(g[:, c], r[:, 0], rtol=1e-14, atol=0, err_msg=f"{name} cell {c}")


@pytest.mark.parametrize("workload,seed,n", [("clean", 0, 1024), ("mixed", 2, 512)])
def test_parity(mos, workload, seed, n):
    import torch
    name, mech, o, s = mos
    d = make_cells(mech, np.arange(n), seed=seed, workload=workload)
    het, sea = _het_inputs(mech, n, seed)
    ref = o.run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"], RTOL, ATOL, het=het, seawater=sea)
    y = _dev(d["y"]).clone()
    na = torch.zeros(n, dtype=torch.int32, device="cuda")
    s.run_device(5.0, _dev(d["temp"]), _dev(d["pres"]), _dev(d["h2o"]), _dev(d["rj"]), y, na,
                 het=_dev(het), seawater=_dev(sea), raise_on_fail=False)
    torch.cuda.synchronize()
    yg, nag = y.cpu().numpy(), na.cpu().numpy()
    fail_ref = ref["nacc"] < 0
    ok = ~fail_ref
    tol = 0.1 * (RTOL * np.abs(ref["y"]) + ATOL)
    bad = int((np.abs(yg - ref["y"]) > tol)[:, ok].sum())
    same = float((nag[ok] == ref["nacc"][ok]).mean())
    exact = float(np.all(yg == ref["y"], axis=0)[ok].mean())
    print(f"{name} {workload} ({s.variant}): cells {n}, failed {int(fail_ref.sum())}, species outside tol "
          f"{bad}, accepted-step match {same:.5f}, bit-identical cells {exact:.4f}")
    assert np.array_equal(fail_ref, nag < 0)
    assert same >= 0.999 and bad == 0


def test_adjoint_matches_oracle(mos):
    import torch
    name, mech, o, s = mos
    n = 128
    d = make_cells(mech, np.arange(n), seed=26, workload="mixed")
    rki = torch.zeros((mech["nrxns"], n), dtype=torch.float64, device="cuda")
    s.rates_device(_dev(d["temp"]), _dev(d["pres"]), _dev(d["h2o"]), _dev(d["rj"]), rki)
    y0 = _dev(d["y"])
    y = y0.clone()
    na = torch.zeros(n, dtype=torch.int32, device="cuda")
    s.solve_device(5.0, rki, y, na, raise_on_fail=False)
    names = mech["chemistry_spc"]
    w = np.zeros((len(names), n))
    w[names.index("O3")] = 1.0
    w[names.index("NO2")] = 0.1
    lam = _dev(w).clone()
    gk = torch.zeros_like(rki)
    s.adjoint_device(5.0, rki, y0, na, lam, gk)
    kk = rki.cpu().numpy()
    ref = o.adjoint_run(5.0, kk, d["y"], w, RTOL, ATOL)
    nag, lg, gg = na.cpu().numpy(), lam.cpu().numpy(), gk.cpu().numpy()
    ok = (nag >= 0) & (ref["nacc"] == nag)
    assert ok.mean() >= 0.99
    for c in np.flatnonzero(ok):
        sl, rl = lg[:, c] * d["y"][:, c], ref["lam"][:, c] * d["y"][:, c]
        sg, rg = gg[:, c] * kk[:, c], ref["gk"][:, c] * kk[:, c]
        assert np.abs(sl - rl).max() <= 1e-8 * np.abs(rl).max(), (name, c)
        assert np.abs(sg - rg).max() <= 1e-8 * np.abs(rg).max(), (name, c)
    print(f"adjoint {name}: cells {n}, step counts equal {ok.mean():.4f}")


def test_rate_type7_on_gpu(cb6r5):
    """chemmech type 7 rates (pressure-dependent) in the GPU's rate program
    against the oracle's restatement of the generated CALC_RCONST text."""
    import torch
    from oracle.oracle import Oracle
    from cmaq_amd.ros3 import Ros3Solver
    from test_oracle import _ktype7_mech
    m, r1, r2 = _ktype7_mech(cb6r5)
    o, s = Oracle(m), Ros3Solver(m)
    n = 256
    d = make_cells(m, np.arange(n), seed=6)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")
    rki = torch.zeros((m["nrxns"], n), dtype=torch.float64, device="cuda")
    s.rates_device(dev(d["temp"]), dev(d["pres"]), dev(d["h2o"]), dev(d["rj"]), rki)
    torch.cuda.synchronize()
    g = rki.cpu().numpy()
    for c in range(0, n, 16):
        r, _ = o.rates(d["temp"][c:c + 1], d["pres"][c:c + 1], d["h2o"][c:c + 1], d["rj"][:, c:c + 1])
        np.testing.assert_allclose(g[[r1, r2], c], r[[r1, r2], 0], rtol=1e-14, atol=0)
        np.testing.assert_allclose(g[:, c], r[:, 0], rtol=1e-14, atol=0)
