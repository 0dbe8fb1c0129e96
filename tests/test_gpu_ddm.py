"""CMAQ-DDM-3D's chemistry step on the GPU (cmaq_ros3_ddm_device) against the
oracle's restatement (ddm3d/DDM3D_CHEM.F:155-511): the forward and its YAVE
take the oracle's accepted steps (to the forward parity tolerance), and from
the GPU's YAVE the REAL(4) Jacobian, Crank-Nicolson right-hand sides and
LINPACK SGEFA / SGESL match the oracle's om_ddm_step bit for bit (same order,
same pivots).  Day and
night cells; initial-condition and rate parameters; the matrix in LDS (CB6r5,
CRACMM2) and in global memory (SAPRC07TIC, 230 species).
"""
import numpy as np
import pytest

from cmaq_amd.inputs import make_cells

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1.0e-3, 1.0e-7


def _dev(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def _params(mech):
    """Up to 16 parameters: initial conditions of a few species, then rate
    constants of a few photolysis and thermal reactions."""
    names = mech["chemistry_spc"]
    init = [names.index(s) for s in ("O3", "NO2", "NO", "FORM", "HCHO", "ISOP", "ISO") if s in names][:5]
    rates = [r for r in range(mech["nrxns"]) if mech["nreact"][r] > 0][:: max(1, mech["nrxns"] // 9)][:9]
    return init, rates


@pytest.mark.parametrize("name", ["cb6r5_ae7_aq", "cracmm2", "saprc07tic_ae7i_aq"])
def test_ddm_step_bit_exact_vs_oracle(name):
    import torch
    from cmaq_amd.ros3 import Ros3Solver, load_mechanism
    from oracle.oracle import Oracle
    m = load_mechanism(name)
    o = Oracle(m)
    n = 12
    d = make_cells(m, np.arange(n), seed=8, workload="mixed")
    assert d["night"].any() and (~d["night"]).any()
    rki, _ = o.rates(d["temp"], d["pres"], d["h2o"], d["rj"])
    init, rates = _params(m)
    prate = [0] * len(init) + [r + 1 for r in rates]
    npar, ns = len(prate), m["numb_mech_spc"]
    sens0 = np.zeros((npar, ns, n), np.float32)
    for p, j in enumerate(init):
        sens0[p, j] = d["y"][j].astype(np.float32)           # S0 = C0 (init/initscen.F:306-314)
    s = Ros3Solver(m, RTOL, ATOL)
    y = _dev(d["y"]).clone()
    sens = _dev(sens0).clone()
    yave = torch.zeros_like(y)
    na = torch.zeros(n, dtype=torch.int32, device="cuda")
    s.ddm_device(5.0, _dev(rki), y, sens, prate, na, yave=yave)
    torch.cuda.synchronize()
    sg, yg, ya, nag = sens.cpu().numpy(), y.cpu().numpy(), yave.cpu().numpy(), na.cpu().numpy()
    assert (nag > 0).all()
    for c in range(n):
        k = rki[:, c]
        # the forward and its YAVE: the oracle's accepted steps, to the
        # forward parity tolerance (operation order differs between the two)
        h, yend = o.steps_cell(5.0, k, d["y"][:, c], RTOL, ATOL)
        assert nag[c] == len(h), c
        yy, acc = d["y"][:, c].copy(), np.zeros(ns)
        for hn in h:
            yn = np.maximum(o.replay_cell(k, yy, [hn]), 1e-30)
            acc = acc + 0.5 * (yy + yn) * hn
            yy = yn
        tol = 0.1 * (RTOL * np.abs(yend) + ATOL)
        assert (np.abs(yg[:, c] - yend) <= tol).all(), c
        assert (np.abs(ya[:, c] - acc / 5.0) <= 0.1 * (RTOL * np.abs(acc / 5.0) + ATOL)).all(), c
        # the sensitivity step from the GPU's own YAVE: bit for bit
        ref = o.ddm_step(5.0, k, ya[:, c], sens0[:, :, c], prate)
        for p in range(npar):
            assert np.array_equal(sg[p, :, c], ref[p]), (name, c, p, prate[p])
    print(f"{name}: DDM step on {n} cells ({int(d['night'].sum())} night), {npar} parameters, bit-exact")


def test_ddm_argument_checks(cb6r5):
    import torch
    from cmaq_amd.ros3 import Ros3Error, Ros3Solver
    s = Ros3Solver(cb6r5, RTOL, ATOL)
    n = 4
    d = make_cells(cb6r5, np.arange(n), seed=1)
    rki = torch.zeros((cb6r5["nrxns"], n), dtype=torch.float64, device="cuda")
    s.rates_device(_dev(d["temp"]), _dev(d["pres"]), _dev(d["h2o"]), _dev(d["rj"]), rki)
    y = _dev(d["y"]).clone()
    na = torch.zeros(n, dtype=torch.int32, device="cuda")
    with pytest.raises(Ros3Error) as e:
        s.ddm_device(5.0, rki, y, torch.zeros((17, s.ns, n), dtype=torch.float32, device="cuda"), [0] * 17, na)
    assert e.value.code == -1
    with pytest.raises(Ros3Error) as e:
        s.ddm_device(5.0, rki, y, torch.zeros((1, s.ns, n), dtype=torch.float32, device="cuda"),
                     [cb6r5["nrxns"] + 1], na)
    assert e.value.code == -1
