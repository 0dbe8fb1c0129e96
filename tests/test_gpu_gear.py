"""SMVGEAR on the GPU (cmaq_gear_*) against the oracle's restatement of
grsmvgear.F with per-cell control (blocks of one cell).

The device's step/order selection takes x^y correctly rounded
(cmaq_amd/csrc/ros3_pow.h, tests/test_crpow.py).  With the oracle's pow
switched to a correctly rounded one too (quad-precision powq) and the same
rate constants, every cell is bit-identical: concentrations, step counts,
failure counts and the cells that fail.  Against the oracle with libm's pow
-- the reference's build -- the difference is the reference's own between
two correct libms (glibc's pow misrounds a few calls in ten thousand),
checked with gear_bar (tests/test_oracle_gear.py).  Where the GPU also
computes the rate constants, its exp and pow differ from glibc's in the last
bit now and then, and the bar is the reference's own spread under such a
difference (gear_bar with rates_differ).
"""
import contextlib

import numpy as np
import pytest

from cmaq_amd.inputs import make_cells
from test_oracle_gear import gear_bar

pytestmark = pytest.mark.gpu


def _dev(a, dtype=None):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda", dtype=dtype)


@contextlib.contextmanager
def exact_pow():
    """The oracle's SMVGEAR with a correctly rounded pow."""
    from oracle.oracle import lib
    lib().om_gear_set_pow_exact(1)
    try:
        yield
    finally:
        lib().om_gear_set_pow_exact(0)


@pytest.fixture(scope="module")
def solver_cb6r5():
    from cmaq_amd.ros3 import Ros3Solver
    return Ros3Solver("cb6r5_ae7_aq")


def _sunlit(d):
    return (d["rj"] > 0).any(axis=0)


def _oracle_gear(o, d, rki=None, tol=None):
    return o.gear_run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"], rki=rki, tol=tol)


def _gpu_rates(solver, mech, d):
    import torch
    n = d["temp"].size
    rki = torch.empty((mech["nrxns"], n), dtype=torch.float64, device="cuda")
    solver.rates_device(_dev(d["temp"]), _dev(d["pres"]), _dev(d["h2o"]), _dev(d["rj"]), rki)
    torch.cuda.synchronize()
    return rki.cpu().numpy()


def assert_bit_identical(ref, y, nsteps, nfail):
    bad = np.flatnonzero(~(np.all(y == ref["y"], axis=0) & (nsteps == ref["nsteps"]) & (nfail == ref["nfail"])))
    assert bad.size == 0, (f"{bad.size} cells differ, first {bad[:8]}", ref["nsteps"][bad[:8]], nsteps[bad[:8]])


@pytest.mark.parametrize("workload,n", [("clean", 1024), ("mixed", 1024)])
def test_gear_solve_follows_oracle(cb6r5, oracle_cb6r5, solver_cb6r5, workload, n):
    """Same rate constants: bit-identical to the oracle with a correctly
    rounded pow; within gear_bar of the oracle with libm's."""
    import torch
    o = oracle_cb6r5
    d = make_cells(cb6r5, np.arange(n), seed=11, workload=workload)
    rki = np.zeros((cb6r5["nrxns"], n))
    for c in range(n):
        rki[:, c:c + 1], _ = o.rates(d["temp"][c:c + 1], d["pres"][c:c + 1], d["h2o"][c:c + 1], d["rj"][:, c:c + 1])
    y = _dev(d["y"]).clone()
    ns_ = torch.zeros(n, dtype=torch.int32, device="cuda")
    nf = torch.zeros(n, dtype=torch.int32, device="cuda")
    solver_cb6r5.gear_solve_device(5.0, _dev(rki), y, _dev(_sunlit(d).astype(np.uint8)), ns_, nf,
                                   raise_on_fail=False)
    torch.cuda.synchronize()
    yg, nsg, nfg = y.cpu().numpy(), ns_.cpu().numpy(), nf.cpu().numpy()
    with exact_pow():
        ref = _oracle_gear(o, d, rki=rki)
    assert_bit_identical(ref, yg, nsg, nfg)
    lib_ref = _oracle_gear(o, d, rki=rki)
    good = lib_ref["nsteps"] >= 0
    assert np.array_equal(good, nsg >= 0)
    ok, s = gear_bar(lib_ref["y"][:, good], yg[:, good], lib_ref["nsteps"][good], nsg[good])
    exact = np.all(yg == lib_ref["y"], axis=0).mean()
    print(f"gear {workload}: steps median {np.median(ref['nsteps'])}, failed cells {(~good).sum()}, against "
          f"libm's pow: {s}, bit-identical cells {exact:.4f}")
    assert ok, s


@pytest.mark.parametrize("workload", ["clean", "mixed"])
def test_gear_run_parity(cb6r5, oracle_cb6r5, solver_cb6r5, workload):
    """The GPU's own CALC_RCONST and SMVGEAR: bit-identical to the oracle's
    SMVGEAR (correctly rounded pow) on the GPU's rate constants, which
    tests/test_gpu_parity.py holds to the oracle's CALC_RCONST; and within
    gear_bar of the oracle's CALC_RCONST + SMVGEAR with libm."""
    import torch
    n = 2048
    d = make_cells(cb6r5, np.arange(n), seed=5, workload=workload)
    y = _dev(d["y"]).clone()
    ns_ = torch.zeros(n, dtype=torch.int32, device="cuda")
    nf = torch.zeros(n, dtype=torch.int32, device="cuda")
    solver_cb6r5.gear_run_device(5.0, _dev(d["temp"]), _dev(d["pres"]), _dev(d["h2o"]), _dev(d["rj"]), y, ns_, nf,
                                 raise_on_fail=False)
    torch.cuda.synchronize()
    yg, nsg, nfg = y.cpu().numpy(), ns_.cpu().numpy(), nf.cpu().numpy()
    with exact_pow():
        ref = _oracle_gear(oracle_cb6r5, d, rki=_gpu_rates(solver_cb6r5, cb6r5, d))
    assert_bit_identical(ref, yg, nsg, nfg)
    lib_ref = _oracle_gear(oracle_cb6r5, d)
    ok = lib_ref["nsteps"] >= 0
    assert np.array_equal(ok, nsg >= 0)
    good, s = gear_bar(lib_ref["y"][:, ok], yg[:, ok], lib_ref["nsteps"][ok], nsg[ok], rates_differ=True)
    print(f"gear run {workload}: cells {n}, against the oracle's rates and libm's pow: {s}")
    assert good, s
    # the host entry point equals the device one bit for bit
    yh, nsh, nfh = solver_cb6r5.gear_run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"], raise_on_fail=False)
    assert np.array_equal(yh, yg) and np.array_equal(nsh, nsg) and np.array_equal(nfh, nfg)


@pytest.mark.parametrize("after,cap", [(1, 0), (7, 100), (0, 0)])
def test_gear_long_cell_handoff(cb6r5, oracle_cb6r5, after, cap):
    """Cells moved to the one-cell-per-wavefront LDS kernel after `after`
    attempted steps (1: nearly every cell; 7 with room for 100: those and no
    more; 0: none) end bit-identical to the oracle with a correctly rounded
    pow, as in the cell-per-lane kernel."""
    import torch
    from cmaq_amd.ros3 import Ros3Solver
    n = 1024
    d = make_cells(cb6r5, np.arange(n), seed=21, workload="mixed")
    s = Ros3Solver("cb6r5_ae7_aq")
    s.set_gear_tail(after, cap)
    yh, nsh, nfh = s.gear_run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"], raise_on_fail=False)
    handed = s.query("gear_tail_cells_last")
    print(f"handoff after {after} attempts, room {cap or 'auto'}: {handed} cells, solve {s.gear_last_timing()[1]:.1f} ms")
    if after == 0:
        assert handed == 0
    elif cap:
        assert handed == cap
    else:
        assert handed >= n - 8          # every cell that takes a second step
    with exact_pow():
        ref = _oracle_gear(oracle_cb6r5, d, rki=_gpu_rates(s, cb6r5, d))
    assert_bit_identical(ref, yh, nsh, nfh)
    # a second call on the same context reuses the buffers and clears the counters
    y2, ns2, nf2 = s.gear_run(5.0, d["temp"][:300], d["pres"][:300], d["h2o"][:300], d["rj"][:, :300],
                              d["y"][:, :300], raise_on_fail=False)
    assert np.array_equal(y2, yh[:, :300]) and np.array_equal(ns2, nsh[:300]) and np.array_equal(nf2, nfh[:300])
    with pytest.raises(Exception):
        s.set_gear_tail(-1, 0)


def test_gear_long_cell_handoff_largest_mechanism():
    """SAPRC07TIC, the largest mechanism (230 species, 929 reactions): its
    cell's workspace takes 78 KB of LDS in gear_tail_kernel, beyond the 64 KB
    an LDS instruction's offset reaches.  Every cell handed over after one
    step; bit-identical to the oracle with a correctly rounded pow."""
    from cmaq_amd.ros3 import Ros3Solver, load_mechanism
    from oracle.oracle import Oracle
    m = load_mechanism("saprc07tic_ae7i_aq")
    o = Oracle(m)
    n = 96
    d = make_cells(m, np.arange(n), seed=4, workload="clean")
    s = Ros3Solver("saprc07tic_ae7i_aq")
    s.set_gear_tail(1, 0)
    yh, nsh, nfh = s.gear_run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"], raise_on_fail=False)
    assert s.query("gear_tail_cells_last") >= n - 2
    with exact_pow():
        ref = _oracle_gear(o, d, rki=_gpu_rates(s, m, d))
    assert_bit_identical(ref, yh, nsh, nfh)


def test_gear_regional_262k_sample(cb6r5, oracle_cb6r5, solver_cb6r5):
    """262,144 regional cells in one call, as the bench runs them: a few
    thousand cells pass 2048 attempted steps and finish in gear_tail_kernel.
    A sample -- every 1024th cell, plus the 48 handed-over cells with the
    fewest attempts, which keeps the oracle's share to seconds -- is
    bit-identical to the oracle (correctly rounded pow, the GPU's rate
    constants); every other cell ends finite and at or above CONCMIN, or
    fails as the reference would (n_steps = -1)."""
    import torch
    from cmaq_amd.inputs import make_cells_torch
    n = 262144
    d = make_cells_torch(cb6r5, n, seed=7, workload="regional", nlay_cells=n // 35, device="cuda")
    y = d["y"].clone()
    ns_ = torch.zeros(n, dtype=torch.int32, device="cuda")
    nf = torch.zeros(n, dtype=torch.int32, device="cuda")
    solver_cb6r5.gear_run_device(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], y, ns_, nf, raise_on_fail=False)
    torch.cuda.synchronize()
    handed = solver_cb6r5.query("gear_tail_cells_last")
    ok = ns_ >= 0
    assert handed > 0 and float(ok.float().mean()) > 0.99
    assert bool(torch.isfinite(y[:, ok]).all()) and bool((y[:, ok] >= 1e-30).all())
    att = (ns_.clamp(min=0) + nf).to(torch.int64)
    over = torch.nonzero(att > 2048).flatten()
    over = over[torch.argsort(att[over])[:48]]
    cells = torch.unique(torch.cat([torch.arange(0, n, 1024, device="cuda"), over]))
    host = {k: d[k][..., cells].cpu().numpy() for k in ("temp", "pres", "h2o", "rj", "y")}
    print(f"gear 262144 regional cells: {handed} handed over; sample {cells.numel()} cells, "
          f"{over.numel()} of them handed over ({int(att[over].min())}-{int(att[over].max())} attempts)")
    assert over.numel() == 48
    with exact_pow():
        ref = _oracle_gear(oracle_cb6r5, host, rki=_gpu_rates(solver_cb6r5, cb6r5, host))
    assert_bit_identical(ref, y[:, cells].cpu().numpy(), ns_[cells].cpu().numpy(), nf[cells].cpu().numpy())


def test_gear_ragged_and_empty(cb6r5, oracle_cb6r5, solver_cb6r5):
    import torch
    for n in (1, 63, 65, 130):
        d = make_cells(cb6r5, np.arange(n), seed=n, workload="mixed")
        yh, nsh, nfh = solver_cb6r5.gear_run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"],
                                             raise_on_fail=False)
        with exact_pow():
            ref = _oracle_gear(oracle_cb6r5, d, rki=_gpu_rates(solver_cb6r5, cb6r5, d))
        assert_bit_identical(ref, yh, nsh, nfh)
    y = torch.zeros((cb6r5["numb_mech_spc"], 0), dtype=torch.float64, device="cuda")
    z = torch.zeros(0, dtype=torch.float64, device="cuda")
    solver_cb6r5.gear_run_device(5.0, z, z, z, torch.zeros((cb6r5["nphotab"], 0), dtype=torch.float64,
                                                           device="cuda"), y)


def test_gear_cracmm2(cracmm2, oracle_cracmm2):
    from cmaq_amd.ros3 import Ros3Solver
    n = 512
    d = make_cells(cracmm2, np.arange(n), seed=3, workload="mixed")
    s = Ros3Solver("cracmm2")
    yh, nsh, nfh = s.gear_run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"], raise_on_fail=False)
    with exact_pow():
        ref = _oracle_gear(oracle_cracmm2, d, rki=_gpu_rates(s, cracmm2, d))
    assert_bit_identical(ref, yh, nsh, nfh)
    lib_ref = _oracle_gear(oracle_cracmm2, d)
    ok = lib_ref["nsteps"] >= 0
    assert np.array_equal(ok, nsh >= 0)
    good, st = gear_bar(lib_ref["y"][:, ok], yh[:, ok], lib_ref["nsteps"][ok], nsh[ok], rates_differ=True)
    assert good, st


def test_gear_step_limit_reported(cb6r5):
    """A step limit far below what the cells need: every cell is reported
    (n_steps = -1) and the kernel still drains the queue."""
    from cmaq_amd.ros3 import Ros3Error, Ros3Solver
    d = make_cells(cb6r5, np.arange(200), seed=1)
    s = Ros3Solver("cb6r5_ae7_aq", max_attempts=3)
    yh, nsh, _ = s.gear_run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"], raise_on_fail=False)
    assert np.all(nsh == -1)
    with pytest.raises(Ros3Error):
        s.gear_run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"])


def test_gear_tolerances(cb6r5, oracle_cb6r5):
    """Non-default GEAR_RTOL / ATOL bounds reach the kernel as in the oracle."""
    from cmaq_amd.ros3 import Ros3Solver
    n = 256
    d = make_cells(cb6r5, np.arange(n), seed=9)
    tol = (1.0e-4, 1.0e-8, 1.0e-10, 1.0e-13)
    s = Ros3Solver("cb6r5_ae7_aq")
    s.set_gear_tolerances(*tol)
    yh, nsh, nfh = s.gear_run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"], raise_on_fail=False)
    with exact_pow():
        ref = _oracle_gear(oracle_cb6r5, d, rki=_gpu_rates(s, cb6r5, d), tol=tol)
    assert_bit_identical(ref, yh, nsh, nfh)
    with pytest.raises(Exception):
        s.set_gear_tolerances(1e-3, 1e-9, 1e-7, 1e-12)     # max_atol below min_atol
