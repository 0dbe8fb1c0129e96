"""GPU discrete adjoint against the oracle's adjoint and finite differences.

The oracle adjoint is pinned on the CPU (tests/test_oracle_adjoint.py:
central FD of the fixed-step map, closed form for linear decay).  Here the
GPU's cmaq_ros3_adjoint_device is compared with it on the same rate
constants, and the autograd wrapper with central FD (h = 1e-6 relative) of
the oracle's fixed-step replay.  Bar: 1e-5 relative to the largest gradient
entry of the cell.
"""
import numpy as np
import pytest

from cmaq_amd.inputs import make_cells

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1.0e-3, 1.0e-7
BAR = 1.0e-5


def _dev(a, dtype=None):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda", dtype=dtype)


@pytest.fixture(scope="module")
def solver():
    from cmaq_amd.ros3 import Ros3Solver
    return Ros3Solver("cb6r5_ae7_aq", RTOL, ATOL)


def _seed(mech, n):
    names = mech["chemistry_spc"]
    w = np.zeros((len(names), n))
    w[names.index("O3")] = 1.0
    w[names.index("NO2")] = 0.1
    return w


def _rates(solver, d):
    import torch
    n = d["temp"].size
    rki = torch.zeros((solver.nr, n), dtype=torch.float64, device="cuda")
    solver.rates_device(_dev(d["temp"]), _dev(d["pres"]), _dev(d["h2o"]), _dev(d["rj"]), rki)
    return rki


def test_solve_device_equals_run_device(cb6r5, solver):
    import torch
    n = 1000
    d = make_cells(cb6r5, np.arange(n), seed=21, workload="mixed")
    rki = _rates(solver, d)
    y1, y2 = _dev(d["y"]).clone(), _dev(d["y"]).clone()
    n1 = torch.zeros(n, dtype=torch.int32, device="cuda")
    n2 = torch.zeros(n, dtype=torch.int32, device="cuda")
    solver.run_device(5.0, _dev(d["temp"]), _dev(d["pres"]), _dev(d["h2o"]), _dev(d["rj"]), y1, n1,
                      raise_on_fail=False)
    solver.solve_device(5.0, rki, y2, n2, raise_on_fail=False)
    assert torch.equal(y1, y2) and torch.equal(n1, n2)


@pytest.mark.parametrize("workload,seed", [("mixed", 22), ("clean", 23)])
def test_adjoint_matches_oracle(cb6r5, oracle_cb6r5, solver, workload, seed):
    import torch
    n = 512
    d = make_cells(cb6r5, np.arange(n), seed=seed, workload=workload)
    rki = _rates(solver, d)
    y0 = _dev(d["y"])
    y = y0.clone()
    na = torch.zeros(n, dtype=torch.int32, device="cuda")
    solver.solve_device(5.0, rki, y, na, raise_on_fail=False)
    w = _seed(cb6r5, n)
    lam = _dev(w).clone()
    gk = torch.zeros_like(rki)
    solver.adjoint_device(5.0, rki, y0, na, lam, gk)
    ref = oracle_cb6r5.adjoint_run(5.0, rki.cpu().numpy(), d["y"], w, RTOL, ATOL)
    nag, lg, gg = na.cpu().numpy(), lam.cpu().numpy(), gk.cpu().numpy()
    ok = (nag >= 0) & (ref["nacc"] == nag)
    assert ok.mean() >= 0.99
    assert np.isnan(lg[:, nag < 0]).all()
    for c in np.flatnonzero(ok):
        sl, sg = lg[:, c] * d["y"][:, c], gg[:, c] * rki[:, c].cpu().numpy()
        rl, rg = ref["lam"][:, c] * d["y"][:, c], ref["gk"][:, c] * rki[:, c].cpu().numpy()
        assert np.abs(sl - rl).max() <= 1e-8 * np.abs(rl).max(), c
        assert np.abs(sg - rg).max() <= 1e-8 * np.abs(rg).max(), c
    print(f"adjoint {workload}: cells {n}, step counts equal {ok.mean():.4f}, "
          f"replay/reverse ms {solver.last_adjoint_timing()}")


def test_autograd_two_sync_steps_vs_fd(cb6r5, oracle_cb6r5, solver):
    """torch.autograd through 2 sync steps; loss = sum(O3 + 0.1 NO2).
    dL/dln y0 and dL/dln k checked against central FD of the oracle's
    fixed-step replay on 16 cells, top entries per cell."""
    import torch
    from cmaq_amd.adjoint import integrate
    n = 16
    d = make_cells(cb6r5, np.arange(n), seed=24, workload="mixed")
    rki = _rates(solver, d)
    y0 = _dev(d["y"]).requires_grad_(True)
    logk = torch.log(rki).requires_grad_(True)   # exp(log k) is not k bit for bit: use it on both sides
    k_used = logk.exp()
    yend = integrate(solver, y0, k_used, 5.0, 2)
    names = cb6r5["chemistry_spc"]
    loss = (yend[names.index("O3")] + 0.1 * yend[names.index("NO2")]).sum()
    loss.backward()
    gy = y0.grad.cpu().numpy() * d["y"]
    gl = logk.grad.cpu().numpy()
    kk = k_used.detach().cpu().numpy()
    w = _seed(cb6r5, 1)[:, 0]
    o = oracle_cb6r5
    for c in range(n):
        k = kk[:, c]
        h1, y1 = o.steps_cell(5.0, k, d["y"][:, c], RTOL, ATOL)
        h2, _ = o.steps_cell(5.0, k, y1, RTOL, ATOL)

        def L(y, kv):
            return w @ o.replay_cell(kv, o.replay_cell(kv, y, h1), h2)

        for j in np.argsort(-np.abs(gy[:, c]))[:4]:
            e = 1e-6 * d["y"][j, c]
            yp, ym = d["y"][:, c].copy(), d["y"][:, c].copy()
            yp[j] += e
            ym[j] -= e
            fd = (L(yp, k) - L(ym, k)) / (2 * e) * d["y"][j, c]
            assert abs(gy[j, c] - fd) <= BAR * np.abs(gy[:, c]).max(), (c, names[j], gy[j, c], fd)
        for r in np.argsort(-np.abs(gl[:, c]))[:4]:
            e = 1e-6 * k[r]
            kp, km = k.copy(), k.copy()
            kp[r] += e
            km[r] -= e
            fd = (L(d["y"][:, c], kp) - L(d["y"][:, c], km)) / (2 * e) * k[r]
            assert abs(gl[r, c] - fd) <= BAR * np.abs(gl[:, c]).max(), (c, r, gl[r, c], fd)


def test_adjoint_in_chunks_equals_single_pass(cb6r5, solver, monkeypatch):
    """A tape budget too small for all cells splits them into chunks; the
    gradients are the same bit for bit."""
    import torch
    n = 700
    d = make_cells(cb6r5, np.arange(n), seed=25, workload="mixed")
    rki = _rates(solver, d)
    y0 = _dev(d["y"])
    y = y0.clone()
    na = torch.zeros(n, dtype=torch.int32, device="cuda")
    solver.solve_device(5.0, rki, y, na, raise_on_fail=False)
    outs = []
    for budget in (None, 64 * 1024 * 1024):
        if budget:
            monkeypatch.setenv("CMAQ_ROS3_TAPE_BYTES", str(budget))
        lam = _dev(_seed(cb6r5, n)).clone()
        gk = torch.zeros_like(rki)
        solver.adjoint_device(5.0, rki, y0, na, lam, gk)
        outs.append((lam.cpu().numpy(), gk.cpu().numpy(), solver.last_adjoint_timing()[2]))
    assert outs[0][2] == 1 and outs[1][2] > 1
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


def test_solve_tape_device_integrates_as_solve_device(cb6r5, solver):
    """The recording forward gives the plain forward's y, nacc and nrej bit
    for bit, takes one record per accepted step (failed cells too, up to
    their failure), and links each cell's records into a chain of nacc
    records ending at its head."""
    import torch
    n = 777
    d = make_cells(cb6r5, np.arange(n), seed=26, workload="mixed")
    rki = _rates(solver, d)
    y1, y2 = _dev(d["y"]).clone(), _dev(d["y"]).clone()
    a1, a2, r1, r2 = (torch.zeros(n, dtype=torch.int32, device="cuda") for _ in range(4))
    solver.solve_device(5.0, rki, y1, a1, r1, raise_on_fail=False)
    tape, count = solver.new_tape(n * 64, device="cuda")
    head = torch.empty(n, dtype=torch.int64, device="cuda")
    assert solver.solve_tape_device(5.0, rki, y2, a2, tape, count, head, nrej=r2, raise_on_fail=False)
    assert torch.equal(y1, y2) and torch.equal(a1, a2) and torch.equal(r1, r2)
    na, hd = a2.cpu().numpy(), head.cpu().numpy()
    ok = int(na[na > 0].sum())
    assert ok <= int(count.item()) <= ok + int((na < 0).sum()) * 1000
    rs = solver.query("tape_record_doubles")
    links = tape.view(-1, rs)[:, rs - 1].cpu().numpy()
    for c in np.flatnonzero(na > 0)[::37]:
        r, k = hd[c], 0
        while r >= 0:
            r, k = int(links[r]), k + 1
        assert k == na[c], (c, k, na[c])
    assert (hd[na < 0] == -1).all()


def test_solve_tape_device_overflow(cb6r5, solver):
    """An arena too small for the sync step: the integration is still the
    plain forward's, the call reports the overflow, and the reverse sweep
    refuses the incomplete tape."""
    import torch
    from cmaq_amd.ros3 import Ros3Error
    n = 300
    d = make_cells(cb6r5, np.arange(n), seed=28, workload="mixed")
    rki = _rates(solver, d)
    y1, y2 = _dev(d["y"]).clone(), _dev(d["y"]).clone()
    a1, a2 = (torch.zeros(n, dtype=torch.int32, device="cuda") for _ in range(2))
    solver.solve_device(5.0, rki, y1, a1, raise_on_fail=False)
    tape, count = solver.new_tape(10, device="cuda")
    head = torch.empty(n, dtype=torch.int64, device="cuda")
    assert not solver.solve_tape_device(5.0, rki, y2, a2, tape, count, head, raise_on_fail=False)
    assert torch.equal(y1, y2) and torch.equal(a1, a2)
    assert int(count.item()) > 10
    lam = _dev(_seed(cb6r5, n)).clone()
    with pytest.raises(Ros3Error):
        solver.reverse_device(5.0, rki, a2, tape, count, head, lam)


def test_stored_tape_equals_replay(cb6r5, solver):
    """Gradients through autograd over 3 sync steps: the stored tape (reverse
    sweeps only), the replay path, a stored tape whose arena overflows in the
    first sync step (falls back to the replay) and one too small to attempt
    agree bit for bit."""
    import torch
    from cmaq_amd.adjoint import integrate
    n = 600
    d = make_cells(cb6r5, np.arange(n), seed=27, workload="mixed")
    rki = _rates(solver, d)
    names = cb6r5["chemistry_spc"]
    outs = []
    for tape, frac, recs in (("stored", 0.7, None), ("replay", 0.7, None), ("stored", 0.7, n * 3 * 4),
                             ("stored", 1e-12, None)):
        y0 = _dev(d["y"]).requires_grad_(True)
        kk = rki.clone().requires_grad_(True)
        yend = integrate(solver, y0, kk, 5.0, 3, tape=tape, tape_fraction=frac, tape_records=recs)
        (yend[names.index("O3")] + 0.1 * yend[names.index("NO2")]).sum().backward()
        rep, _, chunks = solver.last_adjoint_timing()
        outs.append((yend.detach().cpu().numpy(), y0.grad.cpu().numpy(), kk.grad.cpu().numpy(), rep, chunks))
    assert outs[0][3] == 0.0 and outs[0][4] == 0        # no replay launched
    assert all(o[4] >= 1 for o in outs[1:])              # replay launched
    for o in outs[1:]:
        for a, b in zip(outs[0][:3], o[:3]):
            np.testing.assert_array_equal(a, b)


def test_adjoint_vs_ddm3d_night_cells(cb6r5, oracle_cb6r5, solver):
    """GPU adjoint against the oracle's restatement of the CMAQ-DDM-3D
    chemistry step, contracted with the same seed vector, on night cells:
    the dominant sensitivity (largest |dL/dln C0|) agrees to 1 %.  The other
    entries carry DDM-3D's own error -- one Crank-Nicolson step per sync step
    for modes that are fast even at night (NO3/N2O5) -- and are reported."""
    import torch
    n = 64
    d = make_cells(cb6r5, np.arange(n), seed=3)
    night = ~d["rj"].any(axis=0)
    assert night.sum() >= 8
    rki = _rates(solver, d)
    y0 = _dev(d["y"])
    y = y0.clone()
    na = torch.zeros(n, dtype=torch.int32, device="cuda")
    solver.solve_device(5.0, rki, y, na, raise_on_fail=False)
    w = _seed(cb6r5, n)
    lam = _dev(w).clone()
    solver.adjoint_device(5.0, rki, y0, na, lam)
    lg = lam.cpu().numpy() * d["y"]
    kk = rki.cpu().numpy()
    worst = 0.0
    for c in np.flatnonzero(night)[:8]:
        s_ddm, _, _ = oracle_cb6r5.ddm_cell(5.0, kk[:, c], d["y"][:, c], w=w[:, c])
        j = int(np.argmax(np.abs(lg[:, c])))
        assert abs(lg[j, c] - s_ddm[j]) <= 1e-2 * abs(lg[j, c]), (c, j, lg[j, c], s_ddm[j])
        worst = max(worst, np.abs(lg[:, c] - s_ddm).max() / np.abs(lg[:, c]).max())
    print(f"DDM-3D vs adjoint, night cells: largest difference over all entries {worst:.3e} of the largest entry")


def test_adjoint_matches_oracle_cracmm2(cracmm2, oracle_cracmm2):
    """The larger mechanism (189 species, 531 reactions, tiled LU schedule):
    GPU adjoint against the oracle's on the same rate constants."""
    import torch
    from cmaq_amd.ros3 import Ros3Solver
    s = Ros3Solver("cracmm2", RTOL, ATOL)
    n = 256
    d = make_cells(cracmm2, np.arange(n), seed=26, workload="mixed", svoc_ppb=(1e-5, 5.0))
    rki = _rates(s, d)
    y0 = _dev(d["y"])
    y = y0.clone()
    na = torch.zeros(n, dtype=torch.int32, device="cuda")
    s.solve_device(5.0, rki, y, na, raise_on_fail=False)
    w = _seed(cracmm2, n)
    lam = _dev(w).clone()
    gk = torch.zeros_like(rki)
    s.adjoint_device(5.0, rki, y0, na, lam, gk)
    ref = oracle_cracmm2.adjoint_run(5.0, rki.cpu().numpy(), d["y"], w, RTOL, ATOL)
    nag, lg, gg = na.cpu().numpy(), lam.cpu().numpy(), gk.cpu().numpy()
    kk = rki.cpu().numpy()
    ok = (nag >= 0) & (ref["nacc"] == nag)
    assert ok.mean() >= 0.99
    for c in np.flatnonzero(ok):
        sl, rl = lg[:, c] * d["y"][:, c], ref["lam"][:, c] * d["y"][:, c]
        sg, rg = gg[:, c] * kk[:, c], ref["gk"][:, c] * kk[:, c]
        assert np.abs(sl - rl).max() <= 1e-8 * np.abs(rl).max(), c
        assert np.abs(sg - rg).max() <= 1e-8 * np.abs(rg).max(), c
    print(f"adjoint cracmm2 ({s.variant}): cells {n}, step counts equal {ok.mean():.4f}")


@pytest.mark.parametrize("name", ["cb6r5_ae7_aq", "saprc07tic_ae7i_aq"])
def test_tlm_matches_oracle_and_adjoint(name):
    """cmaq_ros3_tlm_device with three directions at once: the forward it
    carries equals solve_device bit for bit; each direction equals a
    one-direction call bit for bit, matches the oracle's om_tlm_cell (1e-8 of
    the cell's largest entry) and, on the GPU alone, the duality
    w . dy_end = lambda0 . dy0 + grad_rki . dk with the adjoint."""
    import torch
    from cmaq_amd.ros3 import Ros3Solver, load_mechanism
    from oracle.oracle import Oracle
    mech = load_mechanism(name)
    o = Oracle(mech)
    s = Ros3Solver(mech, RTOL, ATOL)
    n = 256
    d = make_cells(mech, np.arange(n), seed=31, workload="mixed")
    rki = _rates(s, d)
    kk = rki.cpu().numpy()
    rng = np.random.default_rng(3)
    names = mech["chemistry_spc"]
    ino2 = names.index("NO2")
    dy0 = np.stack([d["y"] * rng.normal(size=d["y"].shape),
                    np.where(np.arange(len(names))[:, None] == ino2, d["y"], 0.0),
                    np.zeros_like(d["y"])])
    dk = np.stack([kk * rng.normal(size=kk.shape), np.zeros_like(kk), kk * (rng.uniform(size=kk.shape) < 0.3)])
    y_f, na_f = _dev(d["y"]).clone(), torch.zeros(n, dtype=torch.int32, device="cuda")
    s.solve_device(5.0, rki, y_f, na_f, raise_on_fail=False)
    y_t, na_t = _dev(d["y"]).clone(), torch.zeros(n, dtype=torch.int32, device="cuda")
    dy = _dev(dy0).clone()
    s.tlm_device(5.0, rki, y_t, dy, _dev(dk), nacc=na_t, raise_on_fail=False)
    assert torch.equal(y_f, y_t) and torch.equal(na_f, na_t)
    dyg, nag = dy.cpu().numpy(), na_t.cpu().numpy()
    assert np.isnan(dyg[:, :, nag < 0]).all()
    for p in range(3):   # one direction per call gives the same bits
        y1, d1 = _dev(d["y"]).clone(), _dev(dy0[p]).clone()
        s.tlm_device(5.0, rki, y1, d1, _dev(dk[p]), raise_on_fail=False)
        assert np.array_equal(d1.cpu().numpy(), dyg[p], equal_nan=True), p
    w = _seed(mech, n)
    lam = _dev(w).clone()
    gk = torch.zeros_like(rki)
    s.adjoint_device(5.0, rki, _dev(d["y"]), na_t, lam, gk)
    lg, gg = lam.cpu().numpy(), gk.cpu().numpy()
    ok = np.flatnonzero(nag >= 0)
    for c in ok[:32]:
        h, _ = o.steps_cell(5.0, kk[:, c], d["y"][:, c])
        assert len(h) == nag[c]
        for p in range(3):
            ref = o.tlm_cell(kk[:, c], d["y"][:, c], h, dy0[p, :, c], dk[p, :, c])
            assert np.abs(dyg[p, :, c] - ref).max() <= 1e-8 * np.abs(ref).max(), (c, p)
    for c in ok:
        for p in range(3):
            lhs = w[:, c] @ dyg[p, :, c]
            rhs = lg[:, c] @ dy0[p, :, c] + gg[:, c] @ dk[p, :, c]
            scale = np.abs(w[:, c]).max() * np.abs(dyg[p, :, c]).sum() + 1e-300
            assert abs(lhs - rhs) <= 1e-9 * scale, (c, p, lhs, rhs)
    with pytest.raises(Exception):
        s.tlm_device(5.0, rki, _dev(d["y"]).clone(), _dev(np.zeros((9,) + d["y"].shape)))
    print(f"tlm {name}: cells {n}, 3 directions, failed {int((nag < 0).sum())}, checked vs oracle {min(32, ok.size)}")


def test_forward_mode_autograd_matches_reverse(cb6r5, solver):
    """torch forward-mode AD through integrate (tangent-linear kernel) against
    reverse mode (adjoint kernel) over 2 sync steps: w . jvp = grad . tangent."""
    import torch
    import torch.autograd.forward_ad as fwAD
    from cmaq_amd.adjoint import integrate
    n = 128
    d = make_cells(cb6r5, np.arange(n), seed=41, workload="mixed")
    rki = _rates(solver, d)
    y0 = _dev(d["y"])
    rng = np.random.default_rng(7)
    ty = _dev(d["y"] * rng.normal(size=d["y"].shape))
    tk = rki * _dev(rng.normal(size=tuple(rki.shape)))
    with fwAD.dual_level():
        out = integrate(solver, fwAD.make_dual(y0, ty), fwAD.make_dual(rki, tk), 5.0, 2)
        jv = fwAD.unpack_dual(out).tangent
    w = _dev(_seed(cb6r5, n))
    y0g, kg = y0.clone().requires_grad_(True), rki.clone().requires_grad_(True)
    (integrate(solver, y0g, kg, 5.0, 2) * w).sum().backward()
    lhs = (w * jv).sum(dim=0).cpu().numpy()
    rhs = ((y0g.grad * ty).sum(dim=0) + (kg.grad * tk).sum(dim=0)).cpu().numpy()
    ok = np.isfinite(lhs) & np.isfinite(rhs)
    assert ok.mean() >= 0.99
    scale = np.abs(jv.cpu().numpy()).sum(axis=0)[ok] + 1e-300
    assert (np.abs(lhs[ok] - rhs[ok]) <= 1e-9 * scale).all()
