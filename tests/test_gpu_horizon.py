"""Parity and gradients over the workloads' real horizon: 12 chemistry sync
steps of 5 min (one hour), integrated on the GPU without restarting from the
oracle, against the oracle's own 12-step chain on the same inputs.

* Uneven stiffness (1,048,576 mixed cells in the benchmark; 4096 here): the
  hipGraph sync loop's per-step checkpoints and step counts against 12
  chained oracle steps.
* Differentiation (262,144 mixed cells in the benchmark; 256 here): the
  autograd adjoint (stored tape) over the hour against the oracle's adjoint
  chained backwards through the same 12 steps (all entries, 1e-8 of the
  largest per cell), and its top-8 dL/dln y0 and dL/dln k entries per cell
  against central finite differences (h = 1e-6 relative) of the oracle's
  fixed-step replay of the hour (1e-5 of the largest).
* The larger mechanism (CRACMM2, 2,097,152 mixed cells with the VOC/SVOC
  species in the benchmark; 256 here): the replay-mode adjoint over the hour
  against the oracle's.

Loss = sum over cells of final O3 + 0.1 final NO2.  The oracle runs in
threads (its C calls release the GIL).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from cmaq_amd.inputs import make_cells

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1.0e-3, 1.0e-7
NSYNC = 12
THREADS = 8


def _dev(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def _chunks(n, k=THREADS):
    b = np.linspace(0, n, k + 1).astype(int)
    return [slice(b[i], b[i + 1]) for i in range(k) if b[i + 1] > b[i]]


def _oracle_chain(o, d, nsync):
    """nsync chained oracle sync steps from d's inputs: states y_s (s = 0..nsync)
    and per-step accepted counts."""
    n = d["temp"].size
    ys, counts = [d["y"].copy()], []
    for _ in range(nsync):
        y = ys[-1]
        out_y, out_n = np.empty_like(y), np.empty(n, np.int32)

        def run(sl):
            r = o.run(5.0, d["temp"][sl], d["pres"][sl], d["h2o"][sl], d["rj"][:, sl], y[:, sl], RTOL, ATOL)
            out_y[:, sl], out_n[sl] = r["y"], r["nacc"]

        with ThreadPoolExecutor(THREADS) as ex:
            list(ex.map(run, _chunks(n)))
        ys.append(out_y)
        counts.append(out_n)
    return ys, counts


def test_uneven_stiffness_hour_vs_oracle(cb6r5, oracle_cb6r5):
    """4096 mixed cells, 12 sync steps through the hipGraph loop, compared per
    step: cells whose accepted-step counts have matched so far (>= 99.9 % at
    every step) have every species within 0.1 * (RTOL*|c| + ATOL)."""
    import torch
    from cmaq_amd.ros3 import Ros3Solver
    n = 4096
    d = make_cells(cb6r5, np.arange(n), seed=2, workload="mixed")
    s = Ros3Solver(cb6r5, RTOL, ATOL)
    y = _dev(d["y"]).clone()
    ck = torch.empty((NSYNC,) + tuple(y.shape), dtype=torch.float64, device="cuda")
    na = torch.empty((NSYNC, n), dtype=torch.int32, device="cuda")
    f = s.run_steps_device(5.0, NSYNC, _dev(d["temp"]), _dev(d["pres"]), _dev(d["h2o"]), _dev(d["rj"]), y, ck, na,
                           raise_on_fail=False)
    torch.cuda.synchronize()
    ck, na = ck.cpu().numpy(), na.cpu().numpy()
    ys, counts = _oracle_chain(oracle_cb6r5, d, NSYNC)
    same = np.ones(n, bool)
    for k in range(NSYNC):
        same &= na[k] == counts[k]
        ref = ys[k + 1]
        tol = 0.1 * (RTOL * np.abs(ref) + ATOL)
        bad = int((np.abs(ck[k] - ref) > tol)[:, same].sum())
        print(f"sync step {k + 1}: step counts equal so far {same.mean():.5f}, species outside tol {bad}, "
              f"mean accepted {counts[k].mean():.1f}", flush=True)
        assert (counts[k] >= 0).all() and (na[k] >= 0).all()      # the input rule leaves no failing cell
        assert same.mean() >= 0.999
        assert bad == 0
    assert f is None


def _grad_setup(mech, solver, n, seed, svoc=None):
    import torch
    d = make_cells(mech, np.arange(n), seed=seed, workload="mixed", svoc_ppb=svoc)
    rki = torch.zeros((mech["nrxns"], n), dtype=torch.float64, device="cuda")
    solver.rates_device(_dev(d["temp"]), _dev(d["pres"]), _dev(d["h2o"]), _dev(d["rj"]), rki)
    # rate constants as the optimiser sees them: exp(log k), used on both sides
    logk = torch.log(rki).requires_grad_(True)
    return d, logk


def _gpu_hour_grads(mech, solver, d, logk, tape):
    import torch
    from cmaq_amd.adjoint import integrate
    names = mech["chemistry_spc"]
    y0 = _dev(d["y"]).requires_grad_(True)
    k_used = logk.exp()
    yend = integrate(solver, y0, k_used, 5.0, NSYNC, tape=tape)
    loss = (yend[names.index("O3")] + 0.1 * yend[names.index("NO2")]).sum()
    loss.backward()
    kk = k_used.detach().cpu().numpy()
    # the same forward's per-step counts (same kernel, same steps)
    y = _dev(d["y"]).clone()
    counts = []
    for _ in range(NSYNC):
        na = torch.empty(y.shape[1], dtype=torch.int32, device="cuda")
        solver.solve_device(5.0, k_used.detach(), y, na, raise_on_fail=False)
        counts.append(na.cpu().numpy())
    return y0.grad.cpu().numpy() * d["y"], logk.grad.cpu().numpy(), kk, counts


def _oracle_hour_adjoint(o, mech, d, kk):
    """The oracle's adjoint of the hour: forward chain with the given rate
    constants, then lam chained backwards; returns dL/dln y0, dL/dln k and
    the per-step counts (cells in threads)."""
    n = kk.shape[1]
    names = mech["chemistry_spc"]
    w = np.zeros((len(names), n))
    w[names.index("O3")] = 1.0
    w[names.index("NO2")] = 0.1
    lam_all, gk_all = np.empty_like(w), np.empty_like(kk)
    counts = np.empty((NSYNC, n), np.int32)

    def part(sl):
        k = np.ascontiguousarray(kk[:, sl])
        ys = [np.ascontiguousarray(d["y"][:, sl])]
        for s in range(NSYNC):
            r = o.adjoint_run(5.0, k, ys[-1], np.zeros_like(ys[-1]), RTOL, ATOL)
            ys.append(r["y"])
            counts[s, sl] = r["nacc"]
        lam, gk = np.ascontiguousarray(w[:, sl]), np.zeros_like(k)
        for s in range(NSYNC - 1, -1, -1):
            r = o.adjoint_run(5.0, k, ys[s], lam, RTOL, ATOL)
            lam, gk = r["lam"], gk + r["gk"]
        lam_all[:, sl], gk_all[:, sl] = lam, gk

    with ThreadPoolExecutor(THREADS) as ex:
        list(ex.map(part, _chunks(n)))
    return lam_all * d["y"], gk_all * kk, list(counts)


def _compare_adjoint(gy, gl, ref_y, ref_l, counts, ref_counts, label):
    same = np.ones(gy.shape[1], bool)
    for a, b in zip(counts, ref_counts):
        same &= (a == b) & (a >= 0)
    assert same.mean() >= 0.99, same.mean()
    worst = 0.0
    for c in np.flatnonzero(same):
        for g, r in ((gy[:, c], ref_y[:, c]), (gl[:, c], ref_l[:, c])):
            err = np.abs(g - r).max() / np.abs(r).max()
            worst = max(worst, err)
            assert err <= 1e-8, (label, c, err)
    print(f"{label}: cells with the oracle's steps in all {NSYNC} sync steps {same.mean():.4f}, "
          f"worst difference / largest entry {worst:.2e}", flush=True)
    return same


def test_differentiation_hour_vs_oracle_and_fd(cb6r5, oracle_cb6r5):
    from cmaq_amd.ros3 import Ros3Solver
    o, m = oracle_cb6r5, cb6r5
    s = Ros3Solver(m, RTOL, ATOL)
    n = 256
    d, logk = _grad_setup(m, s, n, seed=3)
    gy, gl, kk, counts = _gpu_hour_grads(m, s, d, logk, "stored")
    ref_y, ref_l, ref_counts = _oracle_hour_adjoint(o, m, d, kk)
    same = _compare_adjoint(gy, gl, ref_y, ref_l, counts, ref_counts, "CB6r5 stored-tape adjoint, 1 h")
    names = m["chemistry_spc"]
    w = np.zeros(len(names))
    w[names.index("O3")] = 1.0
    w[names.index("NO2")] = 0.1

    def check(c):
        k = kk[:, c]
        hs, y = [], d["y"][:, c]
        for _ in range(NSYNC):
            h, y = o.steps_cell(5.0, k, y, RTOL, ATOL)
            hs.append(h)

        def L(y, kv):
            for h in hs:
                y = o.replay_cell(kv, y, h)
            return w @ y

        worst = 0.0
        for j in np.argsort(-np.abs(gy[:, c]))[:8]:
            e = 1e-6 * d["y"][j, c]
            yp, ym = d["y"][:, c].copy(), d["y"][:, c].copy()
            yp[j] += e
            ym[j] -= e
            fd = (L(yp, k) - L(ym, k)) / (2 * e) * d["y"][j, c]
            worst = max(worst, abs(gy[j, c] - fd) / np.abs(gy[:, c]).max())
        for r in np.argsort(-np.abs(gl[:, c]))[:8]:
            e = 1e-6 * k[r]
            kp, km = k.copy(), k.copy()
            kp[r] += e
            km[r] -= e
            fd = (L(d["y"][:, c], kp) - L(d["y"][:, c], km)) / (2 * e) * k[r]
            worst = max(worst, abs(gl[r, c] - fd) / np.abs(gl[:, c]).max())
        return worst

    cells = np.flatnonzero(same)
    with ThreadPoolExecutor(THREADS) as ex:
        worst = list(ex.map(check, cells))
    print(f"FD of the fixed-step hour, {len(cells)} cells x top-8 y0 and k entries: worst {max(worst):.2e}")
    assert max(worst) <= 1e-5


def test_larger_mechanism_replay_adjoint_hour(cracmm2, oracle_cracmm2):
    from cmaq_amd.ros3 import Ros3Solver
    o, m = oracle_cracmm2, cracmm2
    s = Ros3Solver(m, RTOL, ATOL)
    n = 256
    d, logk = _grad_setup(m, s, n, seed=4, svoc=(1e-5, 5.0))
    gy, gl, kk, counts = _gpu_hour_grads(m, s, d, logk, "replay")
    ref_y, ref_l, ref_counts = _oracle_hour_adjoint(o, m, d, kk)
    _compare_adjoint(gy, gl, ref_y, ref_l, counts, ref_counts, "CRACMM2 replay adjoint, 1 h")
