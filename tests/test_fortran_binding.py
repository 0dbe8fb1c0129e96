"""The Fortran side of the boundary: cmaq_amd/fortran/cmaq_ros3_c.F90, the
ISO_C_BINDING module a gas/ros3_gpu CHEM driver uses, compiled with amdflang
in strict Fortran 2018 mode together with a test driver
(tests/fortran/ros3_fdriver.F90) that calls libcmaq_ros3.so through it.

CPU: every BIND(C) type has the size and field offsets of its C struct in
include/cmaq_ros3.h; cmaq_ros3_init called from Fortran refuses SPECIAL rate
operators with CMAQ_ROS3_E_MECH and, on a machine without a GPU, returns
CMAQ_ROS3_E_HIP with the library's message.  GPU: the driver's
cmaq_ros3_run over 256 cells matches the oracle with the solver's parity bar,
and a failing cell's LOGDEV dump names the species rbsolver.F:532-583 would.
"""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "tests", "fortran")
DRIVER = os.path.join(FDIR, "ros3_fdriver")
FLANG = "/opt/rocm/bin/amdflang"
RTOL, ATOL = 1.0e-3, 1.0e-7


def write_tables(path, m, rtol=RTOL, atol=ATOL):
    """The RXNS_DATA tables as a stream file in the driver's read order,
    every 2-D table in Fortran (column-major) order."""
    ns, nr = m["numb_mech_spc"], m["nrxns"]
    lists = [m[k] for k in ("nrxwm", "nrxww", "nrxwo2", "nrxwn2", "nrxwch4", "nrxwh2")]

    def i32(a, fortran2d=False):
        a = np.asarray(a, np.int32)
        if a.size == 0:
            a = np.zeros(1, np.int32)
        return (a.T if fortran2d else a).tobytes()

    def f64(a, fortran2d=False):
        a = np.asarray(a, np.float64)
        return (a.T if fortran2d else a).tobytes()

    def names(lst, n):
        out = [s.encode()[:16].ljust(16) for s in lst]
        out += [b" " * 16] * (n - len(out))
        return b"".join(out)

    with open(path, "wb") as f:
        f.write(np.array([ns, nr, m["mxprd"], m["kunits"], m["nfalloff"], m["nmphot"], m["nphotab"],
                          m["mhetero"], m["nhetero"]] + [len(x) for x in lists], np.int32).tobytes())
        f.write(i32(m["irr"], True) + f64(m["sc"], True))
        for k in ("nreact", "nprdct", "ktype", "irxbits", "iorder"):
            f.write(i32(m[k]))
        f.write(f64(m["rtdat"]))                      # RTDAT(3, NRXNS): rows are already Fortran order
        f.write(i32(m["irrfall"]))
        f.write(f64(m["rfdat"]) if m["nfalloff"] else np.zeros(5).tobytes())
        f.write(i32(m["iph"], True) + i32(m["ihetero"], True))
        for x in lists:
            f.write(i32(x))
        f.write(np.array([m["atm_air"], m["atm_o2"], m["atm_n2"], m["atm_ch4"], m["atm_h2"]]).tobytes())
        f.write(np.broadcast_to(np.float64(rtol), (ns,)).tobytes() + np.broadcast_to(np.float64(atol), (ns,)).tobytes())
        f.write(names(m["chemistry_spc"], ns) + names(m["photab"], max(m["nphotab"], 1)) +
                names(m["rxlabel"], nr))


def write_inputs(path, d, chemstep=5.0):
    n = d["temp"].size
    with open(path, "wb") as f:
        f.write(np.int64(n).tobytes() + np.float64(chemstep).tobytes())
        for k in ("temp", "pres", "h2o", "rj", "y"):
            f.write(np.ascontiguousarray(d[k], np.float64).tobytes())   # [k][n] = Fortran (n, k)


def read_outputs(path, ns, n):
    raw = open(path, "rb").read()
    rc = int(np.frombuffer(raw[:4], np.int32)[0])
    y = np.frombuffer(raw[4:4 + 8 * ns * n], np.float64).reshape(ns, n)
    rest = np.frombuffer(raw[4 + 8 * ns * n:], np.int32)
    return rc, y, rest[:n], rest[n:2 * n]


def _driver(tmp_path_factory):
    """The driver built by __graft_entry__.build(), or built here."""
    if not os.path.exists(FLANG):
        pytest.skip("amdflang not available")
    if not os.path.exists(DRIVER):
        subprocess.run(["make", "-s", "-C", FDIR], check=True)
    return DRIVER


def _run(args, **kw):
    r = subprocess.run([DRIVER] + args, capture_output=True, text=True, timeout=600, **kw)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_module_compiles_as_strict_free_form_f2018(tmp_path):
    if not os.path.exists(FLANG):
        pytest.skip("amdflang not available")
    src = os.path.join(ROOT, "cmaq_amd", "fortran", "cmaq_ros3_c.F90")
    r = subprocess.run([FLANG, "-std=f2018", "-Werror", "-module-dir", str(tmp_path), "-c", src, "-o",
                        str(tmp_path / "m.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "cmaq_ros3_c.mod").exists()


def test_types_match_the_c_header(tmp_path, tmp_path_factory):
    _driver(tmp_path_factory)
    got = {}
    for line in _run(["sizes"]).splitlines():
        kind, name, v = line.split()
        got[(kind, name)] = int(v)
    csrc = tmp_path / "sizes.c"
    fields = [k for k in got if k[0] == "offset"]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "cmaq_ros3.h"', "int main(void) {"]
    for (kind, name) in got:
        if kind == "sizeof":
            ctype = "cmaq_gear_tol" if name == "gear_tol" else f"cmaq_ros3_{name}"
            lines.append(f'printf("sizeof {name} %zu\\n", sizeof({ctype}));')
    for (_, name) in fields:
        t, fld = name.split(".")
        ctype = "cmaq_gear_tol" if t == "gear_tol" else f"cmaq_ros3_{t}"
        lines.append(f'printf("offset {name} %zu\\n", offsetof({ctype}, {fld}));')
    lines.append("return 0; }")
    csrc.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(csrc), "-o", str(exe)], check=True)
    want = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        kind, name, v = line.split()
        want[(kind, name)] = int(v)
    assert got == want
    assert len(fields) >= 20


def _has_gpu():
    from conftest import has_gpu
    return has_gpu()


def test_init_error_paths_through_fortran(tmp_path, tmp_path_factory, cb6r5):
    _driver(tmp_path_factory)
    tab = tmp_path / "cb6r5.tab"
    write_tables(tab, cb6r5)
    out = _run(["init", str(tab), "1"])                 # NSPECIAL_RXN = 1
    assert "rc -2" in out and "SPECIAL" in out
    if not _has_gpu():
        out = _run(["init", str(tab)])
        assert "rc -3" in out and "message " in out and len(out.split("message ", 1)[1].strip()) > 0


@pytest.mark.gpu
def test_fortran_driver_run_matches_oracle(tmp_path, tmp_path_factory, cb6r5, oracle_cb6r5):
    """256 config-0 cells through the Fortran driver's cmaq_ros3_run, against
    the oracle at the parity bar; then cells with ROR drawn log-uniform (which
    the input rule starts at 1e-6 ppb) fail, and the driver's LOGDEV dump
    names PAR as the species holding most of ERR."""
    import cmaq_amd.inputs as I
    _driver(tmp_path_factory)
    m, o = cb6r5, oracle_cb6r5
    tab, fin, fout, flog = (tmp_path / x for x in ("t.tab", "in.bin", "out.bin", "fail.log"))
    write_tables(tab, m)
    n = 256
    d = I.make_cells(m, np.arange(n), seed=5)
    write_inputs(fin, d)
    print(_run(["run", str(tab), str(fin), str(fout), str(flog)]), flush=True)
    rc, y, nacc, nrej = read_outputs(fout, m["numb_mech_spc"], n)
    ref = o.run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"], RTOL, ATOL)
    assert rc == 0 and ref["rc"] == 0
    tol = 0.1 * (RTOL * np.abs(ref["y"]) + ATOL)
    assert int((np.abs(y - ref["y"]) > tol).sum()) == 0
    assert (nacc == ref["nacc"]).mean() >= 0.999
    # a failing cell and its dump
    iror = m["chemistry_spc"].index("ROR")
    d2 = I.make_cells(m, np.arange(64), seed=0)
    src = I.make_cells(m, np.array([1841]), seed=0)
    for k in ("temp", "pres", "h2o"):
        d2[k][7] = src[k][0]
    d2["rj"][:, 7] = src["rj"][:, 0]
    d2["y"][:, 7] = src["y"][:, 0]
    d2["y"][iror, 7] = I._logu(I.uniform(0, I.F_SPC + iror, np.array([1841])), 1e-6, 1e2)[0] * 1e-3
    write_inputs(fin, d2)
    out = _run(["run", str(tab), str(fin), str(fout), str(flog)])
    rc, y, nacc, nrej = read_outputs(fout, m["numb_mech_spc"], 64)
    assert rc == -4 and nacc[7] == -1 and (nacc[np.arange(64) != 7] > 0).all()
    assert "dump -4" in out
    text = flog.read_text()
    print(text[:3000])
    rows = []
    for line in text.splitlines():
        parts = line.split()
        if len(parts) == 5 and parts[0].isdigit() and parts[1] in m["chemistry_spc"]:
            rows.append((float(parts[2]), parts[1]))
    assert len(rows) == m["numb_mech_spc"]
    assert max(rows)[1] == "PAR" and max(rows)[0] > 0.9
    assert abs(sum(r[0] for r in rows) - 1.0) < 1e-6       # the shares of ERR add up to 1
    assert "in cell       8" in text


@pytest.mark.gpu
def test_fortran_driver_gear_matches_oracle(tmp_path, tmp_path_factory, cb6r5, oracle_cb6r5):
    """256 cells through the Fortran driver's cmaq_gear_set_tolerances +
    cmaq_gear_run (SMVGEAR), against the oracle's grsmvgear restatement."""
    import cmaq_amd.inputs as I
    _driver(tmp_path_factory)
    m, o = cb6r5, oracle_cb6r5
    tab, fin, fout = (tmp_path / x for x in ("t.tab", "in.bin", "out.bin"))
    write_tables(tab, m)
    n = 256
    d = I.make_cells(m, np.arange(n), seed=8, workload="clean")
    write_inputs(fin, d)
    print(_run(["gear", str(tab), str(fin), str(fout)]), flush=True)
    rc, y, nsteps, nfail = read_outputs(fout, m["numb_mech_spc"], n)
    ref = o.gear_run(5.0, d["temp"], d["pres"], d["h2o"], d["rj"], d["y"])
    assert rc == 0 and ref["rc"] == 0
    from test_oracle_gear import gear_bar
    ok, st = gear_bar(ref["y"], y, ref["nsteps"], nsteps, rates_differ=True)     # the library's own CALC_RCONST
    assert ok, st
