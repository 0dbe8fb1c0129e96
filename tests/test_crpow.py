"""The correctly rounded pow of the device's SMVGEAR step control
(cmaq_amd/csrc/ros3_pow.h), compiled for the host from the same source.

Pinned against two independent correctly rounded references: Python's
decimal module at 60 digits (correctly rounded power, then one rounding to
double) and libquadmath's 113-bit powq rounded to double.  The host build
performs the same IEEE operations as the device build (explicit fma, no
contraction, correctly rounded division), so its results are the device's.
"""
import os
import shutil
import subprocess
from decimal import Decimal, getcontext

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, os.pardir, "cmaq_amd", "csrc")


@pytest.fixture(scope="module")
def crpow_bin(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not found")
    out = str(tmp_path_factory.mktemp("crpow") / "crpow_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-I", CSRC, "-o", out,
                    os.path.join(HERE, "native", "crpow_check.cpp"), "-lquadmath", "-lm"], check=True)
    return out


def _pow_cr(binary, pairs):
    text = "".join(f"{float(x).hex()} {float(y).hex()}\n" for x, y in pairs)
    out = subprocess.run([binary], input=text, capture_output=True, text=True, check=True).stdout.split()
    return [float.fromhex(v) for v in out]


def test_matches_decimal_rounding(crpow_bin):
    getcontext().prec = 60
    rng = np.random.default_rng(3)
    ys = [float(np.float32(1.0 / k)) for k in (2, 3, 4, 5, 6)]
    pairs = [(float(np.exp(rng.uniform(-60, 60))), ys[i % 5] if i % 2 else float(rng.uniform(-4, 4)))
             for i in range(20000)]
    got = _pow_cr(crpow_bin, pairs)
    exact = [float(Decimal(x) ** Decimal(y)) for x, y in pairs]
    assert got == exact


def test_sweep_against_quad_precision(crpow_bin):
    """One million pairs over the step control's range (x = error ratios from
    e^-70 to e^70, y = 1/(q+1) as REAL(4)), near 1 with large exponents, and
    across binades: no result differs from powq's; libm's pow (glibc's,
    accurate to 0.52 ulp) misrounds a few in ten thousand of them, the
    difference a device pow must not add to."""
    out = subprocess.run([crpow_bin, "--sweep", "1000000", "12345"], capture_output=True, text=True,
                         check=True).stdout
    last = out.strip().splitlines()[-1].split()
    n, bad, bad_libm = int(last[1]), int(last[3]), int(last[5])
    assert n == 1000000 and bad == 0, out
    assert bad_libm < n // 100


def test_special_arguments(crpow_bin):
    cases = [(0.0, 0.5), (1.0, 0.2), (2.0, 0.0), (4.0, 0.5), (0.25, 0.5), (1e-320, 0.5), (float("inf"), 0.5),
             (2.0, 1100.0), (0.5, 1100.0), (-8.0, 1.0 / 3.0)]
    got = _pow_cr(crpow_bin, cases)
    with np.errstate(all="ignore"):
        want = [float(np.power(np.float64(x), np.float64(y))) for x, y in cases]
    for (x, y), g, e in zip(cases, got, want):
        assert g == e or (np.isnan(g) and np.isnan(e)), (x, y, g, e)
