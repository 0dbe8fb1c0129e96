"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer, on the CPU.

* The oracle (oracle/ros3_oracle.c, `make -C oracle san`, gcc's sanitizers):
  its own tests -- sparse-matrix setup of every mechanism, rates, ROS3 runs,
  adjoint, tangent-linear, DDM-3D and the CGRID gather/scatter -- run against
  the sanitized build in a child process with gcc's libasan preloaded.
* The host runtime (cmaq_amd/csrc/cmaq_ros3.cpp, host code only, clang's
  sanitizers): the C-ABI tests, which drive cmaq_ros3_init's table checks,
  structure hashing, rate-program decoding and DDM table setup for every
  mechanism and its error paths up to the first HIP call, against the
  sanitized library with clang's ASan runtime preloaded.

Any report ends the child with a non-zero status (halt_on_error,
-fno-sanitize-recover=all).  Device code is not sanitized: GPU sanitizers are
not available for these kernels.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_ENV = {"ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=0",
           "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}


def _child(cmd, preload, extra_env):
    env = dict(os.environ, LD_PRELOAD=preload, **SAN_ENV, **extra_env)
    return subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)


def _loaded(preload, extra_env, module, name):
    """The child really maps the sanitized library."""
    code = (f"import {module} as m; m.lib(); "
            f"print(sum('{name}' in l for l in open('/proc/self/maps')))")
    r = _child([sys.executable, "-c", code], preload, extra_env)
    assert r.returncode == 0, r.stderr[-3000:]
    assert int(r.stdout.split()[-1]) > 0, r.stdout


def test_sanitizer_harness_reports_a_planted_overflow(tmp_path):
    """The preload and options above do report: a one-past-the-end read in a
    small C library built with the oracle's sanitizer flags ends the child
    with a heap-buffer-overflow report."""
    if shutil.which("gcc") is None:
        pytest.skip("gcc not found")
    src = tmp_path / "planted.c"
    src.write_text("#include <stdlib.h>\nint planted(int n) { int *a = malloc(n * sizeof(int));"
                   " for (int i = 0; i < n; ++i) a[i] = i; int v = a[n]; free(a); return v; }\n")
    so = tmp_path / "libplanted.so"
    r = subprocess.run(["gcc", "-O1", "-g", "-fPIC", "-shared", "-fsanitize=address,undefined",
                        "-fno-sanitize-recover=all", "-o", str(so), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    libasan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                             text=True).stdout.strip()
    r = _child([sys.executable, "-c", f"import ctypes; ctypes.CDLL('{so}').planted(8)"], libasan, {})
    assert r.returncode != 0 and "heap-buffer-overflow" in r.stderr, r.stderr[-2000:]


def test_oracle_under_asan_ubsan():
    if shutil.which("gcc") is None:
        pytest.skip("gcc not found")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    libasan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                             text=True).stdout.strip()
    assert os.path.isabs(libasan), libasan
    env = {"CMAQ_ORACLE_LIB": os.path.join(ROOT, "oracle", "_build", "liboracle_ros3_san.so")}
    _loaded(libasan, env, "oracle.oracle", "liboracle_ros3_san.so")
    files = ["tests/test_oracle.py", "tests/test_oracle_adjoint.py", "tests/test_oracle_tlm.py",
             "tests/test_oracle_ddm.py", "tests/test_chem_io.py"]
    r = _child([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider"] + files, libasan, env)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr


def test_host_runtime_under_asan_ubsan():
    from cmaq_amd import build
    from cmaq_amd.ros3 import LIB_PATH
    if not os.path.exists(LIB_PATH) or not os.path.isdir(build.OBJ):
        pytest.skip("library not built (python -m cmaq_amd.build)")
    rt = build.asan_runtime()
    if rt is None:
        pytest.skip("clang ASan runtime not found")
    lib = build.build_host_sanitized()
    env = {"CMAQ_ROS3_LIB": lib}
    _loaded(rt, env, "cmaq_amd.ros3", "libcmaq_ros3_san.so")
    r = _child([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "tests/test_abi.py"], rt, env)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
