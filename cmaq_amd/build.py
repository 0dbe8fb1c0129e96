"""Build libcmaq_ros3.so: generate the per-mechanism kernels and compile
everything for gfx950 with hipcc, in-tree.

    python -m cmaq_amd.build            # all mechanisms under cmaq_amd/mechs
"""

import concurrent.futures as cf
import json
import os
import subprocess
import sys

from .codegen.emit_hip import emit, emit_gear

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
GEN = os.path.join(CSRC, "gen")
OBJ = os.path.join(GEN, "obj")
LIB = os.path.join(HERE, "libcmaq_ros3.so")
PROBE = os.path.join(HERE, "hbm_rows")       # memory-stream throughput probe (bench.py)
MECHS = {n: n + ".json" for n in (
    "cb6r5_ae7_aq", "cracmm2", "cb6r3_ae7_aq", "cb6r5m_ae7_aq", "cracmm1_aq", "cracmm1amore_aq",
    "racm2_ae6_aq", "saprc07tc_ae6_aq", "saprc07tic_ae7i_aq")}
# code-generation variants per mechanism: (register budget, load lookahead,
# LU schedule of programs.jac_lu_ops with options "+k1" (stage-1 forward
# substitution fused into the LU), "+step" (one function per attempted step),
# "+ldsN" (N doubles per lane of LDS as a second register tier) and "+sbN" (a
# scheduling barrier every N statements)); the first is the default unless
# tuning.json names another (written by cmaq_amd/autotune.py from timings on
# the GPU).  The table keeps the winners; tuning.json records the others that
# were measured.
VARIANTS = {
    "cb6r5_ae7_aq": [(170, 96, "ctile10x10+k1"), (170, 96, "tile10x10+k1"), (170, 96, "ctile14x10+k1")],
    "cracmm2": [(170, 96, "ctile12x10+k1"), (170, 96, "tile12x12+k1")],
    "cb6r3_ae7_aq": [(170, 96, "ctile14x14+k1")],
    "cb6r5m_ae7_aq": [(170, 96, "ctile8x8+k1")],
    "racm2_ae6_aq": [(170, 96, "ctile14x10+k1")],
    "cracmm1_aq": [(170, 96, "ctile14x10+k1")],
    "cracmm1amore_aq": [(170, 96, "ctile14x14+k1")],
    "saprc07tc_ae6_aq": [(170, 96, "ctile8x8+k1")],
    "saprc07tic_ae7i_aq": [(170, 96, "ctile10x10+k1")],
}


def variant_name(b, l, order):
    base, *extra = order.split("+")
    if base == "column":
        suffix = ""
    elif base.startswith("ctile"):
        suffix = "_ct" + base[5:]
    else:
        suffix = "_t" + base[4:]
    for e in extra:
        suffix += "_" + e
    return f"b{b}_l{l}{suffix}"

TUNING = os.path.join(HERE, "tuning.json")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# -ffp-contract=off: the reference rounds every multiply and add separately.
# packed FP32 is disabled for all device code on this platform.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
             "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
# generated kernels: skip the machine scheduler's two re-scheduling stages
# (high register pressure, low occupancy).  They take half the compile time of
# the long straight-line functions (SAPRC07TIC: 637 s -> 227 s) and the
# kernels run as fast without them (CB6r5: 36.6 vs 36.5 M cell-steps/s).
KERNEL_FLAGS = ["-mllvm", "-amdgpu-disable-unclustered-high-rp-reschedule",
                "-mllvm", "-amdgpu-disable-clustered-low-occupancy-reschedule",
                # no `tail` call markers: LLVM then lets every generated
                # (internal, non-recursive) function clobber the callee-saved
                # registers (interprocedural register allocation), instead of
                # saving and restoring up to ~300 of them through scratch
                # memory on every call -- which it does as soon as one call to
                # it is marked `tail`
                "-fno-optimize-sibling-calls"]


def load_mech(name):
    with open(os.path.join(HERE, "mechs", MECHS[name])) as f:
        return json.load(f)


def _write_if_changed(path, text):
    if os.path.exists(path) and open(path).read() == text:
        return False
    with open(path, "w") as f:
        f.write(text)
    return True


def _tuning_table():
    if os.path.exists(TUNING):
        with open(TUNING) as f:
            return json.load(f)
    return {}


def variants_of(name):
    """The variants built for a mechanism: the hand-kept table above plus the
    knobs the autotuner chose (tuning.json "knobs": register budget,
    lookahead, LU schedule), which need not be in the table."""
    vs = list(VARIANTS[name])
    knobs = _tuning_table().get(name, {}).get("knobs")
    if knobs:
        k = (int(knobs[0]), int(knobs[1]), str(knobs[2]))
        if k not in vs:
            vs.insert(0, k)
    return vs


def tuned(name):
    """(default variant, its refill threshold) for a mechanism."""
    entry = _tuning_table().get(name, {})
    built = [variant_name(*v) for v in variants_of(name)]
    var = entry.get("variant")
    if var not in built:       # a variant no longer generated: the table's first
        var = built[0]
    return var, int(entry.get("refill_min", 1))


def generate(names=None, variants=None, gen=GEN, obj=OBJ, gear=True):
    """Writes the generated sources; `variants` (mechanism -> knob list)
    replaces the built set (the autotuner's candidate library)."""
    os.makedirs(gen, exist_ok=True)
    names = names or list(MECHS)
    srcs, cnames, keep = [], [], set()
    for name in names:
        mech = load_mech(name)
        dflt, refill = tuned(name)
        for (b, l, order) in (variants[name] if variants else variants_of(name)):
            var = variant_name(b, l, order)
            cname = f"{name}_{var}"
            units = []
            gear_nws = 0
            if var == dflt and gear:
                # the SMVGEAR kernel, once per mechanism, on the default's schedule
                text_gear, gear_nws = emit_gear(mech, cname, budget=b, lookahead=l, jac_order=order)
                units.append(("_gear", text_gear))
            (text, text_adj), _ = emit(mech, cname, budget=b, lookahead=l, variant=var,
                                       is_default=(var == dflt), refill_min=refill, jac_order=order,
                                       gear_nws=gear_nws)
            for suffix, t in [("", text), ("_adj", text_adj)] + units:
                path = os.path.join(gen, f"ros3_{cname}{suffix}.hip")
                _write_if_changed(path, t)
                srcs.append((path, KERNEL_FLAGS))
                keep.add(f"ros3_{cname}{suffix}")
            cnames.append(cname)
    if set(names) == set(MECHS) or variants:
        # drop sources and objects of units no longer generated (variants
        # out of the table, the SMVGEAR unit of a former default)
        for d, ext in ((gen, ".hip"), (obj, ".o")):
            for f in os.listdir(d) if os.path.isdir(d) else []:
                stem = f[:f.index(".")] if "." in f else f
                if f.startswith("ros3_") and f.endswith((ext, ext + ".cmd")) and stem not in keep:
                    os.remove(os.path.join(d, f))
    reg = ["// Generated by cmaq_amd/build.py; do not edit.", "#include \"../ros3_registry.h\""]
    for name in cnames:
        reg.append(f"extern \"C\" const ros3::MechKernels ros3_kernels_{name};")
    reg.append("extern \"C\" const ros3::MechKernels *const ros3_mech_registry[] = {")
    reg += [f"    &ros3_kernels_{name}," for name in cnames]
    reg.append("};")
    reg.append(f"extern \"C\" const int ros3_mech_registry_size = {len(cnames)};")
    path = os.path.join(gen, "registry.cpp")
    _write_if_changed(path, "\n".join(reg) + "\n")
    srcs.append((path, []))
    return srcs


def _compile(src, obj, host_only=False, extra=(), force=False):
    if host_only:
        cmd = ["g++", "-O2", "-fPIC", "-std=c++17", "-c", src, "-o", obj]
    else:
        cmd = [HIPCC] + HIP_FLAGS + list(extra) + ["-c", src, "-o", obj]
    # an object is reused only if it is newer than its sources and was built
    # with the same command
    stamp = obj + ".cmd"
    same_cmd = os.path.exists(stamp) and open(stamp).read() == " ".join(cmd)
    deps = [src] + [os.path.join(CSRC, f) for f in ("ros3_device.h", "ros3_registry.h", "ros3_solver_kernel.inc",
                                                    "ros3_adjoint_kernel.inc", "ros3_gear.h",
                                                    "ros3_gear_kernel.inc", "ros3_pow.h")] + \
        [os.path.join(HERE, "..", "include", "cmaq_ros3.h")]
    if src.endswith("cmaq_ros3.cpp"):
        deps.append(os.path.join(CSRC, "ros3_chem_io.h"))
    if not force and same_cmd and os.path.exists(obj) and \
            all(os.path.getmtime(obj) >= os.path.getmtime(d) for d in deps if os.path.exists(d)):
        return obj
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    with open(stamp, "w") as f:
        f.write(" ".join(cmd))
    return obj


def build(names=None, verbose=True, force=False, variants=None, lib=LIB, gen=GEN, obj=OBJ):
    """Generate and compile; force=True recompiles every object.  With
    `variants` (mechanism -> knob list) builds only those kernels, without
    SMVGEAR, into `lib` (the autotuner's candidate library)."""
    os.makedirs(obj, exist_ok=True)
    gen_srcs = generate(names, variants=variants, gen=gen, obj=obj, gear=variants is None)
    jobs = []
    for src, extra in gen_srcs:
        base = os.path.splitext(os.path.basename(src))[0]
        jobs.append((src, os.path.join(obj, base + ".o"), src.endswith("registry.cpp"), extra, force))
    jobs.append((os.path.join(CSRC, "cmaq_ros3.cpp"), os.path.join(obj, "cmaq_ros3.o"), False, (), force))
    # largest sources first, so that the longest compiles overlap the others
    jobs.sort(key=lambda j: -os.path.getsize(j[0]))
    workers = int(os.environ.get("MAX_JOBS", "8"))
    with cf.ThreadPoolExecutor(max_workers=max(1, min(workers, len(jobs)))) as ex:
        objs = list(ex.map(lambda j: _compile(*j), jobs))
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", lib + ".tmp"] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(lib + ".tmp", lib)
    if variants:
        if verbose:
            print(f"built {lib}")
        return lib
    cmd = [HIPCC] + HIP_FLAGS + ["-o", PROBE + ".tmp", os.path.join(CSRC, "hbm_rows.hip")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"probe build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(PROBE + ".tmp", PROBE)
    if verbose:
        print(f"built {LIB}")
    return LIB


# the autotuner's candidate library (cmaq_amd/autotune.py): its own sources,
# objects and .so, loaded with CMAQ_ROS3_LIB
TUNE_GEN = os.path.join(CSRC, "gen_tune")
TUNE_LIB = os.path.join(HERE, "libcmaq_ros3_tune.so")


SAN_DIR = os.path.join(HERE, "_build", "san")
SAN_LIB = os.path.join(SAN_DIR, "libcmaq_ros3_san.so")
# host code only: the device code of the same translation unit is built as usual
SAN_HOST_FLAGS = ["-g", "-Wno-unused-value", "-Xarch_host", "-fsanitize=address", "-Xarch_host",
                  "-fsanitize=undefined", "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host",
                  "-fno-omit-frame-pointer"]


def build_host_sanitized():
    """The host runtime (cmaq_ros3.cpp) with AddressSanitizer and
    UndefinedBehaviorSanitizer on its host code, linked with the objects of
    the regular build, as _build/san/libcmaq_ros3_san.so.  Load it with
    CMAQ_ROS3_LIB and LD_PRELOAD of clang's ASan runtime (asan_runtime())."""
    os.makedirs(SAN_DIR, exist_ok=True)
    objs = sorted(os.path.join(OBJ, f) for f in os.listdir(OBJ) if f.endswith(".o") and f != "cmaq_ros3.o")
    if not objs:
        raise RuntimeError("no kernel objects: run the regular build first")
    obj = _compile(os.path.join(CSRC, "cmaq_ros3.cpp"), os.path.join(SAN_DIR, "cmaq_ros3_san.o"),
                   extra=SAN_HOST_FLAGS)
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-Xarch_host", "-fsanitize=address",
           "-Xarch_host", "-fsanitize=undefined", "-o", SAN_LIB + ".tmp", obj] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(SAN_LIB + ".tmp", SAN_LIB)
    return SAN_LIB


def asan_runtime():
    """clang's shared ASan runtime of this ROCm (the one the sanitized host
    objects were compiled against), or None."""
    import glob
    hits = sorted(glob.glob(os.path.join(os.path.dirname(os.path.realpath(HIPCC)), "..", "lib", "llvm", "lib",
                                         "clang", "*", "lib", "linux", "libclang_rt.asan-x86_64.so")))
    hits += sorted(glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return hits[0] if hits else None


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if a != "--force"]
    build(args or None, force="--force" in sys.argv[1:])
