"""C-ABI checks that need no GPU: the library loads, exports every function
that include/cmaq_ros3.h declares, and cmaq_ros3_init's host-side validation
(table checks, structure-hash lookup of the generated kernels) behaves as
documented before any device is touched."""
import copy
import os
import re

import pytest

from cmaq_amd import ros3
from cmaq_amd.codegen.emit_hip import structure_hash

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "cmaq_ros3.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(cmaq_(?:ros3|gear)_[a-z_]+)\s*\(", text)))


def _lib():
    if not os.path.exists(ros3.LIB_PATH):
        pytest.skip("libcmaq_ros3.so not built (python -m cmaq_amd.build)")
    return ros3.lib()


def test_exports_every_declared_symbol():
    L = _lib()
    names = _declared()
    assert len(names) >= 11
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_registry_holds_every_mechanism():
    """Every mechanism table shipped has generated kernels: the C++ lookup by
    structure hash matches the generator's (init fails with E_MECH only when
    no kernels carry the hash, and with E_HIP later when no device exists)."""
    from cmaq_amd.build import MECHS
    L = _lib()
    assert len(MECHS) == 9
    for name in MECHS:
        m = ros3.load_mechanism(name)
        rc, msg = _init_rc(L, m)
        assert rc in (0, -3), (name, rc, msg)   # ok, or no GPU in this process
        assert "no kernels" not in msg


def _init_rc(L, m):
    import ctypes as C
    s, keep = ros3.pack_mechanism(m)
    ctx = C.c_void_p()
    rc = L.cmaq_ros3_init(C.byref(ctx), C.byref(s), None, None, None)
    msg = L.cmaq_ros3_last_error(None).decode()
    if rc == 0:
        L.cmaq_ros3_free(ctx)
    del keep
    return rc, msg


def test_structural_change_is_rejected(cb6r5):
    """A table whose structure differs from every generated kernel is refused
    with E_MECH naming the hash the Python generator also computes."""
    L = _lib()
    m = copy.deepcopy(cb6r5)
    m["sc"][5][0] = m["sc"][5][0] + 0.125
    rc, msg = _init_rc(L, m)
    assert rc == -2
    assert f"{structure_hash(m):016x}" in msg


@pytest.mark.parametrize("field,value,expect", [
    ("numb_mech_spc", 0, "sizes"),
    ("nspecial_rxn", 3, "SPECIAL"),
])
def test_invalid_tables(cb6r5, field, value, expect):
    L = _lib()
    m = copy.deepcopy(cb6r5)
    m[field] = value
    s, keep = ros3.pack_mechanism(m)
    if field == "nspecial_rxn":
        s.nspecial_rxn = value
    import ctypes as C
    ctx = C.c_void_p()
    rc = L.cmaq_ros3_init(C.byref(ctx), C.byref(s), None, None, None)
    msg = L.cmaq_ros3_last_error(None).decode()
    assert rc in (-1, -2), (rc, msg)
    assert expect.lower() in msg.lower(), msg


@pytest.mark.parametrize("ktype", [11, 13])
def test_reserved_rate_types_rejected(cb6r5, ktype):
    """KTYPE 11 (a SPECIAL operator: a rate recomputed from the
    concentrations in every RBFEVAL) and 13 (Fortran text that chemmech
    writes into CALC_RCONST only, absent from RXNS_DATA) are refused rather
    than evaluated wrongly; KTYPE 7 is accepted."""
    L = _lib()
    m = copy.deepcopy(cb6r5)
    r = m["ktype"].index(3)
    m["ktype"][r] = ktype
    rc, msg = _init_rc(L, m)
    assert rc == -2 and "type" in msg.lower(), msg
    m["ktype"][r] = 7
    rc, msg = _init_rc(L, m)
    assert rc in (0, -3), msg


def test_python_binding_matches_header():
    """Every function the ctypes binding declares takes as many arguments as
    its prototype in include/cmaq_ros3.h."""
    L = _lib()
    text = open(HEADER).read()
    decls = re.findall(r"\n(?:int|void|const char \*)\s*(cmaq_(?:ros3|gear)_\w+)\s*\(([^;]*?)\);", text, re.S)
    assert len(decls) == len(_declared())
    for name, params in decls:
        n = 0 if params.strip() in ("", "void") else params.count(",") + 1
        at = getattr(L, name).argtypes
        assert at is not None, name
        assert len(at) == n, (name, len(at), n)


def test_library_carries_device_code_of_every_kernel():
    """Each variant is compiled as two translation units (forward kernels and
    registry record; reverse kernel).  The library's gfx950 code objects must
    hold the kernel descriptor of every kernel the registry record points at:
    the forward, tape-recording, tangent-linear and YAVE-accumulating
    instantiations of solve_kernel_t and reverse_kernel, and for each
    mechanism's default variant the SMVGEAR kernels (a third unit)."""
    from cmaq_amd.build import VARIANTS, tuned, variant_name, variants_of
    _lib()
    blob = open(ros3.LIB_PATH, "rb").read()
    missing = []
    for name in VARIANTS:
        for (b, l, order) in variants_of(name):
            c = f"{name}_{variant_name(b, l, order)}"
            ns = f"_ZN4ros3{len(c)}{c}"
            kds = [f"{ns}14solve_kernel_tILb{tape}ELb{tlm}ELb{yave}ELb{blk}EEEvNS_11SolveParamsE.kd"
                   for tape, tlm, yave, blk in ((0, 0, 0, 0), (1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0),
                                                (0, 0, 0, 1))]
            kds.append(f"{ns}14reverse_kernelENS_11SolveParamsE.kd")
            if variant_name(b, l, order) == tuned(name)[0]:
                g = c + "_gear"
                kds.append(f"_ZN4ros3{len(g)}{g}11gear_kernelENS_10GearParamsE.kd")
                kds.append(f"_ZN4ros3{len(g)}{g}4tail16gear_tail_kernelENS_10GearParamsE.kd")
            missing += [k for k in kds if k.encode() not in blob]
    assert not missing, missing
