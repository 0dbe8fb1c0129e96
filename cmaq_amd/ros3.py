"""Python host side of the gfx950 ROS3 solver (ctypes over libcmaq_ros3.so).

Mirrors the reference call sequence of the ros3 CHEM driver
(CCTM/src/gas/ros3/rbdriver.F): ``Ros3Solver(...)`` plays RBINIT + RBSPARSE
(rbinit.F:26, rbsparse.F:26) and ``run`` plays CALC_RCONST + RBSOLVER over all
cells (rbdriver.F:440-677), with the reference's units (ppm, minutes, 1/min,
atm) and "cell fastest" layouts.  Failures raise ``Ros3Error`` carrying the
code and the failing cell instead of stopping the process (the reference
calls M3EXIT, rbsolver.F:583).

The native library is required: nothing here falls back to a CPU path.
"""

import ctypes as C
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CMAQ_ROS3_LIB names another build of the same library (the host-sanitized
# one in the CPU sanitizer test)
LIB_PATH = os.environ.get("CMAQ_ROS3_LIB") or os.path.join(HERE, "libcmaq_ros3.so")

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)

E_CODES = {-1: "ARG", -2: "MECH", -3: "HIP", -4: "CONVERGENCE", -5: "STEPS", -6: "NOMEM", -7: "TAPE"}


class Ros3Error(RuntimeError):
    def __init__(self, code, msg, fail=None):
        super().__init__(f"cmaq_ros3 error {code} ({E_CODES.get(code, '?')}): {msg}")
        self.code = code
        self.fail = fail


class _Mech(C.Structure):
    _fields_ = [
        ("numb_mech_spc", C.c_int32), ("nrxns", C.c_int32), ("mxprd", C.c_int32), ("kunits", C.c_int32),
        ("irr", _i32p), ("sc", _f64p), ("nreact", _i32p), ("nprdct", _i32p), ("ktype", _i32p),
        ("irxbits", _i32p), ("iorder", _i32p), ("rtdat", _f64p), ("nfalloff", C.c_int32),
        ("irrfall", _i32p), ("rfdat", _f64p), ("nmphot", C.c_int32), ("nphotab", C.c_int32),
        ("iph", _i32p), ("mhetero", C.c_int32), ("nhetero", C.c_int32), ("ihetero", _i32p),
        ("nwm", C.c_int32), ("nww", C.c_int32), ("nwo2", C.c_int32), ("nwn2", C.c_int32),
        ("nwch4", C.c_int32), ("nwh2", C.c_int32),
        ("nrxwm", _i32p), ("nrxww", _i32p), ("nrxwo2", _i32p), ("nrxwn2", _i32p),
        ("nrxwch4", _i32p), ("nrxwh2", _i32p),
        ("atm_air", C.c_double), ("atm_o2", C.c_double), ("atm_n2", C.c_double),
        ("atm_ch4", C.c_double), ("atm_h2", C.c_double), ("nspecial_rxn", C.c_int32),
    ]


class _Opts(C.Structure):
    _fields_ = [("device", C.c_int32), ("max_attempts", C.c_int32), ("waves_per_cu", C.c_int32),
                ("refill_min", C.c_int32)]


class _Fail(C.Structure):
    _fields_ = [("code", C.c_int32), ("pad", C.c_int32), ("cell", C.c_int64), ("maxerr", C.c_double),
                ("dt", C.c_double), ("tnow", C.c_double)]


class _FailDetail(C.Structure):
    _fields_ = [("code", C.c_int32), ("numb_mech_spc", C.c_int32), ("nrxns", C.c_int32),
                ("nphotab", C.c_int32), ("cell", C.c_int64), ("maxerr", C.c_double), ("err", C.c_double),
                ("dt", C.c_double), ("tnow", C.c_double), ("temp_k", C.c_double), ("pres_atm", C.c_double),
                ("h2o_ppm", C.c_double), ("seawater", C.c_double), ("dens", C.c_double)]


class _GearTol(C.Structure):
    _fields_ = [("rtol", C.c_float), ("max_atol", C.c_float), ("min_atol", C.c_float), ("conc_floor", C.c_float)]


class _Grid(C.Structure):
    _fields_ = [("ncols", C.c_int32), ("nrows", C.c_int32), ("nlays", C.c_int32), ("ld_het", C.c_int32),
                ("cgrid", C.c_void_p), ("ta", C.c_void_p), ("dens", C.c_void_p), ("pres", C.c_void_p),
                ("qv", C.c_void_p), ("seawater_zone", C.c_void_p), ("seaice", C.c_void_p),
                ("rj", C.c_void_p), ("khetero", C.c_void_p)]


_lib = None


def lib():
    """Load the native library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m cmaq_amd.build` first")
        # One HIP runtime per process: when PyTorch is present its bundled
        # libamdhip64 (same SONAME) must be loaded first so that the solver
        # binds to it and shares device pointers and streams with torch.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.cmaq_ros3_init.argtypes = [C.POINTER(vp), C.POINTER(_Mech), _f64p, _f64p, C.POINTER(_Opts)]
        L.cmaq_ros3_run.argtypes = [vp, C.c_int64, C.c_double] + [_f64p] * 7 + [_i32p, _i32p, C.POINTER(_Fail)]
        L.cmaq_ros3_run_device.argtypes = [vp, C.c_int64, C.c_double] + [vp] * 9 + [C.POINTER(_Fail), vp]
        L.cmaq_ros3_run_blocks_device.argtypes = ([vp, C.c_int64, C.c_int32, C.c_double] + [vp] * 9 +
                                                  [C.POINTER(_Fail), vp])
        L.cmaq_ros3_rates_device.argtypes = [vp, C.c_int64] + [vp] * 7 + [vp]
        L.cmaq_ros3_solve_device.argtypes = [vp, C.c_int64, C.c_double] + [vp] * 4 + [C.POINTER(_Fail), vp]
        L.cmaq_ros3_solve_device_noalloc.argtypes = [vp, C.c_int64, C.c_double] + [vp] * 4 + [C.POINTER(_Fail), vp]
        L.cmaq_ros3_solve_yave_device.argtypes = [vp, C.c_int64, C.c_double] + [vp] * 5 + [C.POINTER(_Fail), vp]
        L.cmaq_ros3_ddm_device.argtypes = [vp, C.c_int64, C.c_double, vp, vp, vp, C.c_int32, _i32p, vp, vp, vp,
                                           C.POINTER(_Fail), vp]
        L.cmaq_ros3_adjoint_device.argtypes = [vp, C.c_int64, C.c_double] + [vp] * 5 + [C.POINTER(_Fail), vp]
        L.cmaq_ros3_last_adjoint_timing.argtypes = [vp, _f64p, _f64p, C.POINTER(C.c_int64)]
        L.cmaq_ros3_solve_tape_device.argtypes = [vp, C.c_int64, C.c_double] + [vp] * 5 + [C.c_int64] + \
            [vp] * 2 + [C.POINTER(_Fail), vp]
        L.cmaq_ros3_reverse_device.argtypes = [vp, C.c_int64, C.c_double] + [vp] * 3 + [C.c_int64] + \
            [vp] * 4 + [C.POINTER(_Fail), vp]
        L.cmaq_ros3_run_steps_device.argtypes = [vp, C.c_int64, C.c_double, C.c_int32] + [vp] * 10 + \
            [C.POINTER(_Fail), vp]
        L.cmaq_ros3_last_steps_timing.argtypes = [vp, _f64p]
        L.cmaq_ros3_set_queue_order.argtypes = [vp, vp, C.c_int64]
        L.cmaq_ros3_query.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int64)]
        L.cmaq_ros3_species_order.argtypes = [vp, _i32p]
        L.cmaq_ros3_get_fail_detail.argtypes = [vp, C.POINTER(_FailDetail)] + [_f64p] * 5
        L.cmaq_ros3_last_error.argtypes = [vp]
        L.cmaq_ros3_last_error.restype = C.c_char_p
        L.cmaq_ros3_free.argtypes = [vp]
        L.cmaq_ros3_last_timing.argtypes = [vp, _f64p, _f64p]
        L.cmaq_ros3_set_variant.argtypes = [vp, C.c_char_p]
        L.cmaq_ros3_variant.argtypes = [vp, C.c_int]
        L.cmaq_ros3_variant.restype = C.c_char_p
        L.cmaq_ros3_tlm_device.argtypes = [vp, C.c_int64, C.c_double] + [vp] * 4 + [C.c_int32] + [vp] * 2 + \
            [C.POINTER(_Fail), vp]
        L.cmaq_ros3_set_cgrid_map.argtypes = [vp, C.c_int32, _i32p, _i32p, _f64p]
        L.cmaq_ros3_chem.argtypes = [vp, C.POINTER(_Grid), C.c_double, _i32p, _i32p, C.POINTER(_Fail)]
        L.cmaq_ros3_chem_device.argtypes = [vp, C.POINTER(_Grid), C.c_double, vp, vp, C.POINTER(_Fail), vp]
        L.cmaq_ros3_gather_device.argtypes = [vp, C.POINTER(_Grid)] + [vp] * 7 + [vp]
        L.cmaq_ros3_scatter_device.argtypes = [vp, C.POINTER(_Grid), vp, vp]
        L.cmaq_gear_set_tolerances.argtypes = [vp, C.POINTER(_GearTol)]
        L.cmaq_gear_set_tail.argtypes = [vp, C.c_int32, C.c_int32]
        L.cmaq_gear_run.argtypes = [vp, C.c_int64, C.c_double] + [_f64p] * 7 + [_i32p, _i32p, C.POINTER(_Fail)]
        L.cmaq_gear_run_device.argtypes = [vp, C.c_int64, C.c_double] + [vp] * 9 + [C.POINTER(_Fail), vp]
        L.cmaq_gear_solve_device.argtypes = [vp, C.c_int64, C.c_double] + [vp] * 5 + [C.POINTER(_Fail), vp]
        L.cmaq_gear_last_timing.argtypes = [vp, _f64p, _f64p]
        _lib = L
    return _lib


def load_mechanism(name_or_path):
    """Mechanism tables (RXNS_DATA) as a dict; name is e.g. 'cb6r5_ae7_aq'."""
    path = name_or_path
    if not os.path.exists(path):
        path = os.path.join(HERE, "mechs", f"{name_or_path}.json")
    with open(path) as f:
        return json.load(f)


def pack_mechanism(m):
    """cmaq_ros3_mech with Fortran column-major arrays, as CCTM would pass them."""
    keep = []

    def i32(a, order="F"):
        arr = np.asarray(a, dtype=np.int32)
        arr = np.ascontiguousarray(arr.T) if (arr.ndim == 2 and order == "F") else np.ascontiguousarray(arr)
        if arr.size == 0:
            arr = np.zeros(1, np.int32)
        keep.append(arr)
        return arr.ctypes.data_as(_i32p)

    def f64(a, order="F"):
        arr = np.asarray(a, dtype=np.float64)
        arr = np.ascontiguousarray(arr.T) if (arr.ndim == 2 and order == "F") else np.ascontiguousarray(arr)
        keep.append(arr)
        return arr.ctypes.data_as(_f64p)

    s = _Mech()
    s.numb_mech_spc, s.nrxns, s.mxprd, s.kunits = m["numb_mech_spc"], m["nrxns"], m["mxprd"], m["kunits"]
    # JSON holds IRR/SC reaction-major ([r][slot]); Fortran memory is column-major
    s.irr, s.sc = i32(m["irr"]), f64(m["sc"])
    s.nreact, s.nprdct, s.ktype = i32(m["nreact"]), i32(m["nprdct"]), i32(m["ktype"])
    s.irxbits, s.iorder = i32(m["irxbits"]), i32(m["iorder"])
    # RTDAT(3, NRXNS): JSON rtdat[r] = [k1, k2, k3] is already the Fortran memory order
    s.rtdat = f64(m["rtdat"], order="C")
    s.nfalloff, s.irrfall, s.rfdat = m["nfalloff"], i32(m["irrfall"]), f64(m["rfdat"], order="C")
    s.nmphot, s.nphotab, s.iph = m["nmphot"], m["nphotab"], i32(m["iph"])
    s.mhetero, s.nhetero, s.ihetero = m["mhetero"], m["nhetero"], i32(m["ihetero"])
    for key, nkey in (("nrxwm", "nwm"), ("nrxww", "nww"), ("nrxwo2", "nwo2"), ("nrxwn2", "nwn2"),
                      ("nrxwch4", "nwch4"), ("nrxwh2", "nwh2")):
        setattr(s, key, i32(m[key]))
        setattr(s, nkey, len(m[key]))
    s.atm_air, s.atm_o2, s.atm_n2, s.atm_ch4, s.atm_h2 = (m["atm_air"], m["atm_o2"], m["atm_n2"],
                                                          m["atm_ch4"], m["atm_h2"])
    s.nspecial_rxn = 0
    return s, keep


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


class Ros3Solver:
    """One solver context (RBINIT + RBSPARSE) for one mechanism on one GPU."""

    def __init__(self, mech="cb6r5_ae7_aq", rtol=1.0e-3, atol=1.0e-7, device=-1,
                 max_attempts=0, waves_per_cu=0, refill_min=0):
        self.mech = load_mechanism(mech) if isinstance(mech, str) else mech
        self.ns = self.mech["numb_mech_spc"]
        self.nr = self.mech["nrxns"]
        self.nphotab = self.mech["nphotab"]
        self.nhetero = self.mech["nhetero"]
        s, keep = pack_mechanism(self.mech)
        rt = np.ascontiguousarray(np.broadcast_to(np.asarray(rtol, np.float64), (self.ns,)))
        at = np.ascontiguousarray(np.broadcast_to(np.asarray(atol, np.float64), (self.ns,)))
        self.rtol, self.atol = rt, at
        o = _Opts(device, max_attempts, waves_per_cu, refill_min)
        ctx = C.c_void_p()
        rc = lib().cmaq_ros3_init(C.byref(ctx), C.byref(s), rt.ctypes.data_as(_f64p),
                                  at.ctypes.data_as(_f64p), C.byref(o))
        del keep
        if rc != 0:
            raise Ros3Error(rc, lib().cmaq_ros3_last_error(None).decode())
        self._ctx = ctx

    def close(self):
        if getattr(self, "_ctx", None):
            lib().cmaq_ros3_free(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self, rc, fail=None):
        return Ros3Error(rc, lib().cmaq_ros3_last_error(self._ctx).decode(), fail)

    def query(self, key):
        v = C.c_int64()
        rc = lib().cmaq_ros3_query(self._ctx, key.encode(), C.byref(v))
        if rc != 0:
            raise self._err(rc)
        return v.value

    @property
    def variant(self):
        return lib().cmaq_ros3_variant(self._ctx, -1).decode()

    def variants(self):
        out, i = [], 0
        while True:
            v = lib().cmaq_ros3_variant(self._ctx, i)
            if v is None:
                return out
            out.append(v.decode())
            i += 1

    def set_variant(self, name):
        rc = lib().cmaq_ros3_set_variant(self._ctx, name.encode())
        if rc != 0:
            raise self._err(rc)

    def last_timing(self):
        """(rates_ms, solve_ms) of the last run, from HIP events on its stream."""
        a, b = C.c_double(), C.c_double()
        lib().cmaq_ros3_last_timing(self._ctx, C.byref(a), C.byref(b))
        return a.value, b.value

    def fail_detail(self):
        """The first failure of the last call as rbsolver.F:532-583 reports it:
        None without a failure, else a dict with the scalars of
        cmaq_ros3_fail_detail and, in mechanism order, 'spc_err' (MAX_SPC_ERR),
        'y' (last accepted state), 'yp' (the rejected step's state), plus the
        cell's 'rj' and 'rki' (NaN where the call had no such input)."""
        d = _FailDetail()
        a = {"spc_err": np.zeros(self.ns), "y": np.zeros(self.ns), "yp": np.zeros(self.ns),
             "rj": np.zeros(self.nphotab), "rki": np.zeros(self.nr)}
        p = lambda k: a[k].ctypes.data_as(_f64p)
        rc = lib().cmaq_ros3_get_fail_detail(self._ctx, C.byref(d), p("spc_err"), p("y"), p("yp"), p("rj"), p("rki"))
        if rc != 0:
            raise self._err(rc)
        if d.code == 0:
            return None
        out = {k: getattr(d, k) for k, _ in _FailDetail._fields_}
        out.update(a)
        return out

    def species_order(self):
        out = np.zeros(self.ns, np.int32)
        lib().cmaq_ros3_species_order(self._ctx, out.ctypes.data_as(_i32p))
        return out

    def run(self, chemstep, temp, pres_atm, h2o_ppm, rj, y, het=None, seawater=None, raise_on_fail=True):
        """Host arrays (numpy float64); returns (y_new, n_accepted, n_rejected).

        With raise_on_fail=False a convergence failure still returns the
        arrays (failed cells have n_accepted = -1)."""
        temp = np.ascontiguousarray(temp, np.float64)
        n = temp.size
        pres_atm = np.ascontiguousarray(pres_atm, np.float64)
        h2o_ppm = np.ascontiguousarray(h2o_ppm, np.float64)
        rj = np.ascontiguousarray(rj, np.float64).reshape(self.nphotab, n)
        yy = np.array(y, np.float64, copy=True, order="C").reshape(self.ns, n)
        het_p = None if het is None else np.ascontiguousarray(het, np.float64).reshape(self.nhetero, n)
        sw_p = None if seawater is None else np.ascontiguousarray(seawater, np.float64)
        na, nj = np.zeros(n, np.int32), np.zeros(n, np.int32)
        f = _Fail()
        p = lambda a: None if a is None else a.ctypes.data_as(_f64p)
        rc = lib().cmaq_ros3_run(self._ctx, n, chemstep, p(temp), p(pres_atm), p(h2o_ppm), p(rj),
                                 p(het_p), p(sw_p), p(yy), na.ctypes.data_as(_i32p), nj.ctypes.data_as(_i32p),
                                 C.byref(f))
        if rc != 0 and (raise_on_fail or rc not in (-4, -5)):
            raise self._err(rc, {"cell": f.cell, "maxerr": f.maxerr, "dt": f.dt, "tnow": f.tnow})
        return yy, na, nj

    def run_device(self, chemstep, temp, pres_atm, h2o_ppm, rj, y, nacc=None, nrej=None,
                   het=None, seawater=None, stream=None, raise_on_fail=True):
        """torch CUDA tensors (float64; nacc/nrej int32); y updated in place.

        With raise_on_fail=False a convergence failure (cells marked by
        nacc = -1) returns the failure record instead of raising."""
        import torch
        n = temp.numel()
        for t in (temp, pres_atm, h2o_ppm, rj, y):
            if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
                raise ValueError("inputs must be contiguous float64 CUDA tensors")
        if y.numel() != self.ns * n or rj.numel() != self.nphotab * n:
            raise ValueError("shape mismatch")
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        f = _Fail()
        rc = lib().cmaq_ros3_run_device(self._ctx, n, chemstep, _ptr(temp), _ptr(pres_atm), _ptr(h2o_ppm),
                                        _ptr(rj), _ptr(het), _ptr(seawater), _ptr(y), _ptr(nacc), _ptr(nrej),
                                        C.byref(f), C.c_void_p(stream))
        fail = {"code": rc, "cell": f.cell, "maxerr": f.maxerr, "dt": f.dt, "tnow": f.tnow}
        if rc != 0 and (raise_on_fail or rc not in (-4, -5)):
            raise self._err(rc, fail)
        return fail if rc != 0 else None

    def run_blocks_device(self, chemstep, temp, pres_atm, h2o_ppm, rj, y, blksize=50, nacc=None, nrej=None,
                          het=None, seawater=None, stream=None, raise_on_fail=True):
        """run_device with the reference's stock block semantics: blocks of
        consecutive cells (at most blksize, rbinit.F:286-303) sharing
        LSUNLIGHT and one step size (rbsolver.F:443-456)."""
        import torch
        n = temp.numel()
        for t in (temp, pres_atm, h2o_ppm, rj, y):
            if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
                raise ValueError("inputs must be contiguous float64 CUDA tensors")
        if y.numel() != self.ns * n or rj.numel() != self.nphotab * n:
            raise ValueError("shape mismatch")
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        f = _Fail()
        rc = lib().cmaq_ros3_run_blocks_device(self._ctx, n, int(blksize), chemstep, _ptr(temp), _ptr(pres_atm),
                                               _ptr(h2o_ppm), _ptr(rj), _ptr(het), _ptr(seawater), _ptr(y),
                                               _ptr(nacc), _ptr(nrej), C.byref(f), C.c_void_p(stream))
        fail = {"code": rc, "cell": f.cell, "maxerr": f.maxerr, "dt": f.dt, "tnow": f.tnow}
        if rc != 0 and (raise_on_fail or rc not in (-4, -5)):
            raise self._err(rc, fail)
        return fail if rc != 0 else None

    # ---- CHEM level (CGRID in, CGRID out) ----
    def set_cgrid_map(self, nspcsd, cgrid_index, convert_conc, species_molwt):
        """MAP_CHEMISTRY_SPECIES result: 1-based CGRID index, CONVERT_CONC and
        SPECIES_MOLWT per mechanism species."""
        idx = np.ascontiguousarray(cgrid_index, np.int32)
        conv = np.ascontiguousarray(convert_conc, np.int32)
        mw = np.ascontiguousarray(species_molwt, np.float64)
        if idx.size != self.ns or conv.size != self.ns or mw.size != self.ns:
            raise ValueError("one entry per mechanism species expected")
        rc = lib().cmaq_ros3_set_cgrid_map(self._ctx, int(nspcsd), idx.ctypes.data_as(_i32p),
                                           conv.ctypes.data_as(_i32p), mw.ctypes.data_as(_f64p))
        if rc != 0:
            raise self._err(rc)

    def _grid(self, g, device):
        """cmaq_ros3_grid over numpy arrays (host) or torch tensors (device)."""
        if device:
            import torch
        want = {"cgrid": np.float32, "ta": np.float32, "dens": np.float32, "pres": np.float32,
                "qv": np.float32, "seawater_zone": np.float32, "seaice": np.float32, "rj": np.float32,
                "khetero": np.float64}
        st = _Grid()
        st.ncols, st.nrows, st.nlays = int(g["ncols"]), int(g["nrows"]), int(g["nlays"])
        n = st.ncols * st.nrows * st.nlays
        for k, dt in want.items():
            a = g.get(k)
            if a is None:
                setattr(st, k, None)
                continue
            if device:
                tdt = torch.float32 if dt == np.float32 else torch.float64
                if not (a.is_cuda and a.is_contiguous() and a.dtype == tdt):
                    raise ValueError(f"{k} must be a contiguous {tdt} CUDA tensor")
                setattr(st, k, a.data_ptr())
            else:
                if not (isinstance(a, np.ndarray) and a.dtype == dt and a.flags["C_CONTIGUOUS"]):
                    raise ValueError(f"{k} must be a C-contiguous {np.dtype(dt).name} array")
                setattr(st, k, a.ctypes.data)
        kh = g.get("khetero")
        st.ld_het = (kh.numel() if device else kh.size) // n if kh is not None else self.nhetero
        return st

    def chem(self, grid, chemstep, raise_on_fail=True):
        """One CHEM call on host arrays (make_grid's dict of numpy arrays):
        grid["cgrid"] is updated in place; returns (n_accepted, n_rejected)."""
        st = self._grid(grid, device=False)
        n = st.ncols * st.nrows * st.nlays
        na, nj = np.zeros(n, np.int32), np.zeros(n, np.int32)
        f = _Fail()
        rc = lib().cmaq_ros3_chem(self._ctx, C.byref(st), chemstep, na.ctypes.data_as(_i32p),
                                  nj.ctypes.data_as(_i32p), C.byref(f))
        if rc != 0 and (raise_on_fail or rc not in (-4, -5)):
            raise self._err(rc, {"cell": f.cell, "maxerr": f.maxerr, "dt": f.dt, "tnow": f.tnow})
        return na, nj

    def chem_device(self, grid, chemstep, nacc=None, nrej=None, stream=None, raise_on_fail=True):
        """One CHEM call on device arrays (the same dict with torch CUDA tensors)."""
        import torch
        st = self._grid(grid, device=True)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        f = _Fail()
        rc = lib().cmaq_ros3_chem_device(self._ctx, C.byref(st), chemstep, _ptr(nacc), _ptr(nrej), C.byref(f),
                                         C.c_void_p(stream))
        fail = {"code": rc, "cell": f.cell, "maxerr": f.maxerr, "dt": f.dt, "tnow": f.tnow}
        if rc != 0 and (raise_on_fail or rc not in (-4, -5)):
            raise self._err(rc, fail)
        return fail if rc != 0 else None

    def gather_device(self, grid, out=None, stream=None):
        """The block arrays of every cell (dict of float64 CUDA tensors);
        `out` (a dict returned by an earlier call) is reused when given."""
        import torch
        st = self._grid(grid, device=True)
        n = st.ncols * st.nrows * st.nlays
        dev = grid["ta"].device
        if out is None:
            z = lambda r: torch.empty((r, n), dtype=torch.float64, device=dev)
            out = {"temp": z(1)[0], "pres_atm": z(1)[0], "h2o": z(1)[0], "sea": z(1)[0], "rj": z(self.nphotab),
                   "het": z(self.nhetero) if grid.get("khetero") is not None else None, "y": z(self.ns)}
        self._check([("temp", out["temp"], 1), ("pres_atm", out["pres_atm"], 1), ("h2o", out["h2o"], 1),
                     ("sea", out["sea"], 1), ("rj", out["rj"], self.nphotab), ("het", out["het"], self.nhetero),
                     ("y", out["y"], self.ns)], n)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        rc = lib().cmaq_ros3_gather_device(self._ctx, C.byref(st), _ptr(out["temp"]), _ptr(out["pres_atm"]),
                                           _ptr(out["h2o"]), _ptr(out["rj"]), _ptr(out["het"]), _ptr(out["sea"]),
                                           _ptr(out["y"]), C.c_void_p(stream))
        if rc != 0:
            raise self._err(rc)
        return out

    def scatter_device(self, grid, y, stream=None):
        import torch
        st = self._grid(grid, device=True)
        self._check([("y", y, self.ns)], st.ncols * st.nrows * st.nlays)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        rc = lib().cmaq_ros3_scatter_device(self._ctx, C.byref(st), _ptr(y), C.c_void_p(stream))
        if rc != 0:
            raise self._err(rc)

    def _check(self, tensors, n):
        import torch
        for name, t, rows in tensors:
            if t is None:
                continue
            if not (t.is_cuda and t.is_contiguous()):
                raise ValueError(f"{name} must be a contiguous CUDA tensor")
            want = torch.int32 if name in ("nacc", "nrej", "nacc_fwd", "order") else torch.float64
            if t.dtype != want:
                raise ValueError(f"{name} must be {want}")
            if t.numel() != rows * n:
                raise ValueError(f"{name}: expected {rows} x {n} elements, got {t.numel()}")

    def run_steps_device(self, chemstep, nsteps, temp, pres_atm, h2o_ppm, rj, y, y_ckpt=None, nacc=None,
                         nrej=None, het=None, seawater=None, stream=None, raise_on_fail=True):
        """nsteps sync steps through the captured graph; y [ns, n] in place,
        y_ckpt [nsteps, ns, n], nacc / nrej [nsteps, n] (optional)."""
        import torch
        n = temp.numel()
        self._check([("temp", temp, 1), ("pres", pres_atm, 1), ("h2o", h2o_ppm, 1), ("rj", rj, self.nphotab),
                     ("y", y, self.ns), ("y_ckpt", y_ckpt, nsteps * self.ns), ("nacc", nacc, nsteps),
                     ("nrej", nrej, nsteps), ("het", het, self.nhetero), ("sea", seawater, 1)], n)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        f = _Fail()
        rc = lib().cmaq_ros3_run_steps_device(self._ctx, n, chemstep, nsteps, _ptr(temp), _ptr(pres_atm),
                                              _ptr(h2o_ppm), _ptr(rj), _ptr(het), _ptr(seawater), _ptr(y),
                                              _ptr(y_ckpt), _ptr(nacc), _ptr(nrej), C.byref(f), C.c_void_p(stream))
        fail = {"code": rc, "cell": f.cell, "maxerr": f.maxerr, "dt": f.dt, "tnow": f.tnow}
        if rc != 0 and (raise_on_fail or rc not in (-4, -5)):
            raise self._err(rc, fail)
        return fail if rc != 0 else None

    def set_queue_order(self, order):
        """Hands cells to lanes in this order (int32 CUDA tensor, a permutation
        of the cells) for forward calls of the same size; None clears it.
        The tensor must stay alive while set."""
        if order is not None:
            self._check([("order", order, 1)], order.numel())
        self._order = order
        rc = lib().cmaq_ros3_set_queue_order(self._ctx, _ptr(order), 0 if order is None else order.numel())
        if rc != 0:
            raise self._err(rc)

    def last_steps_timing(self):
        v = C.c_double()
        lib().cmaq_ros3_last_steps_timing(self._ctx, C.byref(v))
        return v.value

    def solve_device(self, chemstep, rki, y, nacc=None, nrej=None, stream=None, raise_on_fail=True):
        """RBSOLVER with the rate constants given: rki [nr, n], y [ns, n]
        (updated in place); see cmaq_ros3_solve_device."""
        import torch
        n = y.numel() // self.ns
        self._check([("rki", rki, self.nr), ("y", y, self.ns), ("nacc", nacc, 1), ("nrej", nrej, 1)], n)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        f = _Fail()
        rc = lib().cmaq_ros3_solve_device(self._ctx, n, chemstep, _ptr(rki), _ptr(y), _ptr(nacc), _ptr(nrej),
                                          C.byref(f), C.c_void_p(stream))
        fail = {"code": rc, "cell": f.cell, "maxerr": f.maxerr, "dt": f.dt, "tnow": f.tnow}
        if rc != 0 and (raise_on_fail or rc not in (-4, -5)):
            raise self._err(rc, fail)
        return fail if rc != 0 else None

    # ---- SMVGEAR (CCTM/src/gas/smvgear) on the same context ----
    def set_gear_tolerances(self, rtol=1.0e-3, max_atol=1.0e-7, min_atol=1.0e-9, conc_floor=1.0e-12):
        """GEAR_RTOL, GEAR_MAX_ATOL, GEAR_MIN_ATOL, GEAR_CONC_FLOOR (REAL(4),
        RUNTIME_VARS.F:173-176)."""
        t = _GearTol(rtol, max_atol, min_atol, conc_floor)
        rc = lib().cmaq_gear_set_tolerances(self._ctx, C.byref(t))
        if rc != 0:
            raise self._err(rc)

    def set_gear_tail(self, after_attempts=2048, max_cells=0):
        """Cells still integrating after `after_attempts` attempted Gear steps
        (0 = never) finish in the one-cell-per-wavefront LDS kernel, at most
        `max_cells` per call (0 = automatic).  Results do not change."""
        rc = lib().cmaq_gear_set_tail(self._ctx, int(after_attempts), int(max_cells))
        if rc != 0:
            raise self._err(rc)

    def gear_run(self, chemstep, temp, pres_atm, h2o_ppm, rj, y, het=None, seawater=None, raise_on_fail=True):
        """CALC_RCONST + SMVGEAR on host arrays; returns (y_new, n_steps,
        n_fail); a failed cell has n_steps = -1."""
        temp = np.ascontiguousarray(temp, np.float64)
        n = temp.size
        pres_atm = np.ascontiguousarray(pres_atm, np.float64)
        h2o_ppm = np.ascontiguousarray(h2o_ppm, np.float64)
        rj = np.ascontiguousarray(rj, np.float64).reshape(self.nphotab, n)
        yy = np.array(y, np.float64, copy=True, order="C").reshape(self.ns, n)
        het_p = None if het is None else np.ascontiguousarray(het, np.float64).reshape(self.nhetero, n)
        sw_p = None if seawater is None else np.ascontiguousarray(seawater, np.float64)
        ns_, nf = np.zeros(n, np.int32), np.zeros(n, np.int32)
        f = _Fail()
        p = lambda a: None if a is None else a.ctypes.data_as(_f64p)
        rc = lib().cmaq_gear_run(self._ctx, n, chemstep, p(temp), p(pres_atm), p(h2o_ppm), p(rj), p(het_p),
                                 p(sw_p), p(yy), ns_.ctypes.data_as(_i32p), nf.ctypes.data_as(_i32p), C.byref(f))
        if rc != 0 and (raise_on_fail or rc not in (-4, -5)):
            raise self._err(rc, {"cell": f.cell, "ylowa": f.maxerr, "dt": f.dt, "tnow": f.tnow})
        return yy, ns_, nf

    def gear_run_device(self, chemstep, temp, pres_atm, h2o_ppm, rj, y, nsteps=None, nfail=None, het=None,
                        seawater=None, stream=None, raise_on_fail=True):
        """CALC_RCONST + SMVGEAR on CUDA tensors; y updated in place."""
        import torch
        n = temp.numel()
        self._check([("temp", temp, 1), ("pres", pres_atm, 1), ("h2o", h2o_ppm, 1), ("rj", rj, self.nphotab),
                     ("y", y, self.ns), ("nacc", nsteps, 1), ("nrej", nfail, 1), ("het", het, self.nhetero),
                     ("seawater", seawater, 1)], n)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        f = _Fail()
        rc = lib().cmaq_gear_run_device(self._ctx, n, chemstep, _ptr(temp), _ptr(pres_atm), _ptr(h2o_ppm),
                                        _ptr(rj), _ptr(het), _ptr(seawater), _ptr(y), _ptr(nsteps), _ptr(nfail),
                                        C.byref(f), C.c_void_p(stream))
        fail = {"code": rc, "cell": f.cell, "ylowa": f.maxerr, "dt": f.dt, "tnow": f.tnow}
        if rc != 0 and (raise_on_fail or rc not in (-4, -5)):
            raise self._err(rc, fail)
        return fail if rc != 0 else None

    def gear_solve_device(self, chemstep, rki, y, sunlit=None, nsteps=None, nfail=None, stream=None,
                          raise_on_fail=True):
        """SMVGEAR with the rate constants given: rki [nr, n], sunlit [n]
        uint8 (None = all sunlit), y [ns, n] updated in place."""
        import torch
        n = y.numel() // self.ns
        self._check([("rki", rki, self.nr), ("y", y, self.ns), ("nacc", nsteps, 1), ("nrej", nfail, 1)], n)
        if sunlit is not None and not (sunlit.is_cuda and sunlit.dtype == torch.uint8 and sunlit.numel() == n
                                       and sunlit.is_contiguous()):
            raise ValueError("sunlit must be a contiguous uint8 CUDA tensor of n elements")
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        f = _Fail()
        rc = lib().cmaq_gear_solve_device(self._ctx, n, chemstep, _ptr(rki), _ptr(sunlit), _ptr(y), _ptr(nsteps),
                                          _ptr(nfail), C.byref(f), C.c_void_p(stream))
        fail = {"code": rc, "cell": f.cell, "ylowa": f.maxerr, "dt": f.dt, "tnow": f.tnow}
        if rc != 0 and (raise_on_fail or rc not in (-4, -5)):
            raise self._err(rc, fail)
        return fail if rc != 0 else None

    def gear_last_timing(self):
        r, s = C.c_double(), C.c_double()
        lib().cmaq_gear_last_timing(self._ctx, C.byref(r), C.byref(s))
        return r.value, s.value

    def solve_device_noalloc(self, chemstep, rki, y, nacc=None, nrej=None, stream=None, raise_on_fail=True):
        """solve_device through the allocation-free entry point (its
        workspace in the library's static device array; fewer waves, same
        results); see cmaq_ros3_solve_device_noalloc."""
        import torch
        n = y.numel() // self.ns
        self._check([("rki", rki, self.nr), ("y", y, self.ns), ("nacc", nacc, 1), ("nrej", nrej, 1)], n)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        f = _Fail()
        rc = lib().cmaq_ros3_solve_device_noalloc(self._ctx, n, chemstep, _ptr(rki), _ptr(y), _ptr(nacc),
                                                  _ptr(nrej), C.byref(f), C.c_void_p(stream))
        fail = {"code": rc, "cell": f.cell, "maxerr": f.maxerr, "dt": f.dt, "tnow": f.tnow}
        if rc != 0 and (raise_on_fail or rc not in (-4, -5)):
            raise self._err(rc, fail)
        return fail if rc != 0 else None

    def solve_yave_device(self, chemstep, rki, y, yave, nacc=None, nrej=None, stream=None, raise_on_fail=True):
        """solve_device that also writes YAVE [ns, n], the time-averaged state
        of the -Dsens build; see cmaq_ros3_solve_yave_device."""
        import torch
        n = y.numel() // self.ns
        self._check([("rki", rki, self.nr), ("y", y, self.ns), ("yave", yave, self.ns), ("nacc", nacc, 1),
                     ("nrej", nrej, 1)], n)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        f = _Fail()
        rc = lib().cmaq_ros3_solve_yave_device(self._ctx, n, chemstep, _ptr(rki), _ptr(y), _ptr(yave), _ptr(nacc),
                                               _ptr(nrej), C.byref(f), C.c_void_p(stream))
        fail = {"code": rc, "cell": f.cell, "maxerr": f.maxerr, "dt": f.dt, "tnow": f.tnow}
        if rc != 0 and (raise_on_fail or rc not in (-4, -5)):
            raise self._err(rc, fail)
        return fail if rc != 0 else None

    DDM_MAX_PARAMS = 16

    def ddm_device(self, chemstep, rki, y, sens, prate, nacc, yave=None, nrej=None, stream=None,
                   raise_on_fail=True):
        """One sync step with CMAQ-DDM-3D's chemistry sensitivity step: y
        advanced in place; sens [np, ns, n] float32 (mechanism order) advanced
        by one Crank-Nicolson step; prate[p] = 0 (propagation only) or the
        1-based reaction whose ln k parameter p is.  See cmaq_ros3_ddm_device."""
        import torch
        n = y.numel() // self.ns
        npar = len(prate)
        self._check([("rki", rki, self.nr), ("y", y, self.ns), ("nacc", nacc, 1), ("nrej", nrej, 1),
                     ("yave", yave, self.ns)], n)
        if not (sens.is_cuda and sens.dtype == torch.float32 and sens.is_contiguous()
                and sens.numel() == npar * self.ns * n):
            raise ValueError("sens must be a contiguous float32 CUDA tensor of np x ns x n")
        pr = np.ascontiguousarray(prate, np.int32)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        f = _Fail()
        rc = lib().cmaq_ros3_ddm_device(self._ctx, n, chemstep, _ptr(rki), _ptr(y), _ptr(yave), npar,
                                        pr.ctypes.data_as(_i32p), _ptr(sens), _ptr(nacc), _ptr(nrej), C.byref(f),
                                        C.c_void_p(stream))
        fail = {"code": rc, "cell": f.cell, "maxerr": f.maxerr, "dt": f.dt, "tnow": f.tnow}
        if rc != 0 and (raise_on_fail or rc not in (-4, -5)):
            raise self._err(rc, fail)
        return fail if rc != 0 else None

    TLM_MAX_DIRECTIONS = 8

    def tlm_device(self, chemstep, rki, y, dy, dk=None, nacc=None, nrej=None, stream=None, raise_on_fail=True):
        """solve_device plus tangent-linear directions: y [ns, n] advances in
        place, dy [ndir, ns, n] or [ns, n] (dy0 in, d y_end out) along dk
        [ndir, nr, n] or [nr, n] (optional); ndir <= 8."""
        import torch
        n = y.numel() // self.ns
        ndir = dy.numel() // (self.ns * n) if n else 1
        self._check([("rki", rki, self.nr), ("y", y, self.ns), ("dy", dy, ndir * self.ns),
                     ("dk", dk, ndir * self.nr), ("nacc", nacc, 1), ("nrej", nrej, 1)], n)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        f = _Fail()
        rc = lib().cmaq_ros3_tlm_device(self._ctx, n, chemstep, _ptr(rki), _ptr(y), _ptr(nacc), _ptr(nrej),
                                        ndir, _ptr(dy), _ptr(dk), C.byref(f), C.c_void_p(stream))
        fail = {"code": rc, "cell": f.cell, "maxerr": f.maxerr, "dt": f.dt, "tnow": f.tnow}
        if rc != 0 and (raise_on_fail or rc not in (-4, -5)):
            raise self._err(rc, fail)
        return fail if rc != 0 else None

    def adjoint_device(self, chemstep, rki, y0, nacc_fwd, lam, grad_rki=None, stream=None):
        """Discrete adjoint of solve_device(chemstep, rki, y0): lam = dL/dy_end
        is overwritten with dL/dy0; grad_rki (if given) accumulates dL/drki."""
        import torch
        n = y0.numel() // self.ns
        self._check([("rki", rki, self.nr), ("y0", y0, self.ns), ("nacc_fwd", nacc_fwd, 1),
                     ("lam", lam, self.ns), ("grad_rki", grad_rki, self.nr)], n)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        f = _Fail()
        rc = lib().cmaq_ros3_adjoint_device(self._ctx, n, chemstep, _ptr(rki), _ptr(y0), _ptr(nacc_fwd),
                                            _ptr(lam), _ptr(grad_rki), C.byref(f), C.c_void_p(stream))
        if rc != 0:
            raise self._err(rc, {"code": rc, "cell": f.cell})

    def new_tape(self, records, device=None):
        """An arena for the stored-tape adjoint: (tape, count), with room for
        `records` accepted steps and the record counter at zero."""
        import torch
        rs = self.query("tape_record_doubles")
        tape = torch.empty(max(1, int(records)) * rs, dtype=torch.float64, device=device)
        count = torch.zeros(1, dtype=torch.int64, device=device)
        return tape, count

    def solve_tape_device(self, chemstep, rki, y, nacc, tape, tape_count, tape_head, nrej=None, stream=None,
                          raise_on_fail=True):
        """solve_device that also appends every accepted step to the arena
        (tape, tape_count) from new_tape; tape_head [n] int64 receives each
        cell's last record.  Returns False, with y and nacc complete, if the
        arena overflowed (its tape is then unusable), else True."""
        import torch
        n = y.numel() // self.ns
        self._check([("rki", rki, self.nr), ("y", y, self.ns), ("nacc", nacc, 1), ("nrej", nrej, 1)], n)
        self._check_tape(tape, tape_count, tape_head, n)
        cap = tape.numel() // self.query("tape_record_doubles")
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        f = _Fail()
        rc = lib().cmaq_ros3_solve_tape_device(self._ctx, n, chemstep, _ptr(rki), _ptr(y), _ptr(nacc), _ptr(nrej),
                                               _ptr(tape), cap, _ptr(tape_count), _ptr(tape_head), C.byref(f),
                                               C.c_void_p(stream))
        if rc == -6:          # the call allocates nothing: -6 is the arena overflow
            if f.code != 0 and raise_on_fail:    # ... and a cell also failed
                raise self._err(f.code, {"code": f.code, "cell": f.cell, "tape_overflow": True})
            return False
        if rc != 0 and (raise_on_fail or rc not in (-4, -5)):
            raise self._err(rc, {"code": rc, "cell": f.cell})
        return True

    def reverse_device(self, chemstep, rki, nacc_fwd, tape, tape_count, tape_head, lam, grad_rki=None,
                       stream=None):
        """The adjoint of one solve_tape_device sync step from its stored
        records (no replay): lam and grad_rki as in adjoint_device."""
        import torch
        n = lam.numel() // self.ns
        self._check([("rki", rki, self.nr), ("nacc_fwd", nacc_fwd, 1), ("lam", lam, self.ns),
                     ("grad_rki", grad_rki, self.nr)], n)
        self._check_tape(tape, tape_count, tape_head, n)
        cap = tape.numel() // self.query("tape_record_doubles")
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        f = _Fail()
        rc = lib().cmaq_ros3_reverse_device(self._ctx, n, chemstep, _ptr(rki), _ptr(nacc_fwd), _ptr(tape), cap,
                                            _ptr(tape_count), _ptr(tape_head), _ptr(lam), _ptr(grad_rki),
                                            C.byref(f), C.c_void_p(stream))
        if rc != 0:
            raise self._err(rc, {"code": rc, "cell": f.cell})

    @staticmethod
    def _check_tape(tape, tape_count, tape_head, n):
        import torch
        for name, t, dt, size in (("tape", tape, torch.float64, None), ("tape_count", tape_count, torch.int64, 1),
                                  ("tape_head", tape_head, torch.int64, n)):
            if not (t.is_cuda and t.is_contiguous() and t.dtype == dt) or (size is not None and t.numel() != size):
                raise ValueError(f"{name} must be a contiguous CUDA {dt} tensor" +
                                 (f" of {size} elements" if size is not None else ""))

    def last_adjoint_timing(self):
        """(replay_ms, reverse_ms, chunks) of the last adjoint_device call."""
        a, b, k = C.c_double(), C.c_double(), C.c_int64()
        lib().cmaq_ros3_last_adjoint_timing(self._ctx, C.byref(a), C.byref(b), C.byref(k))
        return a.value, b.value, k.value

    def rates_device(self, temp, pres_atm, h2o_ppm, rj, rki, het=None, seawater=None, stream=None):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        rc = lib().cmaq_ros3_rates_device(self._ctx, temp.numel(), _ptr(temp), _ptr(pres_atm), _ptr(h2o_ppm),
                                          _ptr(rj), _ptr(het), _ptr(seawater), _ptr(rki), C.c_void_p(stream))
        if rc != 0:
            raise self._err(rc)
