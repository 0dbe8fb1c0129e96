"""ctypes wrapper of the CPU oracle (liboracle_ros3.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's CPU-baseline leg, always as the checker / reported baseline,
never as the product path.  The restated reference functions live in
ros3_oracle.c (each cites the Fortran it follows).
"""

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# CMAQ_ORACLE_LIB names another build (the ASan/UBSan one of `make -C oracle san`)
_LIB_PATH = os.environ.get("CMAQ_ORACLE_LIB") or os.path.join(_HERE, "liboracle_ros3.so")

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class _Tables(C.Structure):
    _fields_ = [
        ("ns", C.c_int32), ("nr", C.c_int32), ("mxprd", C.c_int32), ("kunits", C.c_int32),
        ("irr", _i32p), ("sc", _f64p),
        ("nreact", _i32p), ("nprdct", _i32p), ("ktype", _i32p), ("irxbits", _i32p), ("iorder", _i32p),
        ("rtdat", _f64p), ("nfalloff", C.c_int32), ("irrfall", _i32p), ("rfdat", _f64p),
        ("nmphot", C.c_int32), ("nphotab", C.c_int32), ("iph", _i32p),
        ("mhetero", C.c_int32), ("nhetero", C.c_int32), ("ihetero", _i32p),
        ("nwm", C.c_int32), ("nww", C.c_int32), ("nwo2", C.c_int32), ("nwn2", C.c_int32),
        ("nwch4", C.c_int32), ("nwh2", C.c_int32),
        ("nrxwm", _i32p), ("nrxww", _i32p), ("nrxwo2", _i32p), ("nrxwn2", _i32p),
        ("nrxwch4", _i32p), ("nrxwh2", _i32p),
        ("atm_air", C.c_double), ("atm_o2", C.c_double), ("atm_n2", C.c_double),
        ("atm_ch4", C.c_double), ("atm_h2", C.c_double),
    ]


class _GearTol(C.Structure):
    _fields_ = [("rtol", C.c_float), ("max_atol", C.c_float), ("min_atol", C.c_float), ("conc_floor", C.c_float)]


class _GearStats(C.Structure):
    _fields_ = [(k, C.c_long) for k in ("nsteps", "nsubfun", "npderiv", "numnewt", "ncfail", "nefail",
                                         "numbkups", "nylowred")] + [("mxordused", C.c_int)]


# RUNTIME_VARS.F:173-176 defaults (REAL(4)): GEAR_RTOL, GEAR_MAX_ATOL,
# GEAR_MIN_ATOL, GEAR_CONC_FLOOR
GEAR_TOL_DEFAULT = (1.0e-3, 1.0e-7, 1.0e-9, 1.0e-12)


class _Fail(C.Structure):
    _fields_ = [("cell", C.c_int64), ("maxerr", C.c_double), ("dt", C.c_double), ("tnow", C.c_double),
                ("dump", C.c_void_p)]


def build():
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle_ros3.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        L.om_create.restype = vp
        L.om_create.argtypes = [C.POINTER(_Tables), C.c_char_p, C.c_int]
        L.om_destroy.argtypes = [vp]
        for name in ("om_ischan",):
            getattr(L, name).argtypes = [vp]
        for name in ("om_iarray", "om_nuserat", "om_nlu_ops"):
            getattr(L, name).argtypes = [vp, C.c_int]
        L.om_inew2old.argtypes = [vp, _i32p]
        L.om_nkuserat.argtypes = [vp, C.c_int, _i32p]
        L.om_pattern.argtypes = [vp, C.c_int, _i32p, _i32p]
        L.om_lu_ops.argtypes = [vp, C.c_int, _i32p, _i32p, _i32p]
        L.om_rates_block.argtypes = [vp, C.c_int] + [_f64p] * 7
        L.om_feval.argtypes = [vp, C.c_int, C.c_int, _f64p, _f64p, _f64p]
        L.om_jacob.argtypes = [vp, C.c_int, C.c_int, _f64p, _f64p, _f64p]
        L.om_decomp.argtypes = [vp, C.c_int, C.c_int, _f64p, _f64p]
        L.om_solve.argtypes = [vp, C.c_int, C.c_int, _f64p, _f64p, _f64p]
        L.om_rbsolver.argtypes = [vp, C.c_int, C.c_int, C.c_double, _f64p, _f64p, _f64p, _f64p,
                                  _i32p, _i32p, C.POINTER(_Fail)]
        L.om_run.argtypes = [vp, C.c_int64, C.c_int, C.c_double] + [_f64p] * 7 + [_f64p, _f64p,
                                                                               _i32p, _i32p, C.POINTER(_Fail)]
        L.om_stiffness.argtypes = [vp, C.c_int64, C.c_int] + [_f64p] * 10
        L.om_steps_cell.argtypes = [vp, C.c_int, C.c_double, _f64p, _f64p, _f64p, _f64p, _f64p, C.c_int,
                                    C.POINTER(C.c_int)]
        L.om_replay_cell.argtypes = [vp, C.c_int, _f64p, _f64p, _f64p, C.c_int]
        L.om_adjoint_cell.argtypes = [vp, C.c_int, _f64p, _f64p, _f64p, C.c_int, _f64p, _f64p]
        L.om_adjoint_run.argtypes = [vp, C.c_int64, C.c_double] + [_f64p] * 7 + [_i32p]
        L.om_ddm_cell.argtypes = [vp, C.c_double, _f64p, _f64p, _f64p, _f64p, C.c_int, _i32p,
                                  C.POINTER(C.c_float)]
        L.om_ddm_step.argtypes = [vp, C.c_double, _f64p, _f64p, C.c_int, _i32p, C.POINTER(C.c_float)]
        L.om_ddm_step.restype = None
        L.om_tlm_cell.argtypes = [vp, C.c_int, _f64p, _f64p, _f64p, C.c_int, _f64p, _f64p]
        L.om_gear_set_pow_nudge.argtypes = [C.c_int]
        L.om_gear_set_pow_nudge.restype = None
        L.om_gear_set_pow_exact.argtypes = [C.c_int]
        L.om_gear_set_pow_exact.restype = None
        L.om_gear_coefficients.argtypes = [_f64p] * 4
        L.om_gear_coefficients.restype = None
        L.om_subfun.argtypes = [vp, C.c_int, C.c_int, _f64p, _f64p, _f64p]
        L.om_subfun.restype = None
        L.om_pderiv.argtypes = [vp, C.c_int, C.c_int, _f64p, _f64p, C.c_double, _f64p]
        L.om_pderiv.restype = None
        L.om_optimal_atol.argtypes = [vp, C.c_int, _f64p, C.POINTER(_GearTol), _f64p]
        L.om_optimal_atol.restype = None
        L.om_smvgear.argtypes = [vp, C.c_int, C.c_int, C.c_double, _f64p, _f64p, C.POINTER(_GearTol), C.c_int,
                                 C.POINTER(_GearStats)]
        L.om_gear_run.argtypes = [vp, C.c_int64, C.c_int, C.c_double] + [_f64p] * 7 + \
            [C.POINTER(_GearTol), _i32p, _i32p, _f64p]
        _f32p = C.POINTER(C.c_float)
        L.om_gather.argtypes = [C.c_int] * 6 + [_i32p, _i32p, _f64p] + [_f32p] * 8 + [_f64p, C.c_int] + \
            [_f64p] * 7
        L.om_scatter.argtypes = [C.c_int] * 4 + [_i32p, _i32p, _f64p, _f32p, _f64p, _f32p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


class Oracle:
    """RBSPARSE state for one mechanism plus the numeric kernels."""

    def __init__(self, mech):
        m = mech
        self.mech = m
        self.ns, self.nr = m["numb_mech_spc"], m["nrxns"]
        self.nphotab, self.nhetero = m["nphotab"], m["nhetero"]
        keep = {}

        def i32(x):
            a = np.ascontiguousarray(np.asarray(x, dtype=np.int32).reshape(-1)) if len(x) else np.zeros(1, np.int32)
            keep[id(a)] = a
            return _p(a, _i32p)

        def f64(x):
            a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
            keep[id(a)] = a
            return _p(a, _f64p)

        t = _Tables()
        t.ns, t.nr, t.mxprd, t.kunits = self.ns, self.nr, m["mxprd"], m["kunits"]
        t.irr, t.sc = i32(m["irr"]), f64(m["sc"])
        t.nreact, t.nprdct, t.ktype = i32(m["nreact"]), i32(m["nprdct"]), i32(m["ktype"])
        t.irxbits, t.iorder, t.rtdat = i32(m["irxbits"]), i32(m["iorder"]), f64(m["rtdat"])
        t.nfalloff, t.irrfall, t.rfdat = m["nfalloff"], i32(m["irrfall"]), f64(m["rfdat"])
        t.nmphot, t.nphotab, t.iph = m["nmphot"], m["nphotab"], i32(m["iph"])
        t.mhetero, t.nhetero, t.ihetero = m["mhetero"], m["nhetero"], i32(m["ihetero"])
        for key in ("nrxwm", "nrxww", "nrxwo2", "nrxwn2", "nrxwch4", "nrxwh2"):
            setattr(t, key, i32(m[key]))
        t.nwm, t.nww, t.nwo2 = len(m["nrxwm"]), len(m["nrxww"]), len(m["nrxwo2"])
        t.nwn2, t.nwch4, t.nwh2 = len(m["nrxwn2"]), len(m["nrxwch4"]), len(m["nrxwh2"])
        t.atm_air, t.atm_o2, t.atm_n2 = m["atm_air"], m["atm_o2"], m["atm_n2"]
        t.atm_ch4, t.atm_h2 = m["atm_ch4"], m["atm_h2"]
        self._keep, self._t = keep, t
        err = C.create_string_buffer(256)
        self._ctx = lib().om_create(C.byref(t), err, 256)
        if not self._ctx:
            raise ValueError(err.value.decode())

    def __del__(self):
        if getattr(self, "_ctx", None) and _lib is not None:
            _lib.om_destroy(self._ctx)
            self._ctx = None

    # ---- symbolic ----
    @property
    def ischan(self):
        return lib().om_ischan(self._ctx)

    def iarray(self, ncsp):
        return lib().om_iarray(self._ctx, ncsp)

    def inew2old(self):
        out = np.zeros(self.ns, np.int32)
        lib().om_inew2old(self._ctx, _p(out, _i32p))
        return out

    def nkuserat(self, ncsp):
        out = np.zeros(lib().om_nuserat(self._ctx, ncsp), np.int32)
        lib().om_nkuserat(self._ctx, ncsp, _p(out, _i32p))
        return out

    def pattern(self, ncsp):
        n = self.iarray(ncsp)
        rows, cols = np.zeros(n, np.int32), np.zeros(n, np.int32)
        lib().om_pattern(self._ctx, ncsp, _p(rows, _i32p), _p(cols, _i32p))
        return rows, cols

    def lu_ops(self, ncsp):
        n = lib().om_nlu_ops(self._ctx, ncsp)
        a, b, c = (np.zeros(n, np.int32) for _ in range(3))
        lib().om_lu_ops(self._ctx, ncsp, _p(a, _i32p), _p(b, _i32p), _p(c, _i32p))
        return a, b, c

    # ---- numeric (block = all given cells) ----
    def rates(self, temp, pres_atm, h2o_ppm, rj, het=None, seawater=None):
        temp = np.ascontiguousarray(temp, np.float64)
        nc = temp.size
        pres_atm = np.ascontiguousarray(pres_atm, np.float64)
        h2o_ppm = np.ascontiguousarray(h2o_ppm, np.float64)
        rj = np.ascontiguousarray(rj, np.float64).reshape(self.nphotab, nc)
        het = np.zeros((max(self.nhetero, 1), nc)) if het is None else np.ascontiguousarray(het, np.float64)
        seawater = np.zeros(nc) if seawater is None else np.ascontiguousarray(seawater, np.float64)
        rki = np.zeros((self.nr, nc))
        sun = lib().om_rates_block(self._ctx, nc, _p(temp, _f64p), _p(pres_atm, _f64p), _p(h2o_ppm, _f64p),
                                   _p(rj, _f64p), _p(het, _f64p), _p(seawater, _f64p), _p(rki, _f64p))
        return rki, bool(sun)

    def feval(self, ncsp, rki, y_new):
        rki = np.ascontiguousarray(rki, np.float64)
        y_new = np.ascontiguousarray(y_new, np.float64)
        nc = y_new.shape[1]
        ydot = np.zeros_like(y_new)
        lib().om_feval(self._ctx, ncsp, nc, _p(rki, _f64p), _p(y_new, _f64p), _p(ydot, _f64p))
        return ydot

    def jacob(self, ncsp, rki, y_new):
        rki = np.ascontiguousarray(rki, np.float64)
        y_new = np.ascontiguousarray(y_new, np.float64)
        nc = y_new.shape[1]
        cc2 = np.zeros((self.iarray(ncsp) + 1, nc))
        lib().om_jacob(self._ctx, ncsp, nc, _p(rki, _f64p), _p(y_new, _f64p), _p(cc2, _f64p))
        return cc2

    def decomp(self, ncsp, cc2):
        cc2 = np.array(cc2, np.float64, copy=True, order="C")
        nc = cc2.shape[1]
        vdiag = np.zeros((self.ischan, nc))
        lib().om_decomp(self._ctx, ncsp, nc, _p(cc2, _f64p), _p(vdiag, _f64p))
        return cc2, vdiag

    def solve(self, ncsp, cc2, vdiag, rhs):
        cc2 = np.ascontiguousarray(cc2, np.float64)
        vdiag = np.ascontiguousarray(vdiag, np.float64)
        rhs = np.array(rhs, np.float64, copy=True, order="C")
        lib().om_solve(self._ctx, ncsp, rhs.shape[1], _p(cc2, _f64p), _p(vdiag, _f64p), _p(rhs, _f64p))
        return rhs

    def rbsolver(self, ncsp, chemstep, rki, y_new, rtol_new, atol_new, dump=False):
        """RBSOLVER on one block (arrays in NEW order).  With dump=True a
        fifth value holds, after a DT < DTMIN failure, the failing cell's
        ERR, Y, YP, K1, K2, K3 (rbsolver.F:540-560), else None."""
        rki = np.ascontiguousarray(rki, np.float64)
        y = np.array(y_new, np.float64, copy=True, order="C")
        rt = np.ascontiguousarray(rtol_new, np.float64)
        at = np.ascontiguousarray(atol_new, np.float64)
        na, nj, f = C.c_int32(), C.c_int32(), _Fail()
        buf = np.full(1 + 5 * self.ns, np.nan)
        if dump:
            f.dump = buf.ctypes.data
        rc = lib().om_rbsolver(self._ctx, ncsp, y.shape[1], chemstep, _p(rki, _f64p), _p(y, _f64p),
                               _p(rt, _f64p), _p(at, _f64p), C.byref(na), C.byref(nj), C.byref(f))
        if not dump:
            return rc, y, na.value, nj.value
        return rc, y, na.value, nj.value, (buf if rc != 0 else None)

    def run(self, chemstep, temp, pres_atm, h2o_ppm, rj, y, rtol, atol, het=None, seawater=None, blksize=1):
        """Whole-domain driver; y is [ns, ncells] in mechanism order (copied)."""
        temp = np.ascontiguousarray(temp, np.float64)
        nc = temp.size
        args = [np.ascontiguousarray(a, np.float64) for a in (pres_atm, h2o_ppm, rj)]
        het = np.zeros((max(self.nhetero, 1), nc)) if het is None else np.ascontiguousarray(het, np.float64)
        sw = np.zeros(nc) if seawater is None else np.ascontiguousarray(seawater, np.float64)
        yy = np.array(y, np.float64, copy=True, order="C")
        rt = np.ascontiguousarray(np.broadcast_to(rtol, (self.ns,)), np.float64)
        at = np.ascontiguousarray(np.broadcast_to(atol, (self.ns,)), np.float64)
        na, nj = np.zeros(nc, np.int32), np.zeros(nc, np.int32)
        f = _Fail()
        rc = lib().om_run(self._ctx, nc, blksize, chemstep, _p(temp, _f64p), _p(args[0], _f64p),
                          _p(args[1], _f64p), _p(args[2], _f64p), _p(het, _f64p), _p(sw, _f64p),
                          _p(yy, _f64p), _p(rt, _f64p), _p(at, _f64p), _p(na, _i32p), _p(nj, _i32p),
                          C.byref(f))
        return {"rc": rc, "y": yy, "nacc": na, "nrej": nj,
                "fail": {"cell": f.cell, "maxerr": f.maxerr, "dt": f.dt, "tnow": f.tnow}}

    def stiffness(self, temp, pres_atm, h2o_ppm, rj, y, rtol, atol, het=None, seawater=None, blksize=50):
        """ERRMX2 of the LORDERING pre-pass for every cell (y [ns, ncells],
        mechanism order)."""
        temp = np.ascontiguousarray(temp, np.float64)
        nc = temp.size
        args = [np.ascontiguousarray(a, np.float64) for a in (pres_atm, h2o_ppm, rj, y)]
        het = np.zeros((max(self.nhetero, 1), nc)) if het is None else np.ascontiguousarray(het, np.float64)
        sw = np.zeros(nc) if seawater is None else np.ascontiguousarray(seawater, np.float64)
        rt = np.ascontiguousarray(np.broadcast_to(rtol, (self.ns,)), np.float64)
        at = np.ascontiguousarray(np.broadcast_to(atol, (self.ns,)), np.float64)
        out = np.zeros(nc)
        lib().om_stiffness(self._ctx, nc, blksize, _p(temp, _f64p), _p(args[0], _f64p), _p(args[1], _f64p),
                           _p(args[2], _f64p), _p(het, _f64p), _p(sw, _f64p), _p(args[3], _f64p),
                           _p(rt, _f64p), _p(at, _f64p), _p(out, _f64p))
        return out

    # ---- SMVGEAR (CCTM/src/gas/smvgear; gear_oracle.c) ----
    @staticmethod
    def _gear_tol(tol):
        t = _GearTol()
        t.rtol, t.max_atol, t.min_atol, t.conc_floor = tol or GEAR_TOL_DEFAULT
        return t

    def subfun(self, ncsp, rki, c_new):
        rki = np.ascontiguousarray(rki, np.float64)
        c_new = np.ascontiguousarray(c_new, np.float64)
        g = np.zeros_like(c_new)
        lib().om_subfun(self._ctx, ncsp, c_new.shape[1], _p(rki, _f64p), _p(c_new, _f64p), _p(g, _f64p))
        return g

    def pderiv(self, ncsp, rki, c_new, r1delt):
        rki = np.ascontiguousarray(rki, np.float64)
        c_new = np.ascontiguousarray(c_new, np.float64)
        cc2 = np.zeros((self.iarray(ncsp) + 1, c_new.shape[1]))
        lib().om_pderiv(self._ctx, ncsp, c_new.shape[1], _p(rki, _f64p), _p(c_new, _f64p), r1delt,
                        _p(cc2, _f64p))
        return cc2

    def optimal_atol(self, c_new, tol=None):
        c_new = np.ascontiguousarray(c_new, np.float64)
        out = np.zeros(c_new.shape[1])
        t = self._gear_tol(tol)
        lib().om_optimal_atol(self._ctx, c_new.shape[1], _p(c_new, _f64p), C.byref(t), _p(out, _f64p))
        return out

    def smvgear(self, ncsp, chemstep, rki, y_new, sunlit, tol=None):
        """SMVGEAR on one block (y in NEW order, copied).  Returns (rc, y,
        stats dict)."""
        rki = np.ascontiguousarray(rki, np.float64)
        y = np.array(y_new, np.float64, copy=True, order="C")
        t, st = self._gear_tol(tol), _GearStats()
        rc = lib().om_smvgear(self._ctx, ncsp, y.shape[1], chemstep, _p(rki, _f64p), _p(y, _f64p), C.byref(t),
                              int(bool(sunlit)), C.byref(st))
        return rc, y, {k: getattr(st, k) for k, _ in _GearStats._fields_}

    def gear_run(self, chemstep, temp, pres_atm, h2o_ppm, rj, y, het=None, seawater=None, blksize=1, tol=None,
                 rki=None):
        """The SMVGEAR block driver over every cell; y [ns, ncells] mechanism
        order (copied).  rki [nr, ncells] replaces CALC_RCONST if given."""
        temp = np.ascontiguousarray(temp, np.float64)
        nc = temp.size
        args = [np.ascontiguousarray(a, np.float64) for a in (pres_atm, h2o_ppm, rj)]
        het = np.zeros((max(self.nhetero, 1), nc)) if het is None else np.ascontiguousarray(het, np.float64)
        sw = np.zeros(nc) if seawater is None else np.ascontiguousarray(seawater, np.float64)
        yy = np.array(y, np.float64, copy=True, order="C")
        rk = None if rki is None else np.ascontiguousarray(rki, np.float64)
        ns_, nf = np.zeros(nc, np.int32), np.zeros(nc, np.int32)
        t = self._gear_tol(tol)
        rc = lib().om_gear_run(self._ctx, nc, blksize, chemstep, _p(temp, _f64p), _p(args[0], _f64p),
                               _p(args[1], _f64p), _p(args[2], _f64p), _p(het, _f64p), _p(sw, _f64p),
                               _p(yy, _f64p), C.byref(t), _p(ns_, _i32p), _p(nf, _i32p),
                               None if rk is None else _p(rk, _f64p))
        return {"rc": rc, "y": yy, "nsteps": ns_, "nfail": nf}

    # ---- discrete adjoint with the accepted steps held fixed (one cell) ----
    # y / lam arguments here are in MECHANISM order; the day pattern is used.
    def _to_new(self, v):
        return np.ascontiguousarray(np.asarray(v, np.float64)[self.inew2old() - 1])

    def _to_old(self, v_new):
        out = np.empty_like(v_new)
        out[self.inew2old() - 1] = v_new
        return out

    def steps_cell(self, chemstep, rki, y, rtol=1e-3, atol=1e-7, hcap=1 << 20):
        """Accepted step sizes of one cell's per-cell RBSOLVER run, and y_end."""
        rk = np.ascontiguousarray(rki, np.float64)
        yn = self._to_new(y)
        rt = self._to_new(np.broadcast_to(rtol, (self.ns,)))
        at = self._to_new(np.broadcast_to(atol, (self.ns,)))
        h = np.zeros(hcap)
        nh = C.c_int()
        rc = lib().om_steps_cell(self._ctx, 1, chemstep, _p(rk, _f64p), _p(yn, _f64p), _p(rt, _f64p),
                                 _p(at, _f64p), _p(h, _f64p), hcap, C.byref(nh))
        if rc != 0:
            raise RuntimeError(f"om_steps_cell failed ({rc})")
        return h[:nh.value].copy(), self._to_old(yn)

    def replay_cell(self, rki, y, h):
        rk = np.ascontiguousarray(rki, np.float64)
        yn = self._to_new(y)
        h = np.ascontiguousarray(h, np.float64)
        lib().om_replay_cell(self._ctx, 1, _p(rk, _f64p), _p(yn, _f64p), _p(h, _f64p), h.size)
        return self._to_old(yn)

    def adjoint_cell(self, rki, y0, h, lam):
        """(dL/dy0, dL/drki) for dL/dy_end = lam."""
        rk = np.ascontiguousarray(rki, np.float64)
        y0n = self._to_new(y0)
        ln = self._to_new(lam)
        h = np.ascontiguousarray(h, np.float64)
        gk = np.zeros(self.nr)
        lib().om_adjoint_cell(self._ctx, 1, _p(rk, _f64p), _p(y0n, _f64p), _p(h, _f64p), h.size,
                              _p(ln, _f64p), _p(gk, _f64p))
        return self._to_old(ln), gk

    def tlm_cell(self, rki, y0, h, dy0, dk=None):
        """d y_end along (dy0, dk) for the fixed steps h (mechanism order)."""
        rk = np.ascontiguousarray(rki, np.float64)
        yn = self._to_new(y0)
        dy = self._to_new(dy0)
        hh = np.ascontiguousarray(h, np.float64)
        dkp = None if dk is None else np.ascontiguousarray(dk, np.float64)
        lib().om_tlm_cell(self._ctx, 1, _p(rk, _f64p), _p(yn, _f64p), _p(hh, _f64p), len(hh), _p(dy, _f64p),
                          None if dkp is None else _p(dkp, _f64p))
        return self._to_old(dy)

    def adjoint_run(self, chemstep, rki, y0, lam, rtol=1e-3, atol=1e-7):
        """All cells: returns dict(lam=dL/dy0, gk=dL/drki, y=y_end, nacc)."""
        rki = np.ascontiguousarray(rki, np.float64)
        nc = rki.shape[1]
        y0 = np.ascontiguousarray(y0, np.float64)
        lam = np.array(lam, np.float64, copy=True, order="C")
        rt = np.ascontiguousarray(np.broadcast_to(rtol, (self.ns,)), np.float64)
        at = np.ascontiguousarray(np.broadcast_to(atol, (self.ns,)), np.float64)
        gk = np.zeros((self.nr, nc))
        yend = np.zeros_like(y0)
        na = np.zeros(nc, np.int32)
        rc = lib().om_adjoint_run(self._ctx, nc, chemstep, _p(rki, _f64p), _p(y0, _f64p), _p(rt, _f64p),
                                  _p(at, _f64p), _p(lam, _f64p), _p(gk, _f64p), _p(yend, _f64p),
                                  _p(na, _i32p))
        return {"rc": rc, "lam": lam, "gk": gk, "y": yend, "nacc": na}

    # ---- CMAQ-DDM-3D first-order chemistry sensitivities (one cell) ----
    def ddm_cell(self, chemstep, rki, y0, nsteps=1, rtol=1e-3, atol=1e-7, w=None):
        """Sensitivities of y_end to ln C0 of every species and to ln k of every
        reaction over nsteps sync steps, as DDM-3D computes them.  Returns
        (S_init [ns(param), ns(species)], S_rate [nr, ns], y_end), mechanism
        order; with a seed w, returns the contractions w.S instead of S."""
        ns, nr = self.ns, self.nr
        inew = self.inew2old() - 1
        rk = np.ascontiguousarray(rki, np.float64)
        y = self._to_new(y0)
        rt = self._to_new(np.broadcast_to(rtol, (ns,)))
        at = self._to_new(np.broadcast_to(atol, (ns,)))
        npar = ns + nr
        prate = np.zeros(npar, np.int32)
        prate[ns:] = np.arange(1, nr + 1)
        sens = np.zeros((npar, ns), np.float32)
        # INIT parameter j (mechanism species j): S0 = C0 at species j (new order)
        y0m = np.asarray(y0, np.float64)
        for j in range(ns):
            sens[j, np.flatnonzero(inew == j)[0]] = np.float32(y0m[j])
        for _ in range(nsteps):
            rc = lib().om_ddm_cell(self._ctx, chemstep, _p(rk, _f64p), _p(y, _f64p), _p(rt, _f64p), _p(at, _f64p),
                                   npar, _p(prate, _i32p), sens.ctypes.data_as(C.POINTER(C.c_float)))
            if rc != 0:
                raise RuntimeError(f"om_ddm_cell failed ({rc})")
        s_old = np.empty_like(sens)
        s_old[:, inew] = sens
        yend = self._to_old(y)
        if w is None:
            return s_old[:ns].astype(np.float64), s_old[ns:].astype(np.float64), yend
        wv = np.asarray(w, np.float64)
        return s_old[:ns].astype(np.float64) @ wv, s_old[ns:].astype(np.float64) @ wv, yend

    def ddm_step(self, chemstep, rki, yave, sens, prate):
        """The Crank-Nicolson sensitivity step alone (ddm3d/DDM3D_CHEM.F:296-395)
        from a given time-averaged state yave (mechanism order): sens
        [np, ns] float32 in mechanism order, prate[p] = 0 (initial-condition
        parameter) or the 1-based reaction of a rate parameter.  Returns the
        new sensitivities, mechanism order."""
        inew = self.inew2old() - 1
        ya = self._to_new(yave)
        rk = np.ascontiguousarray(rki, np.float64)
        s_new = np.ascontiguousarray(np.asarray(sens, np.float32)[:, inew])
        pr = np.ascontiguousarray(prate, np.int32)
        lib().om_ddm_step(self._ctx, chemstep, _p(rk, _f64p), _p(ya, _f64p), len(pr), _p(pr, _i32p),
                          s_new.ctypes.data_as(C.POINTER(C.c_float)))
        out = np.empty_like(s_new)
        out[:, inew] = s_new
        return out


# ---- CHEM's CGRID gather / scatter (rbdriver.F:460-509, 577-595) ----
def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def chem_gather(mech, grid, cgrid_index, convert_conc, molwt):
    """Block arrays of every cell from CGRID and the met fields.  grid: dict
    with ncols, nrows, nlays, cgrid [nspcsd][n] and ta, dens, pres, qv [n],
    rj [nphotab][n] (float32, cell = col + ncols*(row + nrows*lay)); optional
    seawater_zone, seaice [ncols*nrows] (float32), khetero [n][ld_het]
    (float64).  Returns a dict of float64 arrays; y in mechanism order."""
    nc, nr_, nl = grid["ncols"], grid["nrows"], grid["nlays"]
    n = nc * nr_ * nl
    ns, nph, nhet = mech["numb_mech_spc"], mech["nphotab"], mech["nhetero"]
    idx = np.ascontiguousarray(cgrid_index, np.int32)
    conv = np.ascontiguousarray(convert_conc, np.int32)
    mw = np.ascontiguousarray(molwt, np.float64)
    f = {k: _f32(grid[k]) for k in ("cgrid", "ta", "dens", "pres", "qv", "rj")}
    f32p = C.POINTER(C.c_float)
    zone = _f32(grid["seawater_zone"]) if grid.get("seawater_zone") is not None else None
    ice = _f32(grid["seaice"]) if grid.get("seaice") is not None else None
    kh = grid.get("khetero")
    kh = np.ascontiguousarray(kh, np.float64) if kh is not None else None
    ld = kh.shape[1] if kh is not None else max(nhet, 1)
    out = {k: np.zeros(n) for k in ("temp", "pres_atm", "h2o", "sea")}
    out["rj"] = np.zeros((nph, n))
    out["het"] = np.zeros((max(nhet, 1), n))
    out["y"] = np.zeros((ns, n))

    def fp(a):
        return None if a is None else _p(a, f32p)

    lib().om_gather(ns, nph, nhet, nc, nr_, nl, _p(idx, _i32p), _p(conv, _i32p), _p(mw, _f64p),
                    fp(f["cgrid"]), fp(f["ta"]), fp(f["dens"]), fp(f["pres"]), fp(f["qv"]), fp(zone), fp(ice),
                    fp(f["rj"]), None if kh is None else _p(kh, _f64p), ld,
                    _p(out["temp"], _f64p), _p(out["pres_atm"], _f64p), _p(out["h2o"], _f64p),
                    _p(out["sea"], _f64p), _p(out["rj"], _f64p), _p(out["het"], _f64p), _p(out["y"], _f64p))
    return out


def chem_scatter(mech, grid, cgrid_index, convert_conc, molwt, y):
    """CGRID (a float32 copy of grid["cgrid"]) with the mechanism species
    overwritten from y [ns][n] (mechanism order)."""
    nc, nr_, nl = grid["ncols"], grid["nrows"], grid["nlays"]
    cg = np.array(grid["cgrid"], np.float32, copy=True, order="C")
    idx = np.ascontiguousarray(cgrid_index, np.int32)
    conv = np.ascontiguousarray(convert_conc, np.int32)
    mw = np.ascontiguousarray(molwt, np.float64)
    dens = _f32(grid["dens"])
    yy = np.ascontiguousarray(y, np.float64)
    f32p = C.POINTER(C.c_float)
    lib().om_scatter(mech["numb_mech_spc"], nc, nr_, nl, _p(idx, _i32p), _p(conv, _i32p), _p(mw, _f64p),
                     _p(dens, f32p), _p(yy, _f64p), _p(cg, f32p))
    return cg
