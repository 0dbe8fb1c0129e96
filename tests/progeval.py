"""Evaluate the generated operation lists (cmaq_amd/codegen/programs.py) in
Python double precision.  Python floats round every operation separately, as
the generated HIP (-ffp-contract=off) and the oracle (gcc -ffp-contract=off)
do, so the evaluated programs can be compared with the oracle bit for bit."""
import re

from cmaq_amd.codegen import programs as P

CONST = dict(GAM=0.43586652150845899941601945119356e+00, C21=-0.10156171083877702091975600115545e+01,
             C31=0.40759956452537699824805835358067e+01, C32=0.92076794298330791242156818474003e+01,
             B1=1.0, B2=0.61697947043828245592553615689730e+01, B3=-0.42772256543218573326238373806514e+00,
             CONMIN=1.0e-30)
_TERN = re.compile(r"^(.*?) = (.*) > CONMIN \? (.*) : 0\.0;$")
# SMVGEAR's clamp of GLOSS at species at or below ZBOUND (grsubfun.F:165-171)
_GCLAMP = re.compile(r"^(.*?) = \((.*) <= ZBOUND && (.*) < 0\.0\) \? 0\.0 : (.*);$")


def _name(a):
    return f"{a[0]}[{a[1]}]" if isinstance(a, tuple) else a


def compile_ops(ops, params=()):
    lines = []
    for op in ops:
        s = op.fmt.format(*[_name(a) for a in op.args])
        m = _TERN.match(s)
        g = _GCLAMP.match(s)
        if g:
            s = f"{g.group(1)} = 0.0 if ({g.group(2)} <= ZBOUND and {g.group(3)} < 0.0) else {g.group(4)}"
        elif m:
            s = f"{m.group(1)} = {m.group(3)} if ({m.group(2)}) > CONMIN else 0.0"
        lines.append("    " + s.rstrip(";"))
    src = f"def _f(W, {', '.join(params) + ', ' if params else ''}**_):\n"
    arrays = sorted({a[0] for op in ops for a in op.args if isinstance(a, tuple)})
    src += "".join(f"    {a} = W['{a}']\n" for a in arrays)
    src += "\n".join(lines) + "\n"
    ns = dict(CONST)
    ns["ZBOUND"] = 1.0e-30
    ns["opaque"] = lambda v: v
    exec(compile(src, "<ops>", "exec"), ns)
    return ns["_f"]


class Programs:
    """The forward and adjoint programs of one mechanism as Python functions
    on a workspace dict of lists (species in the solver's order)."""

    def __init__(self, sym, jac_order="column", fuse_k1=False):
        self.sym = sym
        fe = P.feval_ops(sym, "scatter")
        self.feval = compile_ops(fe)
        self.jac_lu = compile_ops(P.jac_lu_ops(sym, jac_order, fuse_k1), ["gdtinv"])
        self.solve = [compile_ops(P.solve_ops(sym, s, fuse_k1 and s == 1), ["x1", "x2"]) for s in (1, 2, 3)]
        self.yp = compile_ops(P.yp_ops(sym))
        self.seed = compile_ops(P.adj_seed_ops(sym))
        self.ts = {x: compile_ops(P.tsolve_ops(sym, x)) for x in ("U1", "U2", "U3")}
        self.u2 = compile_ops(P.adj_u2_ops(sym), ["x32"])
        self.a = compile_ops(P.adj_a_ops(sym))
        self.u1 = compile_ops(P.adj_u1_ops(sym), ["x21", "x31"])
        self.b = compile_ops(P.adj_b_ops(sym))
        self.t_rhs1 = compile_ops(P.tlm_rhs1_ops(sym))
        self.t_stage = compile_ops(P.tlm_stage_ops(sym))
        self.t_rhs = {s: compile_ops(P.tlm_rhs_ops(sym, s), ["x1", "x2"]) for s in (2, 3)}
        self.ms = {x: compile_ops(P.msolve_ops(sym, x)) for x in ("U1", "U2", "U3")}
        self.t_update = compile_ops(P.tlm_update_ops(sym))

    def workspace(self, rk, y):
        ns, nr, nnz = self.sym.ns, self.sym.nr, self.sym.day.nnz
        W = {k: [0.0] * ns for k in ("Y", "YP", "F", "K1", "K2", "K3", "VD", "LAM", "U1", "U2", "U3", "YB")}
        W.update(RK=list(map(float, rk)), RATE=[0.0] * nr, LU=[0.0] * nnz, GK=[0.0] * nr,
                 S2=[0.0] * nr, S3=[0.0] * nr, T=[0.0] * 8)
        W["Y"] = list(map(float, y))
        return W

    def stages(self, W, h):
        dtinv = 1.0 / h
        gdtinv = dtinv * (1.0 / CONST["GAM"])
        W["YS"] = W["Y"]
        self.feval(W)
        self.jac_lu(W, gdtinv)
        self.solve[0](W, 0.0, 0.0)
        self.yp(W)
        W["YS"] = W["YP"]
        self.feval(W)
        self.solve[1](W, CONST["C21"] * dtinv, 0.0)
        self.solve[2](W, CONST["C31"] * dtinv, CONST["C32"] * dtinv)

    def ynew(self, W):
        c = CONST
        out = []
        for y, a, b, d in zip(W["Y"], W["K1"], W["K2"], W["K3"]):
            v = ((y + c["B1"] * a) + c["B2"] * b) + c["B3"] * d
            out.append(v if v > c["CONMIN"] else c["CONMIN"])
        return out

    def adjoint_step(self, W, h):
        """W holds Y (step start), LAM (dL/dYnew); on exit LAM = dL/dY, GK += dL/dk."""
        self.stages(W, h)
        dtinv = 1.0 / h
        self.seed(W)
        self.ts["U3"](W)
        self.u2(W, CONST["C32"] * dtinv)
        self.ts["U2"](W)
        self.a(W)
        self.u1(W, CONST["C21"] * dtinv, CONST["C31"] * dtinv)
        self.ts["U1"](W)
        self.b(W)

    def tlm_step(self, W, h):
        """W holds Y (step start), LAM = dY, GK = dk; on exit LAM = dYnew
        (Y itself is not advanced)."""
        self.stages(W, h)
        dtinv = 1.0 / h
        c = CONST
        W["YP"] = [((y + c["B1"] * a) + c["B2"] * b) + c["B3"] * d
                   for y, a, b, d in zip(W["Y"], W["K1"], W["K2"], W["K3"])]
        self.t_rhs1(W)
        self.ms["U1"](W)
        self.t_stage(W)
        self.t_rhs[2](W, c["C21"] * dtinv, 0.0)
        self.ms["U2"](W)
        self.t_rhs[3](W, c["C31"] * dtinv, c["C32"] * dtinv)
        self.ms["U3"](W)
        self.t_update(W)
