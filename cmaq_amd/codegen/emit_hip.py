"""Emit the fully unrolled gfx950 ROS3 kernels for one mechanism.

The generated translation unit holds, for the mechanism's day pattern:

* ``feval``   -- reaction rates then f(y) in gather form: each species
  accumulates its loss / production terms in exactly the order in which
  RBFEVAL scatters them (rbfeval.F:93-149), so the rounding is the
  reference's;
* ``jac_lu``  -- Jacobian, diagonal shift and sparse LU fused column by
  column: the Jacobian entries of column j (built in RBJACOB's scatter order,
  rbjacob.F:112-183), the shift CC2(jj) -= 1/(gamma*h) (rbsolver.F:377-381),
  then RBDECOMP's loop-1 updates, the pivot reciprocal and the L scaling
  (rbdecomp.F:86-136).  Each factored entry is written once;
* ``solve``   -- forward then backward substitution in RBSOLVE's term order
  (rbsolve.F:78-187) with the stage right-hand side formed on the fly;
* the persistent solver kernel instantiated for these functions.

Night cells are integrated with the day pattern: the extra entries are
exact zeros (their only contributions are photolysis terms with J = 0), so
every night result is bit-identical to the reference's night pattern.
"""

import struct

from .symbolic import analyse

# fence spacing in the generated code (rates, species, solve rows)
FEVAL_RATE_GROUP = 16
FEVAL_SPC_GROUP = 8
SOLVE_ROW_GROUP = 8

HASH_OFFSET = 0xcbf29ce484222325
HASH_PRIME = 0x100000001b3


def structure_hash(m):
    """FNV-1a over the tables the generated code depends on.

    The C++ runtime computes the same value from the arrays it receives in
    cmaq_ros3_init (cmaq_ros3.cpp: structure_hash)."""
    h = HASH_OFFSET

    def feed(b):
        nonlocal h
        for byte in b:
            h ^= byte
            h = (h * HASH_PRIME) & 0xFFFFFFFFFFFFFFFF

    def i32(v):
        feed(struct.pack("<i", int(v)))

    ns, nr, mxprd = m["numb_mech_spc"], m["nrxns"], m["mxprd"]
    i32(ns); i32(nr); i32(mxprd)
    for r in range(nr):
        i32(m["nreact"][r]); i32(m["nprdct"][r]); i32(m["irxbits"][r])
        for k in range(m["nreact"][r]):
            i32(m["irr"][r][k])
        for k in range(m["nprdct"][r]):
            i32(m["irr"][r][3 + k])
            feed(struct.pack("<d", float(m["sc"][r][k])))
    return h


def _lit(x):
    s = repr(float(x))
    if "e" not in s and "." not in s and "inf" not in s and "nan" not in s:
        s += ".0"
    return s


def emit(m, cname):
    sym = analyse(m)
    if sym.ischan != sym.ns:
        raise ValueError("species without Jacobian entries are not supported")
    ns, nr = sym.ns, sym.nr
    day = sym.day
    nnz = day.nnz
    out = []
    w = out.append
    h = structure_hash(m)

    # workspace layout (entries per lane)
    off = {}
    pos = 0
    for name, size in (("RK", nr), ("RATE", nr), ("Y", ns), ("YP", ns), ("F", ns),
                       ("K1", ns), ("K2", ns), ("K3", ns), ("LU", nnz), ("VD", ns)):
        off[name] = pos
        pos += size
    nws = pos

    w(f"// Generated by cmaq_amd/codegen/emit_hip.py for {m['mechname']}; do not edit.")
    w("#include \"../ros3_device.h\"")
    w(f"namespace ros3 {{ namespace {cname} {{")
    w(f"constexpr int NS = {ns}, NR = {nr}, NNZ = {nnz}, NWS = {nws};")
    w(f"constexpr unsigned long long STRUCT_HASH = 0x{h:016x}ull;")
    for k, v in off.items():
        w(f"constexpr int OFF_{k} = {v};")
    w("// every per-lane vector is addressed from the lane base pointer wb with a")
    w("// compile-time offset, so the compiler can tell all of them apart")
    w("#define AT(off, k) wb[((off) + (k)) * kWave]")
    w("#define W(p, k) AT(OFF_##p, k)")
    w("// compiler-only memory fence: bounds how far loads are merged or hoisted,")
    w("// which keeps the per-lane register footprint (and so occupancy) in check")
    w("#define FENCE() asm volatile(\"\" ::: \"memory\")")
    w("")

    # ---------------- feval ----------------
    w("// f(y) with y at workspace offset yoff (OFF_Y or OFF_YP)")
    w("__device__ __noinline__ void feval(double *wb, int yoff)")
    w("{")
    for q, r in enumerate(day.active):
        terms = [f"W(RK, {r})"] + [f"AT(yoff, {a})" for a in sym.reactants[r]]
        w(f"    W(RATE, {r}) = {' * '.join(terms)};")
        if q % FEVAL_RATE_GROUP == FEVAL_RATE_GROUP - 1:
            w("    FENCE();")
    w("    FENCE();")
    contrib = [[] for _ in range(ns)]
    for r in day.active:
        for a in sym.reactants[r]:
            contrib[a].append(f"a = a - W(RATE, {r});")
        for k, p in enumerate(sym.products[r]):
            c = sym.sc[r][k]
            if c == 1.0:
                contrib[p].append(f"a = a + W(RATE, {r});")
            else:
                contrib[p].append(f"a = a + {_lit(c)} * W(RATE, {r});")
    for i in range(ns):
        if contrib[i]:
            w("    { double a = 0.0; " + " ".join(contrib[i]) + f" W(F, {i}) = a; }}")
        else:
            w(f"    W(F, {i}) = 0.0;")
        if i % FEVAL_SPC_GROUP == FEVAL_SPC_GROUP - 1:
            w("    FENCE();")
    w("}")
    w("")

    # ---------------- jac_lu ----------------
    def explic(r, ial):
        re = sym.reactants[r]
        if len(re) == 1:
            return f"W(RK, {r})"
        if len(re) == 2:
            return f"(W(RK, {r}) * W(Y, {re[1 - ial]}))"
        if ial == 0:
            return f"(W(RK, {r}) * W(Y, {re[1]}) * W(Y, {re[2]}))"
        if ial == 1:
            return f"(W(RK, {r}) * W(Y, {re[0]}) * W(Y, {re[2]}))"
        return f"(W(RK, {r}) * W(Y, {re[1]}) * W(Y, {re[0]}))"

    jac_terms = [[] for _ in range(nnz)]
    for r in day.active:
        for slot, ial in day.jac_loss[r]:
            jac_terms[slot].append(f"a = a - {explic(r, ial)};")
        for slot, ial, iap in day.jac_prod[r]:
            if iap < 0:
                jac_terms[slot].append(f"a = a + {explic(r, ial)};")
            else:
                jac_terms[slot].append(f"a = a + {_lit(sym.sc[r][iap])} * {explic(r, ial)};")

    col_slots = [[] for _ in range(ns)]
    for s in range(nnz):
        col_slots[day.cols[s]].append(s)
    w("__device__ __noinline__ void jac_lu(double *wb, double gdtinv)")
    w("{")
    for j in range(ns):
        w(f"    {{ // column {j}")
        for s in col_slots[j]:
            terms = jac_terms[s]
            if terms:
                body = " ".join(terms)
                if s == day.diag[j]:
                    body += " a = a - gdtinv;"
                w(f"        double e{s}; {{ double a = 0.0; {body} e{s} = a; }}")
            else:
                if s == day.diag[j]:
                    w(f"        double e{s} = 0.0 - gdtinv;")
                else:
                    w(f"        double e{s} = 0.0;")
        beg, end = day.lu_col_ops[j]
        for (ij, ik, kj) in day.lu_ops[beg:end]:
            w(f"        e{ij} = e{ij} - W(LU, {ik}) * e{kj};")
        d = day.diag[j]
        w(f"        const double v = 1.0 / e{d}; W(VD, {j}) = v;")
        for s in day.l_scale[j]:
            w(f"        e{s} = e{s} * v;")
        for s in col_slots[j]:
            w(f"        W(LU, {s}) = e{s};")
        w("    }")
        w("    FENCE();")
    w("}")
    w("")

    # ---------------- solve ----------------
    w("// stage solve into x at workspace offset xoff;")
    w("// mode 1: rhs = -f;  2: rhs = -f - x1*k1;  3: rhs = -f - x1*k1 - x2*k2")
    w("__device__ __noinline__ void solve(double *wb, int xoff, double x1, double x2, int mode)")
    w("{")
    for i in range(ns):
        terms = " ".join(f"r = r - W(LU, {s}) * AT(xoff, {c});" for (s, c) in day.lower[i])
        w(f"    {{ double r = -W(F, {i}); if (mode >= 2) r = r - x1 * W(K1, {i}); "
          f"if (mode == 3) r = r - x2 * W(K2, {i}); {terms} AT(xoff, {i}) = r; }}")
        if i % SOLVE_ROW_GROUP == SOLVE_ROW_GROUP - 1:
            w("    FENCE();")
    for i in range(ns - 1, -1, -1):
        terms = " ".join(f"r = r - W(LU, {s}) * AT(xoff, {c});" for (s, c) in day.upper[i])
        w(f"    {{ double r = AT(xoff, {i}); {terms} AT(xoff, {i}) = r * W(VD, {i}); }}")
        if i % SOLVE_ROW_GROUP == 0:
            w("    FENCE();")
    w("}")
    w("")

    # ---------------- cell load / store ----------------
    w("__device__ __noinline__ void load_cell(const SolveParams &P, int64_t cell, double *wb)")
    w("{")
    w("    const int64_t n = P.ncells;")
    w(f"    for (int r = 0; r < NR; ++r) W(RK, r) = P.rki[r * n + cell];")
    for i in range(ns):
        w(f"    W(Y, {i}) = P.y[{sym.inew2old[i]} * n + cell];")
    w("}")
    w("__device__ __noinline__ void store_cell(const SolveParams &P, int64_t cell, const double *wb)")
    w("{")
    w("    const int64_t n = P.ncells;")
    for i in range(ns):
        w(f"    P.y[{sym.inew2old[i]} * n + cell] = W(Y, {i});")
    w("}")
    w(f"const int32_t INEW2OLD_HOST[NS] = {{{', '.join(str(v) for v in sym.inew2old)}}};")
    w("")
    w("#include \"../ros3_solver_kernel.inc\"")
    w("#undef W")
    w("#undef AT")
    w(f"}} }} // namespace ros3::{cname}")
    w("")
    # registration record consumed by the runtime
    w(f"extern \"C\" const ros3::MechKernels ros3_kernels_{cname} = {{")
    w(f"    \"{m['mechname']}\", ros3::{cname}::STRUCT_HASH, ros3::{cname}::NS, ros3::{cname}::NR,")
    w(f"    ros3::{cname}::NNZ, {len(day.lu_ops)}, {sym.night.nnz}, {len(sym.night.lu_ops)}, ros3::{cname}::NWS,")
    w(f"    (const void *)&ros3::{cname}::solve_kernel, ros3::{cname}::INEW2OLD_HOST}};")
    return "\n".join(out) + "\n", sym
