"""Emit the fully unrolled gfx950 ROS3 kernels for one mechanism.

Two translation units are generated per mechanism and variant, so that the
halves of a large mechanism compile in parallel.  The first holds, for the
mechanism's day pattern, straight-line device functions built from
programs.py's operation lists:

* ``feval``    -- f(y) in RBFEVAL's order (rbfeval.F:93-149);
* ``jac_lu``   -- Jacobian, diagonal shift and sparse LU (rbjacob.F,
  rbsolver.F:377-381, rbdecomp.F);
* ``solve1..3`` -- the three stage solves (rbsolve.F) with the stage
  right-hand sides of rbsolver.F:386-422 formed in place;
* the tangent-linear step (programs.py: tlm_rhs1_ops .. tlm_update_ops);
* the persistent solver kernel (ros3_solver_kernel.inc) that calls them, in
  its forward, tape-replay and tangent-linear forms, and the registry record.
The second holds the discrete adjoint of one accepted step (programs.py:
yp_ops .. adj_b_ops) and the persistent reverse kernel
(ros3_adjoint_kernel.inc), whose host stub the registry record of the first
refers to.

Registers are assigned by regalloc.allocate: at most REG_BUDGET doubles of a
lane's workspace are register-resident, loads are issued LOOKAHEAD operations
ahead of their use, and every workspace access is a buffer load/store with a
compile-time scalar offset from the wave's workspace base, so no per-access
address arithmetic is spent in the vector ALU.

Night cells are integrated with the day pattern: the extra entries are exact
zeros (their only contributions are photolysis terms with J = 0), so every
night result is bit-identical to the reference's night pattern.
"""

import functools
import re
import struct

from .programs import (adj_a_ops, adj_b_ops, adj_seed_ops, adj_u1_ops, adj_u2_ops, feval_ops,
                       jac_lu_ops, msolve_ops, solve_ops, tlm_rhs1_ops, tlm_rhs_ops, tlm_stage_ops,
                       tlm_update_ops, tsolve_ops, yp_ops)
from .regalloc import Op, allocate as _allocate
from .symbolic import analyse

# register budget (doubles per lane held by a generated function) and load
# lookahead (operations); the kernel's occupancy target follows the budget
REG_BUDGET = 48
LOOKAHEAD = 8
JAC_ORDER = "column"
FEVAL_ORDER = "scatter"

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


TLM_MAX = 8   # tangent-linear directions carried by one call (kTlmMaxDirections)


def _function(name, params, items, off, budget, dynamic=None, sched=0, load_wrap=None):
    """HIP source of one generated function from allocator items.

    sched > 0 closes the compiler's instruction-scheduling region after every
    `sched` statements: the allocator has already placed each load `lookahead`
    operations ahead of its use, and the scheduler's cost grows faster than
    linearly with the region length (one region per function otherwise)."""
    dynamic = dynamic or {}
    w = [f"__device__ __noinline__ void {name}(double *base_{params})", "{",
         "    ROS3_PROLOGUE"]
    for d, src in dynamic.items():
        w.append(f"    const int {d}_off = __builtin_amdgcn_readfirstlane({src});")
    w.append("    double " + ", ".join(f"r{i}" for i in range(budget)) + ";")

    def addr(slot):
        arr, i = slot
        if arr in dynamic:
            return f"({arr}_off + {i})"
        return str(off[arr] + i)

    if any(it[0] in ("lds_load", "lds_store") for it in items):
        # this wave's slice of the LDS tier: entry q of the lane at lt[q * 64]
        w.append("    double *const lt = lds_tier + (int)(threadIdx.x >> 6) * (LDS_SLOTS * kWave) + __lane_id();")
    for k, it in enumerate(items):
        if sched and k and k % sched == 0:
            w.append("    __builtin_amdgcn_sched_barrier(0);")
        if it[0] == "load" and load_wrap and it[2][0] in load_wrap:
            # a read-only derived array: its slots are loaded from the
            # source array's rows through a macro (e.g. CEFF of CNEW)
            src, macro = load_wrap[it[2][0]]
            w.append(f"    r{it[1]} = {macro}(LD({off[src] + it[2][1]}));")
        elif it[0] == "load":
            w.append(f"    r{it[1]} = LD({addr(it[2])});")
        elif it[0] == "store":
            w.append(f"    ST({addr(it[2])}, r{it[1]});")
        elif it[0] == "lds_load":
            w.append(f"    r{it[1]} = lt[{it[2]} * kWave];")
        elif it[0] == "lds_store":
            w.append(f"    lt[{it[2]} * kWave] = r{it[1]};")
        else:
            w.append("    " + it[1])
    w.append("}")
    return w


def _remap(ops, arrays=None, scalars=None):
    """The same operations on renamed arrays / scalar parameters."""
    arrays, scalars = arrays or {}, scalars or {}
    pat = re.compile(r"\b(" + "|".join(map(re.escape, scalars)) + r")\b") if scalars else None
    out = []
    for op in ops:
        fmt = pat.sub(lambda m: scalars[m.group(1)], op.fmt) if pat else op.fmt
        args = [(arrays.get(a[0], a[0]), a[1]) if isinstance(a, tuple) else scalars.get(a, a) for a in op.args]
        out.append(Op(fmt, args, op.rmw))
    return out


def _stat(ops, counts):
    return (len(ops), counts.get("load", 0), counts.get("store", 0)) + \
        ((counts.get("lds_load", 0), counts.get("lds_store", 0)) if "lds_load" in counts else ())


def waves_per_eu(budget):
    """Occupancy the kernel is compiled for: VGPRs ~ 2 per resident double plus
    ~28 for addresses, loop state and temporaries, out of 512 per lane."""
    vgprs = 2 * budget + 28
    vgprs = (vgprs + 7) // 8 * 8
    return max(1, min(8, 512 // vgprs))


def emit(m, cname, budget=None, lookahead=None, variant="", is_default=True, refill_min=0,
         jac_order=None, gear_nws=0):
    budget = budget or REG_BUDGET
    lookahead = lookahead or LOOKAHEAD
    jac_order = jac_order or JAC_ORDER
    sym = analyse(m)
    if sym.ischan != sym.ns:
        raise ValueError("species without Jacobian entries are not supported")
    ns, nr = sym.ns, sym.nr
    day = sym.day
    nnz = day.nnz

    off = {}
    pos = 0
    for name, size in (("RK", nr), ("RATE", nr), ("Y", ns), ("YP", ns), ("F", ns),
                       ("K1", ns), ("K2", ns), ("K3", ns), ("LU", nnz), ("VD", ns),
                       ("LAM", ns), ("U1", ns), ("U2", ns), ("U3", ns), ("YB", ns),
                       ("GK", nr), ("S2", nr), ("S3", nr), ("T", 8),
                       # tangent-linear directions: dY (ns) then dk (nr) per direction
                       ("TLM", TLM_MAX * (ns + nr))):
        off[name] = pos
        pos += size
    nws = pos
    h = structure_hash(m)
    stats = {}
    entry_all = lambda slot: True
    body = []
    b = body.append

    # "<order>+k1": stage 1's forward substitution fused into the factorization;
    # "+step": the whole attempted step (f(Y), Jacobian + LU, three stage
    # solves, f(YP)) as one function, so registers carry values across the
    # phases and the callee-saved registers are saved once per step
    # the factored pivots: the solves read their reciprocals VD, never the
    # pivots themselves, so the factorization does not write them back
    piv = frozenset(("LU", day.diag[j]) for j in range(ns))
    base_order = jac_order.split("+")[0]
    flags = jac_order.split("+")[1:]
    merge = "step" in flags
    fuse_k1 = "k1" in flags or merge
    # "+ldsN": N doubles per lane of LDS as a second tier for the forward
    # functions (clean copies of evicted values that are read again)
    lds = next((int(f[3:]) for f in flags if f.startswith("lds")), 0)
    # "+fN": prefetch as far ahead (at most `lookahead` operations) as keeps N
    # loads outstanding, instead of a fixed distance in operations: the same
    # requests in flight with fewer registers held by prefetched values
    inflight = next((int(f[1:]) for f in flags if f[:1] == "f" and f[1:].isdigit()), 0)
    # "+sbN": a scheduling barrier every N statements (see _function)
    sched = next((int(f[2:]) for f in flags if f.startswith("sb")), 0)
    fn = functools.partial(_function, sched=sched)
    allocate = functools.partial(_allocate, inflight=inflight)

    def stage1_backward():
        # YP = Y + K1 (rbsolver.F:395-398) as soon as K1(i) is final
        out = []
        for op in solve_ops(sym, 1, backward_only=fuse_k1):
            out.append(op)
            if op.fmt == "{0} = {0} * {1};" and op.args[0][0] == "K1":
                i = op.args[0][1]
                out.append(Op("{0} = {1} + {2};", [("YP", i), ("Y", i), ("K1", i)], False))
        return out

    if merge:
        step = (_remap(feval_ops(sym, FEVAL_ORDER), {"YS": "Y"}) + jac_lu_ops(sym, base_order, True) +
                stage1_backward() + _remap(feval_ops(sym, FEVAL_ORDER), {"YS": "YP"}) +
                _remap(solve_ops(sym, 2), scalars={"x1": "x21"}) +
                _remap(solve_ops(sym, 3), scalars={"x1": "x31", "x2": "x32"}))
        items, cnt = allocate(step, {"K1", "K2", "K3"}, budget, lookahead, entry_all, lds)
        stats["ros3_step"] = _stat(step, cnt)
        body += fn("ros3_step", ", double gdtinv, double x21, double x31, double x32", items, off, budget)
        b("")
    else:
        # f(y): y at a runtime workspace offset (OFF_Y or OFF_YP)
        ops = feval_ops(sym, FEVAL_ORDER)
        items, cnt = allocate(ops, {"F"}, budget, lookahead, entry_all, lds)
        stats["feval"] = _stat(ops, cnt)
        body += fn("feval", ", int ys_off_", items, off, budget, {"YS": "ys_off_"})
        b("")
        ops = jac_lu_ops(sym, base_order, fuse_k1)
        items, cnt = allocate(ops, {"LU", "VD", "K1"} if fuse_k1 else {"LU", "VD"}, budget, lookahead, entry_all, lds,
                              dead=piv)
        stats["jac_lu"] = _stat(ops, cnt)
        jac_lu_text = fn("jac_lu", ", double gdtinv", items, off, budget)
        body += jac_lu_text
        b("")
        for stage in (1, 2, 3):
            ops = stage1_backward() if stage == 1 else solve_ops(sym, stage)
            outs = {"K1", "YP"} if stage == 1 else {f"K{stage}"}
            items, cnt = allocate(ops, outs, budget, lookahead, entry_all, lds)
            stats[f"solve{stage}"] = _stat(ops, cnt)
            params = "" if stage == 1 else (", double x1" if stage == 2 else ", double x1, double x2")
            body += fn(f"solve{stage}", params, items, off, budget)
            b("")
    # the reverse kernel and its functions form a second translation unit so
    # that the two halves of a large mechanism compile in parallel
    body_adj = []
    if fuse_k1:
        # the adjoint's refactorization has no f(Y) to substitute
        ops = jac_lu_ops(sym, base_order)
        items, cnt = allocate(ops, {"LU", "VD"}, budget, lookahead, entry_all, dead=piv)
        body_adj += fn("jac_lu_plain", ", double gdtinv", items, off, budget)
        body_adj.append("")
    elif not merge:
        body_adj += jac_lu_text
        body_adj.append("")

    # discrete adjoint of one accepted step (second translation unit), then
    # the tangent-linear step (first)
    for fname, ops, outs, params, dyn in (
            ("yp_form", yp_ops(sym), {"YP"}, "", None),
            ("adj_seed", adj_seed_ops(sym), {"LAM", "U3"}, "", None),
            ("tsolve", tsolve_ops(sym, "X"), {"X"}, ", int x_off_", {"X": "x_off_"}),
            ("adj_u2", adj_u2_ops(sym), {"U2"}, ", double x32", None),
            ("adj_a", adj_a_ops(sym), {"YB", "S2", "S3"}, "", None),
            ("adj_u1", adj_u1_ops(sym), {"U1"}, ", double x21, double x31", None),
            ("adj_b", adj_b_ops(sym), {"LAM", "GK"}, "", None),
            # tangent-linear direction through one accepted step; the
            # direction's dY and dk sit at runtime offsets in the TLM region
            ("tlm_rhs1", tlm_rhs1_ops(sym), {"U1"}, ", int dy_off_, int dk_off_",
             {"LAM": "dy_off_", "GK": "dk_off_"}),
            ("msolve", msolve_ops(sym, "X"), {"X"}, ", int x_off_", {"X": "x_off_"}),
            ("tlm_stage", tlm_stage_ops(sym), {"S2"}, ", int dy_off_, int dk_off_",
             {"LAM": "dy_off_", "GK": "dk_off_"}),
            ("tlm_rhs2", tlm_rhs_ops(sym, 2), {"U2"}, ", int dy_off_, int dk_off_, double x1",
             {"LAM": "dy_off_", "GK": "dk_off_"}),
            ("tlm_rhs3", tlm_rhs_ops(sym, 3), {"U3"}, ", int dy_off_, int dk_off_, double x1, double x2",
             {"LAM": "dy_off_", "GK": "dk_off_"}),
            ("tlm_update", tlm_update_ops(sym), {"LAM"}, ", int dy_off_", {"LAM": "dy_off_"})):
        items, cnt = allocate(ops, outs, budget, lookahead, entry_all)
        stats[fname] = _stat(ops, cnt)
        dest = body if fname.startswith(("tlm_", "msolve")) else body_adj
        dest += fn(fname, params, items, off, budget, dyn)
        dest.append("")
    b("__device__ __noinline__ void tlm_step(double *base, double dtinv, int ndir)")
    b("{")
    b("    for (int p = 0; p < ndir; ++p) {")
    b("        const int dy = OFF_TLM + p * TLM_RS, dk = dy + NS;")
    b("        tlm_rhs1(base, dy, dk);")
    b("        msolve(base, OFF_U1);")
    b("        tlm_stage(base, dy, dk);")
    b("        tlm_rhs2(base, dy, dk, C21 * dtinv);")
    b("        msolve(base, OFF_U2);")
    b("        tlm_rhs3(base, dy, dk, C31 * dtinv, C32 * dtinv);")
    b("        msolve(base, OFF_U3);")
    b("        tlm_update(base, dy);")
    b("    }")
    b("}")
    b("")

    out = []
    w = out.append
    w(f"// Generated by cmaq_amd/codegen/emit_hip.py for {m['mechname']}; do not edit.")
    w("// @UNIT@")
    w(f"// register budget {budget} doubles, load lookahead {lookahead} operations, LU schedule {jac_order};")
    w("// per call (operations, loads, stores): " + ", ".join(f"{k} {v}" for k, v in stats.items()))
    w("#include \"../ros3_device.h\"")
    w("#include \"../ros3_gear.h\"")
    w(f"namespace ros3 {{ namespace {cname} {{")
    w(f"constexpr int NS = {ns}, NR = {nr}, NNZ = {nnz}, NWS = {nws};")
    w(f"constexpr int TLM_MAX = {TLM_MAX}, TLM_RS = NS + NR;   // tangent-linear directions per call")
    w("static_assert(TLM_MAX == kTlmMaxDirections, \"TLM_MAX\");")
    w(f"constexpr int REG_BUDGET = {budget}, WAVES_PER_EU = {waves_per_eu(budget)};")
    w(f"constexpr unsigned long long STRUCT_HASH = 0x{h:016x}ull;")
    for k, v in off.items():
        w(f"constexpr int OFF_{k} = {v};")
    w("// Workspace access: the wave's workspace is one buffer resource; entry k of")
    w("// the calling lane sits at byte (k * 64 + lane) * 8 from the wave's base.")
    w("#define ROS3_PROLOGUE \\")
    w("    double *base = (double *)((((unsigned long long)(unsigned int)__builtin_amdgcn_readfirstlane( \\")
    w("                                  (int)((unsigned long long)base_ >> 32))) << 32) | \\")
    w("                              (unsigned long long)(unsigned int)__builtin_amdgcn_readfirstlane( \\")
    w("                                  (int)(unsigned long long)base_)); \\")
    w("    const __amdgpu_buffer_rsrc_t rs = \\")
    w("        __builtin_amdgcn_make_buffer_rsrc(base, (short)0, NWS * kWave * 8, 0x00020000); \\")
    w("    const int vo = __lane_id() * 8;")
    w("#define LD(k) __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, vo, (k) * (kWave * 8), 0))")
    w("#define ST(k, v) __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, (double)(v)), rs, vo, (k) * (kWave * 8), 0)")
    if lds:
        w(f"constexpr int LDS_SLOTS = {lds};")
        w("__shared__ double lds_tier[(256 / kWave) * LDS_SLOTS * kWave];   // 4 waves per workgroup")
    w(f"constexpr int TAPE_RS = {4 * ns + 2};   // doubles per adjoint tape record")
    w("")
    header = list(out)
    # ---- second unit: the reverse sweep of the discrete adjoint ----
    adj = [line.replace("// @UNIT@", "// unit 2 of 2: discrete adjoint (reverse_kernel)") for line in header
           if not line.startswith(("constexpr int LDS_SLOTS", "__shared__ double lds_tier"))]
    wa = adj.append
    adj += body_adj
    # adjoint cell load / store; tape records (one per accepted step: Y, K1,
    # K2, K3 and the step size, contiguous per cell) are written by unit 1
    wa("__device__ __noinline__ void load_adj_cell(const SolveParams &P, int64_t cell, double *base_)")
    wa("{")
    wa("    ROS3_PROLOGUE")
    wa("    const int64_t n = P.stride;")
    wa("    for (int r = 0; r < NR; ++r) { ST(OFF_RK + r, P.rki[r * n + cell]); ST(OFF_GK + r, 0.0); }")
    for i in range(ns):
        wa(f"    ST(OFF_LAM + {i}, P.lam[{sym.inew2old[i]} * n + cell]);")
    wa("}")
    wa("__device__ __noinline__ void store_adj_cell(const SolveParams &P, int64_t cell, double *base_)")
    wa("{")
    wa("    ROS3_PROLOGUE")
    wa("    const int64_t n = P.stride;")
    for i in range(ns):
        wa(f"    P.lam[{sym.inew2old[i]} * n + cell] = LD(OFF_LAM + {i});")
    wa("    if (P.grad_rki)")
    wa("        for (int r = 0; r < NR; ++r) P.grad_rki[r * n + cell] += LD(OFF_GK + r);")
    wa("}")
    wa("__device__ __noinline__ double tape_load(double *base_, const double *rec)")
    wa("{")
    wa("    ROS3_PROLOGUE")
    for k, arr in enumerate(("Y", "K1", "K2", "K3")):
        wa(f"    for (int i = 0; i < NS; ++i) ST(OFF_{arr} + i, rec[{k} * NS + i]);")
    wa("    return rec[4 * NS];")
    wa("}")
    wa(f"#define ROS3_JAC_LU_PLAIN {'jac_lu_plain' if fuse_k1 else 'jac_lu'}")
    wa("#include \"../ros3_adjoint_kernel.inc\"")
    for u in ("LD", "ST", "ROS3_PROLOGUE", "ROS3_JAC_LU_PLAIN"):
        wa(f"#undef {u}")
    wa(f"}} }} // namespace ros3::{cname}")
    # ---- first unit: forward, tape replay and tangent-linear kernels ----
    out[1] = "// unit 1 of 2: forward, replay and tangent-linear kernels; registry record"
    out += body
    # cell load / store: inputs and results are cell-fastest global arrays
    w("__device__ __noinline__ void load_cell(const SolveParams &P, int64_t cell, double *base_)")
    w("{")
    w("    ROS3_PROLOGUE")
    w("    const int64_t n = P.stride;")
    w("    for (int r = 0; r < NR; ++r) ST(OFF_RK + r, P.rki[r * n + cell]);")
    for i in range(ns):
        w(f"    ST(OFF_Y + {i}, P.y[{sym.inew2old[i]} * n + cell]);")
    w("}")
    w("__device__ __noinline__ void store_cell(const SolveParams &P, int64_t cell, double *base_)")
    w("{")
    w("    ROS3_PROLOGUE")
    w("    const int64_t n = P.stride;")
    for i in range(ns):
        w(f"    P.y[{sym.inew2old[i]} * n + cell] = LD(OFF_Y + {i});")
    w("}")
    w("__device__ __noinline__ void load_tlm_cell(const SolveParams &P, int64_t cell, double *base_)")
    w("{")
    w("    ROS3_PROLOGUE")
    w("    const int64_t n = P.stride;")
    w("    for (int p = 0; p < P.ndir; ++p) {")
    w("        const int o = OFF_TLM + p * TLM_RS;")
    w("        const double *dy = P.lam + (int64_t)p * NS * n + cell;")
    w("        for (int r = 0; r < NR; ++r) ST(o + NS + r, P.dk ? P.dk[((int64_t)p * NR + r) * n + cell] : 0.0);")
    for i in range(ns):
        w(f"        ST(o + {i}, dy[{sym.inew2old[i]} * n]);")
    w("    }")
    w("}")
    w("__device__ __noinline__ void store_tlm_cell(const SolveParams &P, int64_t cell, double *base_, bool failed)")
    w("{")
    w("    ROS3_PROLOGUE")
    w("    const int64_t n = P.stride;")
    w("    for (int p = 0; p < P.ndir; ++p) {")
    w("        const int o = OFF_TLM + p * TLM_RS;")
    w("        double *dy = P.lam + (int64_t)p * NS * n + cell;")
    for i in range(ns):
        w(f"        dy[{sym.inew2old[i]} * n] = failed ? __builtin_nan(\"\") : LD(o + {i});")
    w("    }")
    w("}")
    # the time-averaged state of the -Dsens build, mechanism order (rbsolver.F:592-596)
    w("__device__ __noinline__ void store_yave(const SolveParams &P, int64_t cell, double *base_, double chemstep,")
    w("                                        bool failed)")
    w("{")
    w("    ROS3_PROLOGUE")
    w("    const int64_t n = P.stride;")
    for i in range(ns):
        w(f"    P.yave[{sym.inew2old[i]} * n + cell] = failed ? __builtin_nan(\"\") : LD(OFF_LAM + {i}) / chemstep;")
    w("}")
    w("__device__ __noinline__ void tape_store(double *base_, double *rec, double h)")
    w("{")
    w("    ROS3_PROLOGUE")
    for k, arr in enumerate(("Y", "K1", "K2", "K3")):
        w(f"    for (int i = 0; i < NS; ++i) rec[{k} * NS + i] = LD(OFF_{arr} + i);")
    w("    rec[4 * NS] = h;")
    w("}")
    w(f"const int32_t INEW2OLD_HOST[NS] = {{{', '.join(str(v) for v in sym.inew2old)}}};")
    w("")
    if merge:
        w("#define ROS3_MERGED_STEP 1")
    w("#include \"../ros3_solver_kernel.inc\"")
    w("__global__ __launch_bounds__(256, WAVES_PER_EU) void reverse_kernel(SolveParams P);   // unit 2")
    if gear_nws:
        w(f"}} namespace {cname}_gear {{")
        w("__global__ __launch_bounds__(256, " + str(waves_per_eu(budget)) + ") void gear_kernel(GearParams P);   // SMVGEAR unit")
        w("namespace tail { __global__ __launch_bounds__(kWave) void gear_tail_kernel(GearParams P); }")
        w(f"}} namespace {cname} {{")
    w("#undef LD")
    w("#undef ST")
    w("#undef ROS3_PROLOGUE")
    w("#undef ROS3_MERGED_STEP")
    w(f"}} }} // namespace ros3::{cname}")
    w("")
    # the record's kernel addresses make the device pass instantiate the
    # forward kernels; reverse_kernel's device code lives in unit 2, so only
    # the host pass takes its address
    w("#ifdef __HIP_DEVICE_COMPILE__")
    w("#define ROS3_REVERSE_KERNEL nullptr")
    w("#define ROS3_GEAR_KERNEL nullptr")
    w("#define ROS3_GEAR_TAIL_KERNEL nullptr")
    w("#else")
    w(f"#define ROS3_REVERSE_KERNEL (const void *)&ros3::{cname}::reverse_kernel")
    w(f"#define ROS3_GEAR_KERNEL (const void *)&ros3::{cname}_gear::gear_kernel" if gear_nws else
      "#define ROS3_GEAR_KERNEL nullptr")
    w(f"#define ROS3_GEAR_TAIL_KERNEL (const void *)&ros3::{cname}_gear::tail::gear_tail_kernel" if gear_nws else
      "#define ROS3_GEAR_TAIL_KERNEL nullptr")
    w("#endif")
    w(f"extern \"C\" const ros3::MechKernels ros3_kernels_{cname} = {{")
    w(f"    \"{m['mechname']}\", ros3::{cname}::STRUCT_HASH, ros3::{cname}::NS, ros3::{cname}::NR,")
    w(f"    ros3::{cname}::NNZ, {len(day.lu_ops)}, {sym.night.nnz}, {len(sym.night.lu_ops)}, ros3::{cname}::NWS,")
    w(f"    (const void *)&ros3::{cname}::solve_kernel_t<false>, ros3::{cname}::INEW2OLD_HOST,")
    w(f"    \"{variant}\", {budget}, {lookahead}, ros3::{cname}::WAVES_PER_EU, {1 if is_default else 0}, {refill_min},")
    w(f"    (const void *)&ros3::{cname}::solve_kernel_t<true>, ROS3_REVERSE_KERNEL,")
    w(f"    ros3::{cname}::TAPE_RS, (const void *)&ros3::{cname}::solve_kernel_t<false, true>,")
    w(f"    ros3::{cname}::OFF_LAM,   // rows of the workspace the forward kernel touches")
    w(f"    (const void *)&ros3::{cname}::solve_kernel_t<false, false, true>,")
    w(f"    ROS3_GEAR_KERNEL, {gear_nws}, ROS3_GEAR_TAIL_KERNEL,")
    w(f"    (const void *)&ros3::{cname}::solve_kernel_t<false, false, false, true>}};")
    w("#undef ROS3_REVERSE_KERNEL")
    w("#undef ROS3_GEAR_KERNEL")
    w("#undef ROS3_GEAR_TAIL_KERNEL")
    return ("\n".join(out) + "\n", "\n".join(adj) + "\n"), sym


def emit_gear(m, cname, budget=None, lookahead=None, jac_order=None):
    """The SMVGEAR translation unit of one mechanism (namespace
    ros3::<cname>_gear): SUBFUN, PDERIV + DECOMP and BACKSUB as straight-line
    functions over the same day pattern and LU schedule as the ROS3 kernels,
    and the persistent Gear kernel (ros3_gear_kernel.inc) that calls them."""
    budget = budget or REG_BUDGET
    lookahead = lookahead or LOOKAHEAD
    base_order = (jac_order or JAC_ORDER).split("+")[0]
    sym = analyse(m)
    ns, nr, nnz = sym.ns, sym.nr, sym.day.nnz
    from .programs import gear_pderiv_lu_ops, gear_subfun_ops
    from .regalloc import allocate as _alloc
    # rows that the generated functions touch: [entry][64 lanes] per wave, so
    # that a converged wavefront loads 512 contiguous bytes per row
    off, pos = {}, 0
    for name, size in (("RK", nr), ("RATE", nr), ("CN", ns), ("GL", ns), ("LU", nnz), ("VD", ns)):
        off[name] = pos
        pos += size
    nws = pos
    # rows only the state machine touches (CONC's 7 columns, DTLOS, CEST,
    # CHOLD, BLKCONC), which its lanes reach at different labels: contiguous
    # per lane after the wave's rows, so that a lane's chunk of 16 rows is one
    # 128-byte line whichever lanes are active
    loff, lpos = {}, 0
    for name, size in (("CO", 7 * ns), ("DTL", ns), ("CE", ns), ("CH", ns), ("B0", ns)):
        loff[name] = lpos
        lpos += size
    lws = (lpos + 15) // 16 * 16
    piv = frozenset(("LU", sym.day.diag[j]) for j in range(ns))
    entry_all = lambda slot: True
    stats, body = {}, []
    for fname, ops, outs, params, dead in (
            ("gear_subfun", gear_subfun_ops(sym), {"GL"}, "", frozenset()),
            ("gear_pderiv_lu", gear_pderiv_lu_ops(sym, base_order), {"LU", "VD"}, ", double r1delt", piv),
            ("gear_backsub", msolve_ops(sym, "GL"), {"GL"}, "", frozenset())):
        items, cnt = _alloc(ops, outs, budget, lookahead, entry_all, dead=dead)
        if any(it[0] == "store" and it[2][0] == "CF" for it in items):
            raise RuntimeError("CEFF slots are read-only")
        stats[fname] = _stat(ops, cnt)
        body += _function(fname, params, items, off, budget, load_wrap={"CF": ("CN", "CEFF")})
        body.append("")
    gname = cname + "_gear"
    w = []
    a = w.append
    a(f"// Generated by cmaq_amd/codegen/emit_hip.py for {m['mechname']}; do not edit.")
    a("// SMVGEAR kernel (gear_kernel) over the ROS3 kernels' species order and sparse pattern;")
    a(f"// register budget {budget} doubles, load lookahead {lookahead} operations, LU schedule {base_order};")
    a("// per call (operations, loads, stores): " + ", ".join(f"{k} {v}" for k, v in stats.items()))
    a("#include \"../ros3_device.h\"")
    a("#include \"../ros3_gear.h\"")
    a(f"namespace ros3 {{ namespace {gname} {{")
    a(f"constexpr int NS = {ns}, NR = {nr}, NNZ = {nnz}, NWS = {nws}, LWS = {lws};")
    a(f"constexpr int WAVES_PER_EU = {waves_per_eu(budget)};")
    a("static_assert(kGearCols == 7, \"CONC columns\");")
    for k, v in off.items():
        a(f"constexpr int GOFF_{k} = {v};")
    for k, v in loff.items():
        a(f"constexpr int LOFF_{k} = {v};")
    a("#define ROS3_PROLOGUE \\")
    a("    double *base = (double *)((((unsigned long long)(unsigned int)__builtin_amdgcn_readfirstlane( \\")
    a("                                  (int)((unsigned long long)base_ >> 32))) << 32) | \\")
    a("                              (unsigned long long)(unsigned int)__builtin_amdgcn_readfirstlane( \\")
    a("                                  (int)(unsigned long long)base_)); \\")
    a("    const __amdgpu_buffer_rsrc_t rs = \\")
    a("        __builtin_amdgcn_make_buffer_rsrc(base, (short)0, NWS * kWave * 8, 0x00020000); \\")
    a("    const int vo = __lane_id() * 8; \\")
    a("    double *const lw_ = base + NWS * kWave + __lane_id() * LWS; \\")
    a("    (void)lw_;")
    a("#define LD(k) __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, vo, (k) * (kWave * 8), 0))")
    a("#define ST(k, v) __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, (double)(v)), rs, vo, (k) * (kWave * 8), 0)")
    a("#define LDL(k) (lw_[(k)])")
    a("#define STL(k, v) (lw_[(k)] = (double)(v))")
    a("__device__ __forceinline__ double ceff(double v) { return v <= kGearZbound ? 0.0 : v; }   // grpderiv.F:121-129")
    a("#define CEFF(v) ceff(v)")
    a("#define ZBOUND kGearZbound")
    a("__device__ __forceinline__ double opaque(double v) { asm volatile(\"\" : \"+v\"(v)); return v; }")
    a("")
    w += body
    a("__device__ __noinline__ void load_gear_cell(const GearParams &P, int64_t cell, double *base_)")
    a("{")
    a("    ROS3_PROLOGUE")
    a("    const int64_t n = P.stride;")
    a("    for (int r = 0; r < NR; ++r) ST(GOFF_RK + r, P.rki[r * n + cell]);")
    for i in range(ns):
        a(f"    {{ const double v = P.y[{sym.inew2old[i]} * n + cell]; const double c = v > kGearConcmin ? v : kGearConcmin;"
          f" ST(GOFF_CN + {i}, c); STL(LOFF_B0 + {i}, c); }}")
    a("}")
    a("__device__ __noinline__ void store_gear_cell(const GearParams &P, int64_t cell, double *base_)")
    a("{")
    a("    ROS3_PROLOGUE")
    a("    const int64_t n = P.stride;")
    for i in range(ns):
        a(f"    P.y[{sym.inew2old[i]} * n + cell] = LD(GOFF_CN + {i});")
    a("}")
    a("#include \"../ros3_gear_kernel.inc\"")
    for u in ("LD", "ST", "LDL", "STL", "ROS3_PROLOGUE"):
        a(f"#undef {u}")
    # the same functions over one cell's workspace in LDS, for gear_tail_kernel
    a("namespace tail {")
    a("__shared__ double gear_lds[NWS + LWS];")
    a("static_assert((NWS + LWS) * 8 <= 160 * 1024, \"one cell's workspace fits the LDS of a CU\");")
    a("#define GEAR_TAIL")
    # Every access takes the LDS base through an empty asm that redefines
    # it: the compiler then keeps the allocator's loads and stores as they
    # are (with constant addresses it would forward stored values into the
    # registers the allocator had freed, and spill), and the base stays in
    # one VGPR, so an access is one ds_read / ds_write with a constant offset
    a("typedef __attribute__((address_space(3))) double lds_double;")
    a("#define ROS3_PROLOGUE lds_double *lb_ = (lds_double *)gear_lds; (void)base_;")
    a("#define LDS_AT(k) (*({ asm volatile(\"\" : \"+v\"(lb_)); lb_ + (k); }))")
    a("#define LD(k) LDS_AT(k)")
    a("#define ST(k, v) (LDS_AT(k) = (double)(v))")
    a("#define LDL(k) LDS_AT(NWS + (k))")
    a("#define STL(k, v) (LDS_AT(NWS + (k)) = (double)(v))")
    w += body
    a("__device__ __noinline__ void store_gear_cell(const GearParams &P, int64_t cell, double *base_)")
    a("{")
    a("    ROS3_PROLOGUE")
    a("    const int64_t n = P.stride;")
    for i in range(ns):
        a(f"    P.y[{sym.inew2old[i]} * n + cell] = LD(GOFF_CN + {i});")
    a("}")
    a("#include \"../ros3_gear_kernel.inc\"")
    for u in ("LD", "ST", "LDL", "STL", "LDS_AT", "ROS3_PROLOGUE", "GEAR_TAIL"):
        a(f"#undef {u}")
    a("}  // namespace tail")
    for u in ("CEFF", "ZBOUND"):
        a(f"#undef {u}")
    a(f"}} }} // namespace ros3::{gname}")
    return "\n".join(w) + "\n", nws + lws
