"""Per-step operation lists of the ROS3 kernel, in the reference's rounding order.

Each builder returns a list of regalloc.Op.  The arithmetic of every slot is
the sequence the reference applies to it:

* feval   -- RBFEVAL (rbfeval.F:93-149): zero dc/dt, then reaction by
             reaction: rate = RKI * y_1 [* y_2 [* y_3]]; each reactant loses
             the rate, each product gains SC * rate;
* jac_lu  -- RBJACOB (rbjacob.F:91-183) scatter order per entry, the diagonal
             shift (rbsolver.F:377-381), RBDECOMP loop 1 (rbdecomp.F:86-102),
             the pivot reciprocal (:105-108) and the L scaling (:115-136);
             the op *schedule* across entries is free: "column" keeps the
             reference's column order, "row" runs the same per-entry sequence
             row by row (entry (i,j) still sees its updates in ascending k);
* solve   -- RBSOLVE (rbsolve.F:78-187): each row sums its terms in storage
             order, the backward pass ends with the multiply by VDIAG; the
             stage right-hand side K1 = -F, K2 = -F - x1 K1, K3 = -F - x1 K1 -
             x2 K2 (rbsolver.F:386-422) is formed in the row's first statement.
"""

import re

from .regalloc import Op


def _lit(x):
    s = repr(float(x))
    if "e" not in s and "." not in s and "inf" not in s and "nan" not in s:
        s += ".0"
    return s


def feval_ops(sym, order="scatter"):
    """f(y) with y read from array 'YS' into 'F'."""
    day = sym.day
    ops = []
    if order == "scatter":
        for i in range(sym.ns):
            ops.append(Op("{0} = 0.0;", [("F", i)], False))
        for r in day.active:
            re = sym.reactants[r]
            args = [("RATE", r), ("RK", r)] + [("YS", a) for a in re]
            ops.append(Op("{0} = " + " * ".join("{%d}" % (k + 1) for k in range(len(re) + 1)) + ";",
                          args, False))
            for a in re:
                ops.append(Op("{0} = {0} - {1};", [("F", a), ("RATE", r)], True))
            for k, p in enumerate(sym.products[r]):
                c = sym.sc[r][k]
                if c == 1.0:
                    ops.append(Op("{0} = {0} + {1};", [("F", p), ("RATE", r)], True))
                else:
                    ops.append(Op("{0} = {0} + " + _lit(c) + " * {1};", [("F", p), ("RATE", r)], True))
        return ops
    # gather: all rates, then one species at a time
    for r in day.active:
        re = sym.reactants[r]
        args = [("RATE", r), ("RK", r)] + [("YS", a) for a in re]
        ops.append(Op("{0} = " + " * ".join("{%d}" % (k + 1) for k in range(len(re) + 1)) + ";",
                      args, False))
    contrib = [[] for _ in range(sym.ns)]
    for r in day.active:
        for a in sym.reactants[r]:
            contrib[a].append(("-", r, None))
        for k, p in enumerate(sym.products[r]):
            contrib[p].append(("+", r, sym.sc[r][k]))
    for i in range(sym.ns):
        ops.append(Op("{0} = 0.0;", [("F", i)], False))
        for sign, r, c in contrib[i]:
            if sign == "-":
                ops.append(Op("{0} = {0} - {1};", [("F", i), ("RATE", r)], True))
            elif c == 1.0:
                ops.append(Op("{0} = {0} + {1};", [("F", i), ("RATE", r)], True))
            else:
                ops.append(Op("{0} = {0} + " + _lit(c) + " * {1};", [("F", i), ("RATE", r)], True))
    return ops


def _jac_terms(sym):
    """Per LU slot, the Jacobian contributions in RBJACOB's order:
    (sign, coefficient or None, reaction, reactant position)."""
    day = sym.day
    terms = [[] for _ in range(day.nnz)]
    for r in day.active:
        for slot, ial in day.jac_loss[r]:
            terms[slot].append(("-", None, r, ial))
        for slot, ial, iap in day.jac_prod[r]:
            terms[slot].append(("+", None if iap < 0 else sym.sc[r][iap], r, ial))
    return terms


def _explic(sym, r, ial):
    """Operand slots and expression of RBJACOB's EXPLIC(ial) (rbjacob.F:115-144)."""
    re = sym.reactants[r]
    rk = ("RK", r)
    if len(re) == 1:
        return [rk], "{%d}"
    if len(re) == 2:
        return [rk, ("Y", re[1 - ial])], "({%d} * {%d})"
    if ial == 0:
        return [rk, ("Y", re[1]), ("Y", re[2])], "({%d} * {%d} * {%d})"
    if ial == 1:
        return [rk, ("Y", re[0]), ("Y", re[2])], "({%d} * {%d} * {%d})"
    return [rk, ("Y", re[1]), ("Y", re[0])], "({%d} * {%d} * {%d})"


def _entry_ops(sym, terms, s, diag):
    ops = []
    if not terms[s]:
        ops.append(Op("{0} = 0.0;", [("LU", s)], False))
    else:
        ops.append(Op("{0} = 0.0;", [("LU", s)], False))
        for sign, c, r, ial in terms[s]:
            srcs, ex = _explic(sym, r, ial)
            ex = ex % tuple(range(1, len(srcs) + 1))
            if sign == "-":
                fmt = "{0} = {0} - " + ex + ";"
            elif c is None:
                fmt = "{0} = {0} + " + ex + ";"
            else:
                fmt = "{0} = {0} + " + _lit(c) + " * " + ex + ";"
            ops.append(Op(fmt, [("LU", s)] + srcs, True))
    if s == diag:
        ops.append(Op("{0} = {0} - {1};", [("LU", s), "gdtinv"], True))
    return ops


def _gear_entry_ops(sym, terms, s, diag):
    """One entry of SMVGEAR's predictor matrix P = I + R1DELT * J
    (grpderiv.F:88-224): off-diagonal entries start at 0 and diagonal ones at
    1; each partial derivative EXPLIC is multiplied by R1DELT (or by
    FRACN = SC * R1DELT) as it is added, with the concentrations at or below
    ZBOUND taken as zero (CEFF)."""
    ops = [Op("{0} = 1.0;" if s == diag else "{0} = 0.0;", [("LU", s)], False)]
    for sign, c, r, ial in terms[s]:
        srcs, ex = _explic(sym, r, ial)
        # CEFF(i) = CNEW(i), or 0 at or below ZBOUND: slots of array "CF"
        srcs = [("CF", a[1]) if a[0] == "Y" else a for a in srcs]
        ex = ex % tuple(range(1, len(srcs) + 1))
        if sign == "-":
            fmt = "{0} = {0} - r1delt * " + ex + ";"
        elif c is None:
            fmt = "{0} = {0} + r1delt * " + ex + ";"
        else:
            # FRACN = SC * R1DELT, formed per term (opaque() keeps the compiler
            # from holding every distinct FRACN live across the function)
            fmt = "{0} = {0} + (" + _lit(c) + " * opaque(r1delt)) * " + ex + ";"
        ops.append(Op(fmt, [("LU", s)] + srcs, True))
    return ops


def gear_pderiv_lu_ops(sym, order="column"):
    """PDERIV then DECOMP (grpderiv.F, grdecomp.F) in the given schedule;
    grdecomp.F applies RBDECOMP's operations in the same order."""
    return _jac_lu_ops(sym, order, entry=_gear_entry_ops)


def gear_subfun_ops(sym):
    """SUBFUN (grsubfun.F:83-171) without emissions: RBFEVAL's accumulation
    on CNEW into GLOSS, then a species at or below ZBOUND loses nothing."""
    ops = _remap_arrays(feval_ops(sym, "scatter"), {"YS": "CN", "F": "GL"})
    for i in range(sym.ns):
        ops.append(Op("{0} = ({1} <= ZBOUND && {0} < 0.0) ? 0.0 : {0};", [("GL", i), ("CN", i)], True))
    return ops


def _remap_arrays(ops, names):
    return [Op(op.fmt, [(names.get(a[0], a[0]), a[1]) if isinstance(a, tuple) else a for a in op.args], op.rmw)
            for op in ops]


def _fuse_k1_forward(sym, ops):
    """Interleaves stage 1's forward substitution (rbsolve.F:78-128 with
    K1 = -F, rbsolver.F:386-388) with the factorization: right after L(i,j)
    is scaled, K1(i) -= L(i,j) * K1(j), while L(i,j) is still in a register.
    Every schedule of jac_lu_ops scales a row's L entries in ascending column
    order and only after all of L(j,:) is scaled, so each K1(i) starts from
    -F(i) and takes its terms in ascending j -- the reference's order -- and
    K1(j) is final when it is used."""
    day = sym.day
    started = set()
    out = []

    def start(i):
        if i not in started:
            started.add(i)
            out.append(Op("{0} = -{1};", [("K1", i), ("F", i)], False))

    for op in ops:
        out.append(op)
        if op.fmt == "{0} = {0} * {1};" and op.args[1][0] == "VD" and op.args[0][0] == "LU":
            s_, j = op.args[0][1], op.args[1][1]
            i = day.rows[s_]
            if i > j:
                start(j)
                start(i)
                out.append(Op("{0} = {0} - {1} * {2};", [("K1", i), ("LU", s_), ("K1", j)], True))
    for i in range(sym.ns):      # rows without L entries
        if i not in started:
            start(i)
    return out


def jac_lu_ops(sym, order="column", fuse_k1=False):
    """Jacobian, shift and LU in the given schedule; with fuse_k1 also stage
    1's forward substitution."""
    ops = _jac_lu_ops(sym, order)
    return _fuse_k1_forward(sym, ops) if fuse_k1 else ops


def _jac_lu_ops(sym, order="column", entry=None):
    entry = entry or _entry_ops
    day = sym.day
    n = sym.ns
    terms = _jac_terms(sym)
    ops = []
    if order == "column":
        col_slots = [[] for _ in range(n)]
        for s in range(day.nnz):
            col_slots[day.cols[s]].append(s)
        for j in range(n):
            for s in col_slots[j]:
                ops += entry(sym, terms, s, day.diag[j])
            beg, end = day.lu_col_ops[j]
            for (ij, ik, kj) in day.lu_ops[beg:end]:
                ops.append(Op("{0} = {0} - {1} * {2};", [("LU", ij), ("LU", ik), ("LU", kj)], True))
            ops.append(Op("{0} = 1.0 / {1};", [("VD", j), ("LU", day.diag[j])], False))
            for s in day.l_scale[j]:
                ops.append(Op("{0} = {0} * {1};", [("LU", s), ("VD", j)], True))
        return ops
    if order.startswith("tile") or order.startswith("ctile"):
        # 2-D tiles of br x bc entries, "tile": row block by row block,
        # "ctile": column block by column block (top to bottom within a column
        # block, so the Jacobian terms of a column's reactions and the pivot
        # rows' U entries stay close).  A tile's entries take their updates k
        # by k (each L(i,k) and U(k,j) loaded once per tile), a column's pivot
        # and L scaling come when k reaches it.  Every operand is final when
        # used in either order: L(i,k) lies in an earlier tile of the same row
        # block (or this one), U(k,j) in an earlier row block of the same
        # column block (or this tile).
        colmajor = order.startswith("ctile")
        br, bc = (int(v) for v in order[5 if colmajor else 4:].split("x"))
        upd = {}
        for (ij, ik, kj) in day.lu_ops:
            upd.setdefault(ij, []).append((day.cols[ik], ik, kj))
        tiles = {}
        for s in range(day.nnz):
            tiles.setdefault((day.rows[s] // br, day.cols[s] // bc), []).append(s)
        finished = set()

        def finish(k, in_tile):
            out = []
            if (k, k) in day.slot and day.slot[(k, k)] in in_tile and k not in finished:
                out.append(Op("{0} = 1.0 / {1};", [("VD", k), ("LU", day.diag[k])], False))
                finished.add(k)
            for s in day.l_scale[k]:
                if s in in_tile:
                    out.append(Op("{0} = {0} * {1};", [("LU", s), ("VD", k)], True))
            return out

        for key in sorted(tiles, key=(lambda k: (k[1], k[0])) if colmajor else None):
            ents = sorted(tiles[key], key=lambda s: (day.rows[s], day.cols[s]))
            in_tile = set(ents)
            for s in ents:
                ops += entry(sym, terms, s, day.diag[day.cols[s]])
            cols = sorted({day.cols[s] for s in ents})
            by_k = {}
            for s in ents:
                for (k, ik, kj) in upd.get(s, []):
                    by_k.setdefault(k, []).append((s, ik, kj))
            done_cols = set()
            for k in sorted(by_k):
                for c in cols:
                    if c <= k and c not in done_cols:
                        ops += finish(c, in_tile)
                        done_cols.add(c)
                for (s, ik, kj) in by_k[k]:
                    ops.append(Op("{0} = {0} - {1} * {2};", [("LU", s), ("LU", ik), ("LU", kj)], True))
            for c in cols:
                if c not in done_cols:
                    ops += finish(c, in_tile)
        return ops
    if order.startswith("group"):
        # columns in groups of g: the updates of a group are applied k by k, so
        # each L(i,k) is loaded once per group rather than once per column;
        # column k of the group is finished (pivot, L scaling) when k is reached
        g = int(order[5:] or 2)
        col_slots = [[] for _ in range(n)]
        for s in range(day.nnz):
            col_slots[day.cols[s]].append(s)

        def finish(j):
            out = [Op("{0} = 1.0 / {1};", [("VD", j), ("LU", day.diag[j])], False)]
            for s in day.l_scale[j]:
                out.append(Op("{0} = {0} * {1};", [("LU", s), ("VD", j)], True))
            return out

        for j0 in range(0, n, g):
            grp = list(range(j0, min(n, j0 + g)))
            for j in grp:
                for s in col_slots[j]:
                    ops += entry(sym, terms, s, day.diag[j])
            by_k = {}
            for j in grp:
                beg, end = day.lu_col_ops[j]
                for (ij, ik, kj) in day.lu_ops[beg:end]:
                    by_k.setdefault(day.cols[ik], []).append((j, ij, ik, kj))
            done = set()
            for k in sorted(by_k):
                if k in grp and k not in done:
                    for j in grp:
                        if j < k and j not in done:
                            ops += finish(j)
                            done.add(j)
                    ops += finish(k)
                    done.add(k)
                for (j, ij, ik, kj) in by_k[k]:
                    ops.append(Op("{0} = {0} - {1} * {2};", [("LU", ij), ("LU", ik), ("LU", kj)], True))
            for j in grp:
                if j not in done:
                    ops += finish(j)
        return ops
    # row order: row i's entries, then for k ascending: finish L(i,k), apply its updates
    by_row = [dict() for _ in range(n)]       # i -> k -> [(ij, kj)]
    for (ij, ik, kj) in day.lu_ops:
        i, k = day.rows[ij], day.cols[ik]
        by_row[i].setdefault(k, []).append((ij, ik, kj))
    row_slots = [[] for _ in range(n)]
    for s in range(day.nnz):
        row_slots[day.rows[s]].append(s)
    for i in range(n):
        for s in row_slots[i]:
            ops += entry(sym, terms, s, day.diag[i])
        lks = sorted(day.cols[s] for s in row_slots[i] if day.cols[s] < i)
        for k in lks:
            s_ik = day.slot[(i, k)]
            ops.append(Op("{0} = {0} * {1};", [("LU", s_ik), ("VD", k)], True))
            for (ij, ik, kj) in by_row[i].get(k, []):
                ops.append(Op("{0} = {0} - {1} * {2};", [("LU", ij), ("LU", ik), ("LU", kj)], True))
        ops.append(Op("{0} = 1.0 / {1};", [("VD", i), ("LU", day.diag[i])], False))
    return ops


def solve_ops(sym, stage, backward_only=False):
    """Stage solve: x = K<stage>, right-hand side per rbsolver.F:386-422.
    backward_only: the forward substitution was fused into jac_lu."""
    day = sym.day
    x = f"K{stage}"
    ops = []
    for i in ([] if backward_only else range(sym.ns)):
        if stage == 1:
            ops.append(Op("{0} = -{1};", [(x, i), ("F", i)], False))
        elif stage == 2:
            ops.append(Op("{0} = -{1} - x1 * {2};", [(x, i), ("F", i), ("K1", i)], False))
        else:
            ops.append(Op("{0} = -{1} - x1 * {2} - x2 * {3};", [(x, i), ("F", i), ("K1", i), ("K2", i)], False))
        for (s, c) in day.lower[i]:
            ops.append(Op("{0} = {0} - {1} * {2};", [(x, i), ("LU", s), (x, c)], True))
    for i in range(sym.ns - 1, -1, -1):
        for (s, c) in day.upper[i]:
            ops.append(Op("{0} = {0} - {1} * {2};", [(x, i), ("LU", s), (x, c)], True))
        ops.append(Op("{0} = {0} * {1};", [(x, i), ("VD", i)], True))
    return ops


# ---------------------------------------------------------------------------
# Discrete adjoint of one accepted ROS3 step (no reference counterpart).
#
# Forward step (rbsolver.F:357-473) with M = J(Y) - I/(GAM h):
#   K1 = M^-1 (-F(Y)),  YP = Y + K1,  K2 = M^-1 (-F(YP) - x21 K1),
#   K3 = M^-1 (-F(YP) - x31 K1 - x32 K2),  Ynew = max(Y + B1 K1 + B2 K2 + B3 K3, CONMIN)
# Reverse with LAM = dL/dYnew (step sizes held fixed):
#   LAM *= [Y + B.K > CONMIN];  U3 = M^-T (B3 LAM);  U2 = M^-T (B2 LAM - x32 U3)
#   YB = J(YP)^T (-(U2 + U3));  U1 = M^-T (B1 LAM - x31 U3 - x21 U2 + YB)
#   dL/dY = LAM + YB + J(Y)^T (-U1) - sum_s d/dY [U_s^T J(Y) K_s]
#   dL/dk_r += -s1 p_r(Y) - (s2 + s3) p_r(YP) - sum_s s_s (grad p_r(Y) . K_s)
# where s_s = (S^T U_s)_r (S the stoichiometric matrix) and p_r the product of
# the reaction's reactant concentrations.  Reactions have at most two
# reactants in the supported mechanisms.
# ---------------------------------------------------------------------------

def _check_bimolecular(sym):
    for r in sym.day.active:
        if len(sym.reactants[r]) > 2:
            raise ValueError("adjoint code generation supports at most two reactants per reaction")


def yp_ops(sym):
    """YP = Y + K1, the forward's own expression (rbsolver.F:395-398)."""
    return [Op("{0} = {1} + {2};", [("YP", i), ("Y", i), ("K1", i)], False) for i in range(sym.ns)]


def adj_seed_ops(sym):
    ops = []
    for i in range(sym.ns):
        ops.append(Op("{0} = ((({1} + B1 * {2}) + B2 * {3}) + B3 * {4}) > CONMIN ? {0} : 0.0;",
                      [("LAM", i), ("Y", i), ("K1", i), ("K2", i), ("K3", i)], True))
        ops.append(Op("{0} = B3 * {1};", [("U3", i), ("LAM", i)], False))
    return ops


def tsolve_ops(sym, x):
    """x <- (L U)^T^-1 x with the factors of jac_lu: U^T by rows ascending,
    then L^T by rows descending, both in scatter form."""
    day = sym.day
    ops = []
    for i in range(sym.ns):
        ops.append(Op("{0} = {0} * {1};", [(x, i), ("VD", i)], True))
        for (s, c) in day.upper[i]:
            ops.append(Op("{0} = {0} - {1} * {2};", [(x, c), ("LU", s), (x, i)], True))
    for i in range(sym.ns - 1, -1, -1):
        for (s, c) in day.lower[i]:
            ops.append(Op("{0} = {0} - {1} * {2};", [(x, c), ("LU", s), (x, i)], True))
    return ops


def adj_u2_ops(sym):
    return [Op("{0} = B2 * {1} - x32 * {2};", [("U2", i), ("LAM", i), ("U3", i)], False)
            for i in range(sym.ns)]


def _st_ops(sym, r, dst, src):
    """dst = (S^T src)_r: reactants subtract, products add SC * src."""
    ops = []
    re = sym.reactants[r]
    ops.append(Op("{0} = -{1};", [dst, (src, re[0])], False))
    for a in re[1:]:
        ops.append(Op("{0} = {0} - {1};", [dst, (src, a)], True))
    for k, p in enumerate(sym.products[r]):
        c = sym.sc[r][k]
        if c == 1.0:
            ops.append(Op("{0} = {0} + {1};", [dst, (src, p)], True))
        else:
            ops.append(Op("{0} = {0} + " + _lit(c) + " * {1};", [dst, (src, p)], True))
    return ops


def adj_a_ops(sym):
    """S2, S3 per reaction and YB = J(YP)^T (-(U2 + U3))."""
    _check_bimolecular(sym)
    ops = [Op("{0} = 0.0;", [("YB", i)], False) for i in range(sym.ns)]
    g = ("T", 0)
    for r in sym.day.active:
        re = sym.reactants[r]
        ops += _st_ops(sym, r, ("S2", r), "U2")
        ops += _st_ops(sym, r, ("S3", r), "U3")
        ops.append(Op("{0} = -({1} * ({2} + {3}));", [g, ("RK", r), ("S2", r), ("S3", r)], False))
        if len(re) == 1:
            ops.append(Op("{0} = {0} + {1};", [("YB", re[0]), g], True))
        else:
            a, b = re
            ops.append(Op("{0} = {0} + {1} * {2};", [("YB", a), g, ("YP", b)], True))
            ops.append(Op("{0} = {0} + {1} * {2};", [("YB", b), g, ("YP", a)], True))
    return ops


def adj_u1_ops(sym):
    return [Op("{0} = ((B1 * {1} - x31 * {2}) - x21 * {3}) + {4};",
               [("U1", i), ("LAM", i), ("U3", i), ("U2", i), ("YB", i)], False) for i in range(sym.ns)]


def adj_b_ops(sym):
    """dL/dY into LAM and dL/dk accumulated into GK."""
    _check_bimolecular(sym)
    ops = [Op("{0} = {0} + {1};", [("YB", i), ("LAM", i)], True) for i in range(sym.ns)]
    s1, s23, dk, t = ("T", 1), ("T", 2), ("T", 3), ("T", 4)
    for r in sym.day.active:
        re = sym.reactants[r]
        S2, S3, rk = ("S2", r), ("S3", r), ("RK", r)
        ops += _st_ops(sym, r, s1, "U1")
        ops.append(Op("{0} = {1} + {2};", [s23, S2, S3], False))
        if len(re) == 1:
            a = re[0]
            ops.append(Op("{0} = ((-({1} * {2}) - {3} * {4}) - {1} * {5}) - {6} * {7} - {8} * {9};",
                          [dk, s1, ("Y", a), s23, ("YP", a), ("K1", a), S2, ("K2", a), S3, ("K3", a)], False))
            ops.append(Op("{0} = {0} + {1};", [("GK", r), dk], True))
            ops.append(Op("{0} = {0} - {1} * {2};", [("YB", a), rk, s1], True))
            continue
        a, b = re
        ya, yb_ = ("Y", a), ("Y", b)
        ops.append(Op("{0} = -({1} * ({2} * {3})) - {4} * ({5} * {6});",
                      [dk, s1, ya, yb_, s23, ("YP", a), ("YP", b)], False))
        for sig, K in ((s1, "K1"), (S2, "K2"), (S3, "K3")):
            ops.append(Op("{0} = {0} - {1} * ({2} * {3} + {4} * {5});",
                          [dk, sig, yb_, (K, a), ya, (K, b)], True))
        ops.append(Op("{0} = {0} + {1};", [("GK", r), dk], True))
        # slot a: other reactant b; slot b: other reactant a
        for me, other in ((a, b), (b, a)):
            ops.append(Op("{0} = {1} * {2} + {3} * {4} + {5} * {6};",
                          [t, s1, ("K1", other), S2, ("K2", other), S3, ("K3", other)], False))
            ops.append(Op("{0} = {0} - {1} * ({2} * {3} + {4});",
                          [("YB", me), rk, s1, ("Y", other), t], True))
    ops += [Op("{0} = {1};", [("LAM", i), ("YB", i)], False) for i in range(sym.ns)]
    return ops


# ---------------------------------------------------------------------------
# Arithmetic count of one attempted ROS3 step (the FP64 roofline's numerator)
# ---------------------------------------------------------------------------
_TOK = re.compile(r"\{\d+\}|[A-Za-z_][A-Za-z0-9_]*|\d+\.\d*(?:e[-+]?\d+)?|[-+*/()]")


def op_flops(ops):
    """Binary floating-point operations (+ - * /) in a list of ops; unary
    minus is free (a sign flip folded into the instruction)."""
    n = 0
    for op in ops:
        prev = None
        for t in _TOK.findall(op.fmt.split("=", 1)[1]):
            if t in "*/+" or (t == "-" and prev is not None and prev not in "(*/+-"):
                n += 1
            prev = t
    return n


# ---------------------------------------------------------------------------
# Tangent-linear (forward-mode) sensitivity of one accepted ROS3 step along a
# direction (dY, dk), step size held fixed (no reference counterpart):
#   U1 = -S w1,  w1_r = q_r(Y; dY) + dk_r p_r(Y) + dq_r(Y; K1);  U1 <- M^-1 U1
#   S3 = Y + K1 (stage point), YB = dY + U1,  S2 = S (q_r(S3; YB) + dk_r p_r(S3))
#   U2 = M^-1 (-S2 - x21 U1 - S dq_r(Y; K2))
#   U3 = M^-1 (-S2 - x31 U1 - x32 U2 - S dq_r(Y; K3))
#   dY <- [Ynew > CONMIN] (dY + B1 U1 + B2 U2 + B3 U3)
# with p_r the reactant product, q_r(y; v) = k_r grad p_r(y) . v and
# dq_r(Y; K) = dk_r grad p_r(Y) . K + k_r (d/de grad p_r(Y + e dY)) . K.
# dY lives in LAM and dk in GK (the adjoint's arrays, unused by the forward);
# the first ns entries of S2 / S3 hold dF1 and the stage point.
# ---------------------------------------------------------------------------

def _scatter(sym, r, out, sign):
    """out += sign * (S w)_r with w in RATE(r): -w per reactant, +sc w per product."""
    ops = []
    for a in sym.reactants[r]:
        ops.append(Op("{0} = {0} %s {1};" % ("-" if sign > 0 else "+"), [(out, a), ("RATE", r)], True))
    for k, p in enumerate(sym.products[r]):
        c = sym.sc[r][k]
        op = "+" if sign > 0 else "-"
        if c == 1.0:
            ops.append(Op("{0} = {0} %s {1};" % op, [(out, p), ("RATE", r)], True))
        else:
            ops.append(Op("{0} = {0} %s %s * {1};" % (op, _lit(c)), [(out, p), ("RATE", r)], True))
    return ops


def _tlm_w(sym, r, terms):
    """RATE(r) = sum of the requested per-reaction terms:
    'q'  q_r(YV; DV)     'p'  dk_r p_r(YV)     'dq' dq_r(Y; KV) along (LAM, GK)
    terms: list of (kind, {name: array})."""
    re = sym.reactants[r]
    args = [("RATE", r)]
    index = {}

    def slot(arr, i):
        key = (arr, i)
        if key not in index:
            args.append(key)
            index[key] = len(args) - 1
        return "{%d}" % index[key]

    k, g = (lambda: slot("RK", r)), (lambda: slot("GK", r))
    parts = []
    for kind, arrs in terms:
        if kind == "q":
            y, d = arrs["YV"], arrs["DV"]
            if len(re) == 1:
                parts.append("%s * %s" % (k(), slot(d, re[0])))
            else:
                a, b = re
                parts.append("%s * (%s * %s + %s * %s)" % (k(), slot(d, a), slot(y, b), slot(d, b), slot(y, a)))
        elif kind == "p":
            y = arrs["YV"]
            if len(re) == 1:
                parts.append("%s * %s" % (g(), slot(y, re[0])))
            else:
                parts.append("%s * (%s * %s)" % (g(), slot(y, re[0]), slot(y, re[1])))
        else:  # dq along (LAM, GK) at Y
            kv = arrs["KV"]
            if len(re) == 1:
                parts.append("%s * %s" % (g(), slot(kv, re[0])))
            else:
                a, b = re
                parts.append("%s * (%s * %s + %s * %s)" % (g(), slot(kv, a), slot("Y", b), slot(kv, b), slot("Y", a)))
                parts.append("%s * (%s * %s + %s * %s)" % (k(), slot(kv, a), slot("LAM", b), slot(kv, b),
                                                          slot("LAM", a)))
    return Op("{0} = " + " + ".join(parts) + ";", args, False)


def tlm_rhs1_ops(sym):
    """U1 = -S (q(Y; dY) + dk p(Y) + dq(Y; K1)), before the solve."""
    _check_bimolecular(sym)
    ops = [Op("{0} = 0.0;", [("U1", i)], False) for i in range(sym.ns)]
    for r in sym.day.active:
        ops.append(_tlm_w(sym, r, [("q", {"YV": "Y", "DV": "LAM"}), ("p", {"YV": "Y"}), ("dq", {"KV": "K1"})]))
        ops += _scatter(sym, r, "U1", -1)
    return ops


def tlm_stage_ops(sym):
    """S3 = Y + K1 (the stage point, as yp_ops), YB = dY + U1, then
    S2 = S (q(S3; YB) + dk p(S3))."""
    ops = [Op("{0} = {1} + {2};", [("S3", i), ("Y", i), ("K1", i)], False) for i in range(sym.ns)]
    ops += [Op("{0} = {1} + {2};", [("YB", i), ("LAM", i), ("U1", i)], False) for i in range(sym.ns)]
    ops += [Op("{0} = 0.0;", [("S2", i)], False) for i in range(sym.ns)]
    for r in sym.day.active:
        ops.append(_tlm_w(sym, r, [("q", {"YV": "S3", "DV": "YB"}), ("p", {"YV": "S3"})]))
        ops += _scatter(sym, r, "S2", 1)
    return ops


def tlm_rhs_ops(sym, stage):
    """U2 = -S2 - x1 U1 - S dq(Y; K2), or U3 = -S2 - x1 U1 - x2 U2 - S dq(Y; K3)."""
    _check_bimolecular(sym)
    out, kv = ("U2", "K2") if stage == 2 else ("U3", "K3")
    ops = []
    for i in range(sym.ns):
        if stage == 2:
            ops.append(Op("{0} = -{1} - x1 * {2};", [(out, i), ("S2", i), ("U1", i)], False))
        else:
            ops.append(Op("{0} = -{1} - x1 * {2} - x2 * {3};", [(out, i), ("S2", i), ("U1", i), ("U2", i)], False))
    for r in sym.day.active:
        ops.append(_tlm_w(sym, r, [("dq", {"KV": kv})]))
        ops += _scatter(sym, r, out, -1)
    return ops


def msolve_ops(sym, x):
    """x <- (L U)^-1 x with the factors of jac_lu (RBSOLVE's substitutions,
    rbsolve.F:78-187, on an arbitrary right-hand side)."""
    day = sym.day
    ops = []
    for i in range(sym.ns):
        for (s, c) in day.lower[i]:
            ops.append(Op("{0} = {0} - {1} * {2};", [(x, i), ("LU", s), (x, c)], True))
    for i in range(sym.ns - 1, -1, -1):
        for (s, c) in day.upper[i]:
            ops.append(Op("{0} = {0} - {1} * {2};", [(x, i), ("LU", s), (x, c)], True))
        ops.append(Op("{0} = {0} * {1};", [(x, i), ("VD", i)], True))
    return ops


def tlm_update_ops(sym):
    """dY <- [Ynew > CONMIN] (dY + B1 U1 + B2 U2 + B3 U3); Ynew is in YP."""
    return [Op("{0} = {1} > CONMIN ? (({0} + B1 * {2}) + B2 * {3}) + B3 * {4} : 0.0;",
               [("LAM", i), ("YP", i), ("U1", i), ("U2", i), ("U3", i)], True) for i in range(sym.ns)]


def flops_per_step(sym):
    """FP64 operations of one attempted step: two f evaluations, Jacobian and
    LU, three stage solves, YP = Y + K1, and the new state and error norm of
    rbsolver.F:425-441 (13 per species) plus sqrt and the controller."""
    return (2 * op_flops(feval_ops(sym)) + op_flops(jac_lu_ops(sym)) +
            sum(op_flops(solve_ops(sym, k)) for k in (1, 2, 3)) + sym.ns + 13 * sym.ns + 2)
