"""Symbolic analysis of a mechanism for the generated ROS3 kernels.

Re-derives, from the RXNS_DATA tables, everything CMAQ's RBSPARSE
(CCTM/src/gas/ros3/rbsparse.F:26-696) fixes once per run:

* the solver species order (rbsparse.F:209-252): species that appear in no
  partial derivative go last; the rest are selection-sorted by their number
  of Jacobian row entries, fewest first;
* the day / night active-reaction lists (rbsparse.F:259-281; night drops
  reactions with bit 1 of IRXBITS, the photolysis flag);
* the symbolic LU fill (rbsparse.F:423-455, column-by-column Doolittle
  without pivoting) and the storage order of the factored matrix
  (rbsparse.F:461-497: L rows 2..N, then U rows N-1..1, then the diagonal);
* the operation order of the numeric factorisation (rbdecomp.F:86-136) and
  of the triangular solves (rbsolve.F:78-187), whose term-by-term order
  fixes the floating-point rounding;
* the Jacobian scatter order (rbjacob.F:112-183, pointers from
  rbsparse.F:626-655) and the f(y) accumulation order (rbfeval.F:93-149).

Indices are 0-based here.  The generated kernels reproduce the reference's
rounding exactly because every per-entry sequence of operations below is
kept in the reference order.
"""

from dataclasses import dataclass, field
from typing import List, Tuple


@dataclass
class Pattern:
    """Sparse structure of one of the two (day/night) matrices."""
    active: List[int]                       # active reactions, reference order
    nnz: int
    slot: dict                              # (row, col) -> storage slot
    rows: List[int]                         # slot -> row
    cols: List[int]                         # slot -> col
    lu_ops: List[Tuple[int, int, int]]      # (ij, ik, kj) slots, execution order
    lu_col_ops: List[Tuple[int, int]]       # per column: [beg, end) into lu_ops
    l_scale: List[List[int]]                # per column: L slots scaled by 1/U_jj
    diag: List[int]                         # per species: diagonal slot
    lower: List[List[Tuple[int, int]]]      # row i: [(slot, col)] forward solve
    upper: List[List[Tuple[int, int]]]      # row i: [(slot, col)] backward solve
    # Jacobian scatter, per active reaction: loss [(slot, reactant_pos)],
    # production [(slot, reactant_pos, product_index or -1 for unit coeff)]
    jac_loss: dict = field(default_factory=dict)
    jac_prod: dict = field(default_factory=dict)


@dataclass
class Symbolic:
    ns: int
    nr: int
    ischan: int
    inew2old: List[int]       # new -> old (0-based)
    iold2new: List[int]
    reactants: List[List[int]]  # per reaction, new-order reactant ids
    products: List[List[int]]   # per reaction, new-order product ids
    sc: List[List[float]]       # per reaction, product coefficients
    day: Pattern
    night: Pattern


def analyse(m) -> Symbolic:
    ns, nr, mxprd = m["numb_mech_spc"], m["nrxns"], m["mxprd"]
    irr, nreact, nprdct = m["irr"], m["nreact"], m["nprdct"]

    # rows of the Jacobian pattern touched by each species (rbsparse.F:185-202)
    touched = [set() for _ in range(ns)]
    for r in range(nr):
        row = irr[r]
        for a in row[:3]:
            if a:
                for b in row[:3 + mxprd]:
                    if b:
                        touched[b - 1].add(a - 1)
    npd = [len(s) for s in touched]

    # species order (rbsparse.F:209-252)
    inew2old = [0] * ns
    iold2new = [0] * ns
    top, bottom = 0, ns - 1
    for old in range(ns):
        if npd[old] > 0:
            inew2old[top], iold2new[old] = old, top
            top += 1
        else:
            inew2old[bottom], iold2new[old] = old, bottom
            bottom -= 1
    ischan = top
    for j in range(ischan):
        best_new, best_val = j, npd[inew2old[j]]
        for i in range(j + 1, ischan):
            if npd[inew2old[i]] < best_val:
                best_new, best_val = i, npd[inew2old[i]]
        a_old, b_old = inew2old[j], inew2old[best_new]
        inew2old[best_new], inew2old[j] = a_old, b_old
        iold2new[a_old], iold2new[b_old] = best_new, j

    reactants = [[iold2new[irr[r][k] - 1] for k in range(nreact[r])] for r in range(nr)]
    products = [[iold2new[irr[r][3 + k] - 1] for k in range(nprdct[r])] for r in range(nr)]
    sc = [list(m["sc"][r][:nprdct[r]]) for r in range(nr)]

    day = [r for r in range(nr) if nreact[r] > 0]
    night = [r for r in day if not (m["irxbits"][r] & 2)]

    def pattern(active):
        lz = [[False] * ns for _ in range(ns)]
        for i in range(ns):
            lz[i][i] = True
        for r in active:
            for col in reactants[r]:
                for row in reactants[r]:
                    lz[row][col] = True
                for row in products[r]:
                    lz[row][col] = True
        ops_ikj, lsc = [], []
        for j in range(ischan):
            col_ops = []
            for i in range(1, ischan):
                for k in range(min(i, j)):
                    if lz[i][k] and lz[k][j]:
                        col_ops.append((i, k))
                        lz[i][j] = True
            ops_ikj.append(col_ops)
            lsc.append([i for i in range(j + 1, ischan) if lz[i][j]])
        slot, rows, cols = {}, [], []

        def put(i, j):
            slot[(i, j)] = len(rows)
            rows.append(i)
            cols.append(j)
        for i in range(1, ischan):
            for j in range(i):
                if lz[i][j]:
                    put(i, j)
        for i in range(ischan - 2, -1, -1):
            for j in range(i + 1, ischan):
                if lz[i][j]:
                    put(i, j)
        for i in range(ischan):
            put(i, i)
        lu_ops, col_rng = [], []
        for j in range(ischan):
            beg = len(lu_ops)
            for (i, k) in ops_ikj[j]:
                lu_ops.append((slot[(i, j)], slot[(i, k)], slot[(k, j)]))
            col_rng.append((beg, len(lu_ops)))
        lower = [[(slot[(i, j)], j) for j in range(i) if (i, j) in slot] for i in range(ischan)]
        upper = [[(slot[(i, j)], j) for j in range(i + 1, ischan) if (i, j) in slot] for i in range(ischan)]
        p = Pattern(active=list(active), nnz=len(rows), slot=slot, rows=rows, cols=cols,
                    lu_ops=lu_ops, lu_col_ops=col_rng,
                    l_scale=[[slot[(i, j)] for i in lsc[j]] for j in range(ischan)],
                    diag=[slot[(i, i)] for i in range(ischan)], lower=lower, upper=upper)
        for r in active:
            loss, prod = [], []
            for ial, ir in enumerate(reactants[r]):
                for jr in reactants[r]:
                    loss.append((slot[(jr, ir)], ial))
                for iap, jp in enumerate(products[r]):
                    unit = abs(sc[r][iap] - 1.0) <= 1.0e-06
                    prod.append((slot[(jp, ir)], ial, -1 if unit else iap))
            p.jac_loss[r] = loss
            p.jac_prod[r] = prod
        return p

    return Symbolic(ns=ns, nr=nr, ischan=ischan, inew2old=inew2old, iold2new=iold2new,
                    reactants=reactants, products=products, sc=sc,
                    day=pattern(day), night=pattern(night))
