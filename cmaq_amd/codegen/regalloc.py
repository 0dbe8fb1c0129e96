"""Register allocation and load scheduling for the generated straight-line code.

Every generated function is first built as a list of operations on named
values.  A value is either a workspace slot -- a per-lane double with a home
in the lane's workspace, written as a tuple such as ("LU", 17) -- or a
uniform scalar parameter, written as a plain string such as "gdtinv".  This
module decides where each slot value lives at each point:

* registers: at most ``budget`` values, so the compiler never spills and the
  kernel keeps its occupancy;
* LDS: an optional second tier of ``lds_slots`` values per lane holding
  copies of values that left the registers but are read again (its copies are
  always clean: a dirty value is written back to the workspace when it leaves
  the registers);
* the workspace in HBM (the home of every slot).

Eviction follows Belady's rule (dead values first, then the farthest next
use; clean values before dirty ones); a dirty value is written back only if
it is read again later or belongs to one of the function's output arrays.
Each operand is fetched ``lookahead`` operations before it is needed, which
keeps several independent requests in flight per wave.

The result is a flat list of items -- ("load", reg, slot), ("store", reg,
slot), ("lds_load", reg, lds), ("lds_store", reg, lds), ("op", text) -- that
emit_hip turns into HIP source.  The arithmetic of every operation is the
caller's, in the caller's order, so the schedule never changes a rounding.
"""

from collections import defaultdict

RESERVE = 6     # registers kept free of prefetches (at least one op's operands + result)


class Op:
    """One statement.  `fmt` names its operands {0}, {1}, ... in `args`;
    args[0] is the destination slot; `rmw` says whether the destination's
    previous value is read."""
    __slots__ = ("fmt", "args", "rmw")

    def __init__(self, fmt, args, rmw):
        self.fmt = fmt
        self.args = args
        self.rmw = rmw


def _reads(op):
    return [a for k, a in enumerate(op.args)
            if isinstance(a, tuple) and (k > 0 or op.rmw)]


def allocate(ops, outputs, budget, lookahead, defined_on_entry, lds_slots=0, dead=frozenset(), inflight=0):
    """Returns (items, counts) with counts of each item kind.

    outputs: array names whose final values must be in memory at the end,
    except for the slots in `dead` (values no later function reads).
    defined_on_entry(slot): True if memory holds the slot's value on entry."""
    R, S = budget, lds_slots
    n = len(ops)
    reserve = max([RESERVE] + [len(op.args) + 1 for op in ops])
    uses = defaultdict(list)
    for t, op in enumerate(ops):
        for s in _reads(op):
            uses[s].append(t)
    cursor = defaultdict(int)

    def next_use(slot, t):
        lst = uses.get(slot, ())
        i = cursor[slot]
        while i < len(lst) and lst[i] < t:
            i += 1
        cursor[slot] = i
        return lst[i] if i < len(lst) else None

    reg_of, slot_of = {}, [None] * R
    dirty, pinned = [False] * R, [-1] * R
    lds_of, lds_slot = {}, [None] * S        # clean LDS copies
    in_mem = set()
    free = list(range(R - 1, -1, -1))
    items = []
    counts = defaultdict(int)

    def emit(kind, a, b):
        items.append((kind, a, b))
        counts[kind] += 1

    def needs_store(r, t):
        s = slot_of[r]
        return dirty[r] and (next_use(s, t) is not None or (s[0] in outputs and s not in dead))

    def lds_put(r, t):
        """Keep a copy of register r's (clean) value in LDS if it is worth it."""
        s = slot_of[r]
        nu = next_use(s, t)
        if S == 0 or nu is None or s in lds_of:
            return
        victim, far = None, nu
        for q in range(S):
            v = lds_slot[q]
            if v is None:
                victim, far = q, None
                break
            vu = next_use(v, t)
            if vu is None:
                victim, far = q, None
                break
            if vu > far:
                victim, far = q, vu
        if victim is None:
            return
        if lds_slot[victim] is not None:
            del lds_of[lds_slot[victim]]
        lds_slot[victim] = s
        lds_of[s] = victim
        emit("lds_store", r, victim)

    def evict(t, protect):
        best, best_key = None, None
        for r in range(R):
            s = slot_of[r]
            if s is None or pinned[r] >= t or s in protect:
                continue
            nu = next_use(s, t)
            st = 1 if needs_store(r, t) else 0
            key = (0, st, 0) if nu is None else (1, -nu, st)
            if best_key is None or key < best_key:
                best, best_key = r, key
        if best is None:
            return None
        s = slot_of[best]
        if needs_store(best, t):
            emit("store", best, s)
            in_mem.add(s)
        lds_put(best, t)
        del reg_of[s]
        slot_of[best] = None
        dirty[best] = False
        return best

    def take(t, protect):
        return free.pop() if free else evict(t, protect)

    def ensure(slot, t, protect, until):
        if slot in reg_of:
            r = reg_of[slot]
            pinned[r] = max(pinned[r], until)
            return True
        if slot not in lds_of and slot not in in_mem:
            if not defined_on_entry(slot):
                raise RuntimeError(f"slot {slot} read before it is defined")
            in_mem.add(slot)
        r = take(t, protect)
        if r is None:
            return False
        if slot in lds_of:
            emit("lds_load", r, lds_of[slot])
        else:
            emit("load", r, slot)
        reg_of[slot] = r
        slot_of[r] = slot
        dirty[r] = False
        pinned[r] = until
        return True

    def invalidate_lds(s):
        q = lds_of.pop(s, None)
        if q is not None:
            lds_slot[q] = None

    pf = 0
    pending = []      # use times of prefetched loads not yet used (inflight > 0)
    for t in range(n):
        # prefetch the operands of the next `lookahead` operations; with
        # inflight > 0, as far ahead as keeps `inflight` prefetched loads
        # outstanding (at most `lookahead` operations)
        if inflight:
            pending = [u for u in pending if u > t]
        while pf < n and pf <= t + lookahead and (not inflight or len(pending) < inflight):
            window, written = set(), set()
            for q in range(t, pf + 1):
                window.update(_reads(ops[q]))
                if q < pf and isinstance(ops[q].args[0], tuple):
                    written.add(ops[q].args[0])
            # operands produced inside the window are computed, not loaded;
            # prefetches never pin more than R - RESERVE registers
            if sum(1 for r in range(R) if pinned[r] > t) + len(_reads(ops[pf])) > R - reserve:
                break
            nload = counts["load"]
            if not all(ensure(s, t, window, pf) for s in _reads(ops[pf]) if s not in written):
                break
            if inflight and pf > t:
                pending += [pf] * (counts["load"] - nload)
            pf += 1
        op = ops[t]
        need = _reads(op)
        protect = set(need)
        for s in need:
            if not ensure(s, t, protect, t):
                raise RuntimeError("register budget too small for one operation")
        dst = op.args[0]
        if isinstance(dst, tuple) and dst not in reg_of:
            r = take(t, protect)
            if r is None:
                raise RuntimeError("register budget too small for one operation")
            reg_of[dst] = r
            slot_of[r] = dst
        names = [f"r{reg_of[a]}" if isinstance(a, tuple) else a for a in op.args]
        items.append(("op", op.fmt.format(*names)))
        if isinstance(dst, tuple):
            r = reg_of[dst]
            dirty[r] = True
            in_mem.discard(dst)
            invalidate_lds(dst)
            pinned[r] = max(pinned[r], t)
    for r in range(R):
        s = slot_of[r]
        if s is not None and dirty[r] and s[0] in outputs and s not in dead:
            emit("store", r, s)
    return items, dict(counts)
