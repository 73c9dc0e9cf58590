"""Label planning: the host side of the path, restated from
src/planning/labels.jl and src/layout/axis_group.jl.

Labels are classified into M (in A and C), N (in B and C) and K (in A and B);
free labels are ordered by |stride| in C, contracted labels by the reference's
two-operand pack cost model; the M/N roles swap for a real compute type when
only N gives C a unit-stride run long enough for the register tile.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

from . import _lib
from .errors import ArgumentError, BoundsError, DimensionMismatch, raise_status

INT64_MAX = (1 << 63) - 1
INT64_MIN = -(1 << 63)

# Contracted-label order constants (src/hardware/target.jl:205-208).
K_WALK_FAR_BYTES = 4096
K_WALK_FAR_PENALTY = 3
# The L2 share the cost model assumes when none is given: the reference's own
# value for an undetected cache (target.jl l2_core_bytes), and a 64-byte line.
DEFAULT_L2_BYTES = 1 << 20
DEFAULT_LINE_BYTES = 64


def _checked(x: int) -> int:
    if x > INT64_MAX or x < INT64_MIN:
        raise OverflowError("AxisGroup: Int64 overflow")
    return x


@dataclass(frozen=True)
class AxisGroup:
    """AxisGroup{D,P} (axis_group.jl:5-41): `lengths` of D dims, first fastest,
    and P stride maps.  Construction runs the reference's validation."""
    lengths: tuple
    strides: tuple  # P tuples of D ints

    def __post_init__(self):
        lengths = tuple(int(x) for x in self.lengths)
        strides = tuple(tuple(int(x) for x in s) for s in self.strides)
        object.__setattr__(self, "lengths", lengths)
        object.__setattr__(self, "strides", strides)
        P = len(strides)
        if P < 1:
            raise ArgumentError(f"AxisGroup requires at least one map (P >= 1), got P = {P}")
        for d, L in enumerate(lengths):
            if L < 0:
                raise ArgumentError(f"AxisGroup lengths must be nonnegative, got lengths[{d + 1}] = {L}")
        for s in strides:
            if len(s) != len(lengths):
                raise ArgumentError("AxisGroup: every map needs one stride per dim")
        # validate_axis_group_bounds (axis_group.jl:27-41)
        if any(L == 0 for L in lengths):
            return
        card = 1
        for L in lengths:
            card = _checked(card * L)
        for S in strides:
            acc = 0
            for d, L in enumerate(lengths):
                if L <= 1:
                    continue
                if S[d] == INT64_MIN:
                    raise OverflowError("checked_abs overflow")
                acc = _checked(acc + _checked((L - 1) * abs(S[d])))

    @property
    def ndim(self) -> int:
        return len(self.lengths)

    def axis_length(self) -> int:
        q = 1
        for L in self.lengths:
            q *= L
        return q

    def offsets(self, q: int):
        """Offsets of coordinate `q` in every map, by direct decoding."""
        Q = self.axis_length()
        if not (0 <= q < Q):
            raise BoundsError(f"coordinate {q} out of range [0, {Q})")
        out = [0] * len(self.strides)
        for d, L in enumerate(self.lengths):
            x = q % L
            q //= L
            for p, S in enumerate(self.strides):
                out[p] += x * S[d]
        return tuple(out)

    def fill_offsets(self, first: int, count: int):
        """fill_offsets! through the library (two-map groups)."""
        if len(self.strides) != 2:
            raise ArgumentError("fill_offsets: the library enumerates two-map groups")
        g = _lib.make_group(self.lengths, self.strides[0], self.strides[1])
        n = max(count, 0)
        o0 = (C.c_int64 * max(n, 1))()
        o1 = (C.c_int64 * max(n, 1))()
        st = _lib.lib().qs_fill_offsets(C.byref(g), first, count, o0, o1)
        if st != 0:
            raise_status(st, _lib.last_error())
        return list(o0[:n]), list(o1[:n])


def axis_length(g: AxisGroup) -> int:
    return g.axis_length()


def map_ramp_step(g: AxisGroup, p: int):
    """The step of map `p` if it is a single ramp (axis_group.jl:149-168), else None."""
    step, run, started = 0, 1, False
    for d, L in enumerate(g.lengths):
        if L == 0:
            return 0
        if L == 1:
            continue
        S = g.strides[p][d]
        if not started:
            step, run, started = S, L, True
        else:
            if run * step != S:
                return None
            run *= L
    return step


def affine_ramp(g: AxisGroup):
    """(is every map a ramp, steps) -- axis_group.jl:172-180."""
    steps = []
    for p in range(len(g.strides)):
        s = map_ramp_step(g, p)
        if s is None:
            return False, tuple(0 for _ in g.strides)
        steps.append(s)
    return True, tuple(steps)


def group_from_labels(labels, op1, op2) -> AxisGroup:
    """AxisGroup(labels, (ind1, v1), (ind2, v2)) (axis_group.jl:50-69)."""
    ind1, v1 = op1
    ind2, v2 = op2
    lens, s1, s2 = [], [], []
    for lbl in labels:
        p1 = ind1.index(lbl)
        p2 = ind2.index(lbl)
        l1, l2 = v1.size[p1], v2.size[p2]
        if l1 != l2:
            raise DimensionMismatch(f"label {lbl} has mismatched axis length: {l1} vs {l2}")
        lens.append(l1)
        s1.append(v1.strides[p1])
        s2.append(v2.strides[p2])
    return AxisGroup(tuple(lens), (tuple(s1), tuple(s2)))


# --------------------------------------------------------------- labels

def group_ranks(NA: int, NB: int, NC: int):
    """(|M|, |N|, |K|) from the tuple lengths (labels.jl:11-12)."""
    return ((NA + NC - NB) // 2, (NB + NC - NA) // 2, (NA + NB - NC) // 2)


def _allunique(t) -> bool:
    return len(set(t)) == len(t)


def classify_labels(indA, indB, indC):
    """(mlabels, nlabels, klabels) in indA/indB order (labels.jl:28-102)."""
    indA, indB, indC = tuple(indA), tuple(indB), tuple(indC)
    if not _allunique(indA):
        raise ArgumentError(f"indA has a repeated label (diagonal), not supported: {indA}")
    if not _allunique(indB):
        raise ArgumentError(f"indB has a repeated label (diagonal), not supported: {indB}")
    if not _allunique(indC):
        raise ArgumentError(f"indC has a repeated label (diagonal), not supported: {indC}")
    m, n, k = [], [], []
    for lbl in indA:
        inB, inC = lbl in indB, lbl in indC
        if inB and inC:
            raise ArgumentError(f"label {lbl} appears in indA, indB, and indC: labels present in all three "
                                "operands (batch-like) are not supported")
        elif inB:
            k.append(lbl)
        elif inC:
            m.append(lbl)
        else:
            raise ArgumentError(f"label {lbl} appears only in indA (not in indB or indC): not a valid free (M) "
                                "or contracted (K) label")
    for lbl in indB:
        if lbl in indA:
            continue
        if lbl in indC:
            n.append(lbl)
        else:
            raise ArgumentError(f"label {lbl} appears only in indB (not in indA or indC): not a valid free (N) "
                                "or contracted (K) label")
    for lbl in indC:
        if lbl not in indA and lbl not in indB:
            raise ArgumentError(f"label {lbl} appears in indC but not in indA or indB")
    return tuple(m), tuple(n), tuple(k)


def order_free_labels(labels, indC, C):
    """`labels` stably sorted by |stride| of their axis in C (labels.jl:109-115)."""
    st = C.strides
    key = {l: abs(st[list(indC).index(l)]) for l in labels}
    return tuple(sorted(labels, key=lambda l: key[l]))


@dataclass
class KOperand:
    len: tuple
    st: tuple
    n: int
    S: int
    wholeline: bool


def k_operand(klabels, ind, v, free, line) -> KOperand:
    """What the K-order model reads of one operand (labels.jl:149-169)."""
    ind = list(ind)
    st, sz = v.strides, v.size
    S = v.parent.element_size()
    ps = [ind.index(l) for l in klabels]
    length = tuple(sz[p] for p in ps)
    kst = tuple(abs(st[p]) for p in ps)
    line_elems = max(1, line // S)
    wholeline = False
    for l in free:
        p = ind.index(l)
        if sz[p] == 1:
            continue
        wholeline = st[p] == 1 and sz[p] >= line_elems
        break
    n = 1
    for s in sz:
        n *= s
    return KOperand(length, kst, n, S, wholeline)


def k_order_cost(perm, op: KOperand, Qfree: int, l2bytes: int, line: int) -> int:
    """Cost of packing one operand under K order `perm` (labels.jl:172-211)."""
    length, st, S = op.len, op.st, op.S
    kfast = None
    for i in perm:
        if length[i] > 1:
            kfast = i
            break
    if kfast is None:
        return 0
    walk = K_WALK_FAR_PENALTY if st[kfast] * S > K_WALK_FAR_BYTES else 1
    if op.wholeline:
        return op.n * walk
    u, su = None, None
    for i in perm:
        if length[i] == 1:
            continue
        if su is None or st[i] < su:
            su, u = st[i], i
    suS = su * S
    if 2 * suS > line or u == kfast:
        return op.n * walk
    faster = 1
    for i in perm:
        if i == u:
            break
        faster *= length[i]
    if faster * Qfree * line <= l2bytes:
        return op.n * walk
    share = min(length[u], line // suS)
    return op.n * walk * share


def _sortperm(keys):
    return tuple(sorted(range(len(keys)), key=lambda i: keys[i]))


def order_contract_labels(klabels, indA, A, morder, indB, B, norder, m_length, n_length,
                          l2bytes=None, line=DEFAULT_LINE_BYTES):
    """The contracted-label order among indA order, A-sorted and B-sorted, by the
    pack cost model; ties keep indA order (labels.jl:215-243)."""
    klabels = tuple(klabels)
    DK = len(klabels)
    if DK <= 1:
        return klabels
    opA = k_operand(klabels, indA, A, morder, line)
    opB = k_operand(klabels, indB, B, norder, line)
    ident = tuple(range(DK))
    permA, permB = _sortperm(opA.st), _sortperm(opB.st)
    if permA == ident and permB == ident:
        return klabels
    l2 = DEFAULT_L2_BYTES if l2bytes is None else l2bytes

    def cost(perm):
        return k_order_cost(perm, opA, m_length, l2, line) + k_order_cost(perm, opB, n_length, l2, line)

    best, bestcost = ident, cost(ident)
    for cand in (permA, permB):
        if cand == best:
            continue
        c = cost(cand)
        if c < bestcost:
            best, bestcost = cand, c
    return tuple(klabels[i] for i in best)


def leading_unit_run(labels, indC, C) -> int:
    """Length of the leading unit-stride run of `labels` in C (labels.jl:249-263)."""
    st = C.strides
    run = 1
    for l in labels:
        p = list(indC).index(l)
        L = C.size[p]
        if L == 1:
            continue
        if L == 0:
            return 0
        if st[p] != run:
            break
        run *= L
    return run


def prefer_swap(run_m: int, run_n: int, m_tile_asis: int, m_tile_swapped: int = None) -> bool:
    """Swap when the as-is M run cannot fill a register sliver and the swapped
    one can (labels.jl:269-270)."""
    if m_tile_swapped is None:
        m_tile_swapped = m_tile_asis
    return run_m < m_tile_asis and run_n >= m_tile_swapped
