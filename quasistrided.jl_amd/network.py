"""Device-resident tensor networks: `ncon` over this engine.

TensorOperations' `ncon(tensors, network; backend)` reaches the reference
through `tensorcontract!` one pairwise contraction at a time
(test/integrations/test_tensoroperations.jl:132-140), with the permutes and
traces forwarded to StridedNative (src/integrations/tensoroperations.jl:90-107).
Here the whole network runs on the device: partial traces through
`tensortrace`, pairwise contractions through cached device plans, the final
permutation through `tensoradd`, every intermediate a new tensor in HBM.

Network convention (ncon's): `network[i]` labels the axes of `tensors[i]`;
a positive label appears exactly twice (contracted; twice on one tensor is a
partial trace), a negative label once (open).  The output axes are the open
labels in the order -1, -2, ... unless `output` says otherwise; positive
labels are contracted in increasing order unless `order` says otherwise
(`order="optimal"`: the least-multiplication order, `optimal_order`;
`order="greedy"`: `greedy_order`, for networks past the exhaustive search).
When the next label's two tensors are chosen, every label they share is
contracted at once; tensors left with no shared label are combined by outer
products, first two first.
"""
from __future__ import annotations

from collections import Counter

import torch

from .add import tensoradd, tensortrace
from .errors import ArgumentError, DimensionMismatch
from .plan import execute, planned


def ncon_steps(network, order=None, output=None):
    """The host side of `ncon`: validate the network and return
    `(traces, steps, output)`.

    * `traces[i]` = (kept labels, traced label pairs) for a tensor that traces
      a label with itself, else absent.
    * `steps` = list of `(a, b, shared, labels)`: contract nodes `a` and `b`
      over the labels `shared` into a new node (appended, numbered after the
      inputs) whose axes carry `labels`.
    * `output` = the final label order.
    """
    network = [tuple(int(l) for l in ix) for ix in network]
    counts = Counter(l for ix in network for l in ix)
    if any(l == 0 for l in counts):
        raise ArgumentError("ncon: label 0 is not allowed")
    for l, c in counts.items():
        if l > 0 and c != 2:
            raise ArgumentError(f"ncon: contracted label {l} appears {c} times (must be 2)")
        if l < 0 and c != 1:
            raise ArgumentError(f"ncon: open label {l} appears {c} times (must be 1)")
    open_labels = sorted((l for l in counts if l < 0), reverse=True)
    if output is None:
        output = tuple(open_labels)
    else:
        output = tuple(int(l) for l in output)
        if sorted(output) != sorted(open_labels):
            raise ArgumentError(f"ncon: output {output} is not a permutation of the open labels {tuple(open_labels)}")
    positive = sorted(l for l in counts if l > 0)
    if order is None:
        order = positive
    else:
        order = [int(l) for l in order]
        if sorted(order) != positive:
            raise ArgumentError(f"ncon: order {tuple(order)} is not a permutation of the contracted labels")

    traces = {}
    nodes = []
    for i, ix in enumerate(network):
        c = Counter(ix)
        twice = [l for l in ix if c[l] == 2]
        if twice:
            pairs = sorted(set(twice), key=ix.index)
            kept = tuple(l for l in ix if c[l] == 1)
            traces[i] = (kept, tuple(pairs))
            nodes.append(kept)
        else:
            nodes.append(ix)
    done = {l for pair in traces.values() for l in pair[1]}
    alive = list(range(len(nodes)))
    steps = []
    for l in order:
        if l in done:
            continue
        holders = [n for n in alive if l in nodes[n]]
        if len(holders) != 2:   # cannot happen after the count check and the traces
            raise ArgumentError(f"ncon: label {l} is not on two tensors")
        a, b = holders
        shared = tuple(x for x in nodes[a] if x > 0 and x in nodes[b])
        labels = tuple(x for x in nodes[a] if x not in shared) + tuple(x for x in nodes[b] if x not in shared)
        steps.append((a, b, shared, labels))
        nodes.append(labels)
        alive = [n for n in alive if n not in (a, b)] + [len(nodes) - 1]
        done.update(shared)
    while len(alive) > 1:   # no shared labels left: outer products
        a, b = alive[0], alive[1]
        labels = nodes[a] + nodes[b]
        steps.append((a, b, (), labels))
        nodes.append(labels)
        alive = alive[2:] + [len(nodes) - 1]
    return traces, steps, output


def _free_labels(ix):
    """The labels of one tensor after its partial traces."""
    c = Counter(ix)
    return frozenset(l for l in ix if c[l] == 1)


def contraction_cost(network, sizes, order=None):
    """Multiplications of the pairwise sequence `ncon_steps(network, order)`
    runs: each step costs the product of the lengths of every label on its
    two operands (shared ones once).  Partial traces are not counted: every
    order performs the same ones."""
    _, steps, _ = ncon_steps(network, order)
    nodes = [tuple(_free_labels(tuple(int(l) for l in ix))) for ix in network]
    total = 0
    for a, b, _shared, labels in steps:
        cost = 1
        for l in set(nodes[a]) | set(nodes[b]):
            cost *= int(sizes[l])
        total += cost
        nodes.append(labels)
    return total


def optimal_order(network, sizes):
    """A contraction order (a permutation of the positive labels, as `ncon`'s
    `order`) of least `contraction_cost`.

    Dynamic programming over subsets of tensors: the best pairwise tree of a
    subset S is the cheapest split S = S1 + S2 into two parts that share a
    label, each part contracted by its own best tree first; a step costs the
    product of the lengths of the labels on its two operands.  Disconnected
    components are searched apart (`ncon` joins them by outer products at the
    end).  The tree is turned into an order by listing, children first, one
    shared label per step: `ncon` contracts every label two nodes share when
    it reaches one of them.  Exhaustive, so meant for networks of up to
    about 12 tensors (3^n splits)."""
    network = [tuple(int(l) for l in ix) for ix in network]
    ncon_steps(network)   # the same label checks as ncon
    n = len(network)
    if n > 16:
        raise ArgumentError(f"optimal_order: {n} tensors is beyond the exhaustive search (at most 16)")
    free = [_free_labels(ix) for ix in network]
    labels_of = {}

    def labels(mask):   # open labels of the subnetwork `mask`
        if mask not in labels_of:
            c = Counter(l for i in range(n) if mask >> i & 1 for l in free[i])
            labels_of[mask] = frozenset(l for l, k in c.items() if k == 1)
        return labels_of[mask]

    best = {1 << i: (0, None) for i in range(n)}

    def solve(mask):
        if mask in best:
            return best[mask][0]
        low = mask & -mask
        top = None
        sub = (mask - 1) & mask
        while sub:
            if sub & low:   # each split once: the part holding the lowest tensor
                rest = mask ^ sub
                if labels(sub) & labels(rest):
                    step = 1
                    for l in labels(sub) | labels(rest):
                        step *= int(sizes[l])
                    c1 = solve(sub)
                    c2 = solve(rest)
                    if c1 is not None and c2 is not None and (top is None or c1 + c2 + step < top[0]):
                        top = (c1 + c2 + step, (sub, rest))
            sub = (sub - 1) & mask
        best[mask] = top if top is not None else (None, None)
        return best[mask][0]

    # connected components of the tensor graph
    comps, seen = [], 0
    for i in range(n):
        if seen >> i & 1:
            continue
        comp, grow = 1 << i, True
        while grow:
            grow = False
            for j in range(n):
                if not comp >> j & 1 and labels(comp) & free[j]:
                    comp |= 1 << j
                    grow = True
        seen |= comp
        comps.append(comp)

    order = []

    def emit(mask):
        split = best[mask][1]
        if split is None:
            return
        s1, s2 = split
        emit(s1)
        emit(s2)
        order.append(min(labels(s1) & labels(s2)))

    for comp in comps:
        solve(comp)
        emit(comp)
    positive = sorted({l for ix in network for l in ix if l > 0})
    return order + [l for l in positive if l not in order]


def greedy_order(network, sizes):
    """A contraction order (as `ncon`'s `order`) for networks of any size:
    repeatedly contract the pair of connected nodes whose result is smallest
    relative to its operands (result size minus the two operands' sizes,
    ties to the cheaper step, then to the lower node numbers), listing one
    shared label per step.  No optimality claim; `optimal_order` is exact up
    to 16 tensors."""
    network = [tuple(int(l) for l in ix) for ix in network]
    ncon_steps(network)   # the same label checks as ncon
    nodes = {i: _free_labels(ix) for i, ix in enumerate(network)}

    def size(labels):
        out = 1
        for l in labels:
            out *= int(sizes[l])
        return out

    order = []
    nxt = len(nodes)
    while True:
        best = None
        keys = sorted(nodes)
        for x in range(len(keys)):
            for y in range(x + 1, len(keys)):
                a, b = keys[x], keys[y]
                shared = nodes[a] & nodes[b]
                if not shared:
                    continue
                result = nodes[a] ^ nodes[b]
                score = (size(result) - size(nodes[a]) - size(nodes[b]), size(nodes[a] | nodes[b]), a, b)
                if best is None or score < best[0]:
                    best = (score, a, b, shared, result)
        if best is None:
            break
        _, a, b, shared, result = best
        order.append(min(shared))
        del nodes[a], nodes[b]
        nodes[nxt] = result
        nxt += 1
    positive = sorted({l for ix in network for l in ix if l > 0})
    return order + [l for l in positive if l not in order]


def ncon(tensors, network, conjlist=None, order=None, output=None):
    """`ncon` on the device: returns a new tensor with the open labels in
    `output` order (-1, -2, ... by default).  `conjlist[i]`: use conj of
    `tensors[i]`.  `order="optimal"` contracts in the order `optimal_order`
    finds for these tensors' lengths (at most 16 tensors), `order="greedy"`
    in the order `greedy_order` finds (any number)."""
    tensors = list(tensors)
    network = [tuple(ix) for ix in network]
    if len(tensors) != len(network):
        raise ArgumentError(f"ncon: {len(tensors)} tensors but {len(network)} index lists")
    conjlist = [False] * len(tensors) if conjlist is None else [bool(c) for c in conjlist]
    if len(conjlist) != len(tensors):
        raise ArgumentError("ncon: conjlist has the wrong length")
    sizes = {}
    for t, ix in zip(tensors, network):
        if t.dim() != len(ix):
            raise ArgumentError(f"ncon: a tensor of {t.dim()} axes has {len(ix)} labels")
        if t.device.type != "cuda":
            raise ArgumentError("ncon: the device engine takes tensors in GPU memory")
        for n, l in zip(t.shape, ix):
            if sizes.setdefault(int(l), int(n)) != int(n):
                raise DimensionMismatch(f"ncon: label {l} has lengths {sizes[int(l)]} and {int(n)}")
    if isinstance(order, str):
        if order not in ("optimal", "greedy"):
            raise ArgumentError(f"ncon: order {order!r} (a label sequence, \"optimal\" or \"greedy\")")
        order = optimal_order(network, sizes) if order == "optimal" else greedy_order(network, sizes)
    traces, steps, output = ncon_steps(network, order, output)

    nodes = []   # (tensor, labels, conj pending)
    for i, (t, ix) in enumerate(zip(tensors, network)):
        ix = tuple(int(l) for l in ix)
        if i in traces:
            kept, pairs = traces[i]
            out = torch.empty([sizes[l] for l in kept], dtype=t.dtype, device=t.device)
            p = tuple(ix.index(l) + 1 for l in kept)
            q1 = tuple(ix.index(l) + 1 for l in pairs)
            q2 = tuple(len(ix) - ix[::-1].index(l) for l in pairs)
            tensortrace(out, t, (p, ()), (q1, q2), conjlist[i], 1.0, 0.0)
            nodes.append((out, kept, False))
        else:
            nodes.append((t, ix, conjlist[i]))
    for a, b, shared, labels in steps:
        A, la, ca = nodes[a]
        B, lb, cb = nodes[b]
        dt = torch.promote_types(A.dtype, B.dtype)
        out = torch.empty([sizes[l] for l in labels], dtype=dt, device=A.device)
        plan = planned(out, A, la, B, lb, labels, conjA=ca, conjB=cb, cache=True)
        try:
            execute(plan, 1.0, 0.0)
        finally:
            plan.release()
        nodes.append((out, labels, False))
    X, lx, cx = nodes[-1]
    if lx == output and not cx:
        return X
    res = torch.empty([sizes[l] for l in output], dtype=X.dtype, device=X.device)
    perm = tuple(lx.index(l) + 1 for l in output)
    tensoradd(res, X, (perm, ()), cx, 1.0, 0.0)
    return res
