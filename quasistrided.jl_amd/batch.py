"""Batch labels on the device: a label that appears in A, B and C runs the
contraction once per batch coordinate, all of them in one launch
(`qs_group_create_batched`).

`plan_contract` keeps the reference's rejection of such labels
(src/planning/labels.jl:48-54); this module is the extension.  The batch
axes are dropped, the remaining contraction is planned exactly as
`plan_contract` plans it (same validation, groups, K order and M/N swap), and
every batch coordinate becomes a group member that runs that plan with its A,
B and C bases moved by the coordinate's offsets.  With one alpha and beta for
all coordinates the members run as one batched launch of the plan's tiled
kernel (one grid plane per coordinate); with per-coordinate coefficients, as
one grouped persistent launch.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .errors import ArgumentError, DimensionMismatch, raise_status
from .group import Graph, _coef
from .plan import PathModes, planned
from .strided import StridedView


def _check(st):
    if st != 0:
        raise_status(st, _lib.last_error())


def _disjoint(sizes, strides) -> bool:
    """A sufficient test that a strided index map is one-to-one: sorted by
    |stride|, each stride clears the span of all smaller ones."""
    dims = sorted((abs(s), n) for n, s in zip(sizes, strides) if n > 1)
    span = 0
    for s, n in dims:
        if s <= span:
            return False
        span += s * (n - 1)
    return True


def _drop(v: StridedView, ind, batch):
    keep = [i for i, lbl in enumerate(ind) if lbl not in batch]
    sub = StridedView(v.parent, tuple(v.size[i] for i in keep), tuple(v.strides[i] for i in keep), v.offset, v.op)
    return sub, tuple(ind[i] for i in keep)


class BatchedPlan:
    """One inner plan run for every batch coordinate: one batched launch of
    the plan's kernel when all coordinates share alpha and beta, else one
    grouped launch."""

    def __init__(self, plan, handle, n, batch, extents, offsets, tile):
        self.plan = plan            # the inner ContractPlan (batch coordinate 0)
        self.batch = batch          # batch labels, first one fastest
        self.extents = extents
        self.n = n
        self.tile = tile            # workgroup tile of the batched launch (-1: the plan's)
        self._offsets = offsets
        self._h = handle
        f = C.c_int64()
        _check(_lib.lib().qs_group_flops(handle, C.byref(f)))
        self.flops = f.value

    def _rebuild(self, tile):
        h = _create(self.plan, self.n, self._offsets, tile)
        _lib.lib().qs_group_destroy(self._h)
        self._h, self.tile = h, tile

    def autotune(self, reps=5, tiles=(-1, 0, 1, 2, 3, 4, 5)):
        """Time the batched launch with each workgroup tile (C is saved and
        restored) and keep the fastest; returns {tile: seconds}.  Each tile is
        timed as a captured graph of `reps` launches, so the per-call host
        work of building the coefficient arrays stays out of the figure."""
        import torch
        saved = self.plan.Cstorage.clone()
        times = {}
        for t in tiles:
            self._rebuild(t)
            g = self.capture(repeats=reps, alpha=1.0, beta=0.0)
            g.launch()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            g.launch()
            e1.record()
            torch.cuda.synchronize()
            g.release()
            times[t] = e0.elapsed_time(e1) * 1e-3 / reps
        self._rebuild(min(times, key=times.get))
        self.plan.Cstorage.copy_(saved)
        torch.cuda.synchronize()
        return times

    def execute(self, alpha=1.0, beta=0.0, stream=None):
        """C[.., b, ..] = alpha A[.., b, ..] B[.., b, ..] + beta C[.., b, ..] for every batch coordinate b."""
        import torch
        s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
        _check(_lib.lib().qs_group_execute(self._h, _coef(alpha, self.n), _coef(beta, self.n), C.c_void_p(s)))
        return self.plan.Cstorage

    def capture(self, repeats=1, alpha=1.0, beta=0.0, stream=None) -> Graph:
        import torch
        s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
        g = C.c_void_p()
        _check(_lib.lib().qs_group_capture(self._h, _coef(alpha, self.n), _coef(beta, self.n), int(repeats),
                                           C.c_void_p(s), C.byref(g)))
        return Graph(g, self)

    def release(self):
        if self._h is not None:
            _lib.lib().qs_group_destroy(self._h)
            self._h = None
        if self.plan is not None:
            self.plan.release()
            self.plan = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def batch_labels(indA, indB, indC):
    """Labels present in all three operands, in indC order."""
    return tuple(lbl for lbl in indC if lbl in indA and lbl in indB)


def plan_batched(C_, A_, indA, B_, indB, indC, *, tile=-1, **kw) -> BatchedPlan:
    """Plan `C[indC] = A[indA] * B[indB]` where some labels appear in all
    three (batch labels).  Device memory only; keyword arguments as for
    `plan_contract` (plain precision).  Each batch coordinate's C slice must be
    disjoint from the others'."""
    Cv, Av, Bv = StridedView.of(C_), StridedView.of(A_), StridedView.of(B_)
    indA, indB, indC = tuple(indA), tuple(indB), tuple(indC)
    for name, ind, v in (("indA", indA, Av), ("indB", indB, Bv), ("indC", indC, Cv)):
        if len(ind) != v.ndim:
            raise ArgumentError(f"{name} has {len(ind)} labels for an operand of {v.ndim} axes")
        if len(set(ind)) != len(ind):
            raise ArgumentError(f"{name} has a repeated label (diagonal), not supported: {ind}")
    batch = batch_labels(indA, indB, indC)
    if not batch:
        raise ArgumentError("plan_batched: no label appears in indA, indB and indC (use plan_contract)")
    if kw.get("precision", "plain") != "plain":
        raise ArgumentError("plan_batched: batched contractions run in plain precision")
    ext, sa, sb, sc = [], [], [], []
    for lbl in batch:
        ia, ib, ic = indA.index(lbl), indB.index(lbl), indC.index(lbl)
        n = Cv.size[ic]
        if Av.size[ia] != n or Bv.size[ib] != n:
            raise DimensionMismatch(f"batch label {lbl}: lengths {Av.size[ia]} (A), {Bv.size[ib]} (B), {n} (C)")
        ext.append(n)
        sa.append(Av.strides[ia])
        sb.append(Bv.strides[ib])
        sc.append(Cv.strides[ic])
    if not _disjoint(Cv.size, Cv.strides):
        raise ArgumentError("plan_batched: C's elements (batch slices included) may overlap")
    Av0, iA = _drop(Av, indA, batch)
    Bv0, iB = _drop(Bv, indB, batch)
    Cv0, iC = _drop(Cv, indC, batch)
    kw.setdefault("path_modes", PathModes("never", "never"))   # the grouped tiled GEMM path
    plan = planned(Cv0, Av0, iA, Bv0, iB, iC, **kw)
    n = int(np.prod(ext)) if ext else 1
    coords = np.array(np.unravel_index(np.arange(n), ext, order="F"), dtype=np.int64).reshape(len(ext), n)
    off = lambda s: (coords * np.asarray(s, dtype=np.int64)[:, None]).sum(axis=0)
    aoff, boff, coff = off(sa), off(sb), off(sc)
    if plan.swapped:
        aoff, boff = boff, aoff
    offsets = (aoff, boff, coff)
    try:
        h = _create(plan, n, offsets, int(tile))
    except Exception:
        plan.release()
        raise
    return BatchedPlan(plan, h, n, batch, tuple(ext), offsets, int(tile))


def _create(plan, n, offsets, tile):
    arr = lambda x: (C.c_int64 * max(n, 1))(*[int(v) for v in x])
    h = C.c_void_p()
    aoff, boff, coff = offsets
    _check(_lib.lib().qs_group_create_batched(plan.handle, n, arr(aoff), arr(boff), arr(coff), int(tile), C.byref(h)))
    return h


def contract_batched(C_, alpha, A_, indA, B_, indB, beta, indC, **kw):
    """Plan, run and release a batched contraction; returns C."""
    bp = plan_batched(C_, A_, indA, B_, indB, indC, **kw)
    try:
        bp.execute(alpha, beta)
    finally:
        bp.release()
    return C_
