"""Permuted add and partial trace on the device: TensorOperations'
`tensoradd!` / `tensortrace!`, which the reference forwards to
`StridedNative` (src/integrations/tensoroperations.jl:90-107).

    tensoradd(C, A, pA, conjA, alpha, beta):
        C = beta * C + alpha * permutedims(op(A), linearize(pA))
    tensortrace(C, A, p, q, conjA, alpha, beta):
        C = beta * C + alpha * (op(A) traced over the axis pairs (q[0][i], q[1][i]),
                                permuted by linearize(p))

Axis numbers are one-based, as in Julia.  C's axes become two groups for the
library: M starts with C's unit-stride axis, N with A's, so a transposing
permutation is read and written coalesced; the traced pairs form K, each with
the sum of its two strides in A.
"""
from __future__ import annotations

import ctypes as C
from collections import OrderedDict

import torch

from . import _lib
from .errors import ArgumentError, DimensionMismatch, UnsupportedError, raise_status
from .plan import _scalar_buffer, compute_type
from .strided import DTYPE_CODE, StridedView


def _linearize(p):
    return tuple(p[0]) + tuple(p[1])


def _check_perm(seq, n, what):
    if sorted(seq) != list(range(1, n + 1)):
        raise ArgumentError(f"invalid permutation of {what}: {seq}")


def _overlap(Cv: StridedView, Av: StridedView) -> bool:
    return Cv.parent.untyped_storage().data_ptr() == Av.parent.untyped_storage().data_ptr()


def _group_split(dims, lead, limit=64):
    """Move dims[lead] (if any) first, then take dims in order until the
    group reaches `limit` elements; returns (taken, rest)."""
    order = list(range(len(dims)))
    if lead is not None:
        order.remove(lead)
        order.insert(0, lead)
    taken, n = [], 1
    for i in order:
        if n >= limit and taken:
            break
        taken.append(i)
        n *= dims[i][0]
    return taken, [i for i in order if i not in taken]


def _run(Cv, Av, amap, pairs, conjA, alpha, beta):
    """C axis d <- A axis amap[d] (zero-based); pairs: traced (q1, q2)."""
    if Cv.op in ("conj", "adjoint"):
        raise ArgumentError("tensoradd!: a conjugated output view is not supported")
    if Av.dtype.is_complex and not Cv.dtype.is_complex:
        raise ArgumentError("tensoradd!: a complex A needs a complex C")
    if _overlap(Cv, Av):
        raise ArgumentError("tensoradd!: C must not alias A")
    if Cv.parent.device.type != "cuda" or Av.parent.device.type != "cuda":
        raise UnsupportedError("tensoradd!: the device engine runs on operands resident in GPU memory")
    T = compute_type(Av.dtype, Av.dtype, Cv.dtype, None)
    conj = (Av.op in ("conj", "adjoint")) != bool(conjA)

    dims = [(Cv.size[d], Av.strides[amap[d]], Cv.strides[d]) for d in range(len(amap))]
    cu = next((d for d, x in enumerate(dims) if x[0] > 1 and abs(x[2]) == 1), None)
    au = next((d for d, x in enumerate(dims) if x[0] > 1 and abs(x[1]) == 1 and d != cu), None)
    # M: C's unit-stride axis first; A's unit-stride axis is kept for N.
    pool = [d for d in range(len(dims)) if d != au]
    mi, _ = _group_split([dims[d] for d in pool], pool.index(cu) if cu in pool else None)
    mdims = [pool[i] for i in mi]
    ndims_ = [d for d in range(len(dims)) if d not in mdims]
    if au is not None and au in ndims_:
        ndims_.remove(au)
        ndims_.insert(0, au)
    grp = lambda idx: _lib.make_group([dims[d][0] for d in idx], [dims[d][1] for d in idx],
                                      [dims[d][2] for d in idx])
    kg = _lib.make_group([Av.size[q1] for q1, _ in pairs], [Av.strides[q1] + Av.strides[q2] for q1, q2 in pairs],
                         [0] * len(pairs))
    d = _lib.QSAddDesc()
    d.a_type, d.c_type, d.compute_type = DTYPE_CODE[Av.dtype], DTYPE_CODE[Cv.dtype], DTYPE_CODE[T]
    d.conj_a = 1 if (conj and T.is_complex) else 0
    d.m, d.n, d.k = grp(mdims), grp(ndims_), kg
    d.a, d.a_base, d.a_len = Av.parent.data_ptr(), Av.offset, Av.parent.numel()
    d.c, d.c_base, d.c_len = Cv.parent.data_ptr(), Cv.offset, Cv.parent.numel()
    lib = _lib.lib()
    handle = _cached_plan(d)
    st = lib.qs_add_plan_rebind(handle, C.c_void_p(d.a), d.a_base, d.a_len, C.c_void_p(d.c), d.c_base, d.c_len)
    if st != 0:
        raise_status(st, _lib.last_error())
    a, b = _scalar_buffer(alpha, T), _scalar_buffer(beta, T)
    s = C.c_void_p(torch.cuda.current_stream(Cv.parent.device).cuda_stream)
    st = lib.qs_add_execute(handle, a, b, s)
    if st != 0:
        raise_status(st, _lib.last_error())


# Add plans by layout: types, conj and the three groups (lengths and both
# stride maps), with the device they live on.  A hit rebinds the storages and
# reuses the device tables; the oldest plan is freed past CACHE_SIZE.
CACHE_SIZE = 64
_cache: "OrderedDict[tuple, C.c_void_p]" = OrderedDict()


def _group_key(g):
    n = g.ndim
    return (n, tuple(g.lengths[:n]), tuple(g.strides[0][:n]), tuple(g.strides[1][:n]))


def _cached_plan(d):
    key = (torch.cuda.current_device(), d.a_type, d.c_type, d.compute_type, d.conj_a,
           _group_key(d.m), _group_key(d.n), _group_key(d.k))
    h = _cache.get(key)
    if h is not None:
        _cache.move_to_end(key)
        return h
    h = C.c_void_p()
    st = _lib.lib().qs_add_plan_create(C.byref(d), C.byref(h))
    if st != 0:
        raise_status(st, _lib.last_error())
    _cache[key] = h
    while len(_cache) > CACHE_SIZE:
        _, old = _cache.popitem(last=False)
        _lib.lib().qs_add_plan_destroy(old)
    return h


def clear_plan_cache():
    """Free every cached add plan."""
    while _cache:
        _, h = _cache.popitem(last=False)
        _lib.lib().qs_add_plan_destroy(h)


def tensoradd(C_, A_, pA, conjA, alpha, beta):
    """`TO.tensoradd!(C, A, pA, conjA, alpha, beta, backend)`; returns C."""
    Cv, Av = StridedView.of(C_), StridedView.of(A_)
    p = _linearize(pA)
    _check_perm(p, len(Av.size), "A")
    if len(p) != len(Cv.size):
        raise ArgumentError(f"invalid permutation of C: {pA}")
    for c, a in enumerate(p):
        if Cv.size[c] != Av.size[a - 1]:
            raise DimensionMismatch(f"axis {c + 1} of C has length {Cv.size[c]}, expected {Av.size[a - 1]}")
    _run(Cv, Av, [a - 1 for a in p], [], conjA, alpha, beta)
    return C_


def tensortrace(C_, A_, p, q, conjA, alpha, beta):
    """`TO.tensortrace!(C, A, p, q, conjA, alpha, beta, backend)`; returns C."""
    Cv, Av = StridedView.of(C_), StridedView.of(A_)
    pl = _linearize(p)
    q1, q2 = tuple(q[0]), tuple(q[1])
    if len(q1) != len(q2):
        raise ArgumentError(f"invalid trace pairs: {q}")
    _check_perm(pl + q1 + q2, len(Av.size), "A")
    if len(pl) != len(Cv.size):
        raise ArgumentError(f"invalid permutation of C: {p}")
    for a, b in zip(q1, q2):
        if Av.size[a - 1] != Av.size[b - 1]:
            raise DimensionMismatch(f"traced axes {a} and {b} of A differ in length")
    for c, a in enumerate(pl):
        if Cv.size[c] != Av.size[a - 1]:
            raise DimensionMismatch(f"axis {c + 1} of C has length {Cv.size[c]}, expected {Av.size[a - 1]}")
    _run(Cv, Av, [a - 1 for a in pl], [(a - 1, b - 1) for a, b in zip(q1, q2)], conjA, alpha, beta)
    return C_
