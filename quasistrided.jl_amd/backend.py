"""`QuasiStridedBackend`: the TensorOperations plugin boundary
(src/integrations/tensoroperations.jl:24-107) over the gfx950 engine.

`tensorcontract` takes TensorOperations' (pA, pB, pAB) index tuples, which are
one-based axis numbers as in Julia, turns them into labels exactly as
`contraction_labels` does, and runs `planned`.
"""
from __future__ import annotations

from . import add as _add
from .errors import ArgumentError, DimensionMismatch
from .plan import compute_type, planned, execute
from .strided import StridedView


def linearize(p):
    return tuple(p[0]) + tuple(p[1])


def invperm(p):
    out = [0] * len(p)
    for i, v in enumerate(p):
        out[v - 1] = i + 1
    return tuple(out)


def contraction_labels(pA, pB, pAB):
    """TO's (pA, pB, pAB) -> one Int label per axis (tensoroperations.jl:40-47):
    1..NoA for A's open axes, NoA+1..NoA+NoB for B's, -1..-Nk for contracted
    pairs; indC = linearize(pAB)."""
    NoA, Nk = len(pA[0]), len(pA[1])
    qA = invperm(linearize(pA))
    qB = invperm(linearize(pB))
    indA = tuple(s if s <= NoA else -(s - NoA) for s in qA)
    indB = tuple(-s if s <= Nk else NoA + (s - Nk) for s in qB)
    return indA, indB, linearize(pAB)


def _is_perm(p, n):
    return sorted(p) == list(range(1, n + 1))


def _argcheck(C, A, pA, B, pB, pAB):
    """TensorOperations' argcheck/dimcheck_tensorcontract."""
    if not _is_perm(linearize(pA), A.ndim):
        raise ArgumentError(f"invalid permutation of A: {pA}")
    if not _is_perm(linearize(pB), B.ndim):
        raise ArgumentError(f"invalid permutation of B: {pB}")
    if len(pA[1]) != len(pB[0]):
        raise ArgumentError("pA and pB contract different numbers of axes")
    if not _is_perm(linearize(pAB), len(pA[0]) + len(pB[1])) or len(linearize(pAB)) != C.ndim:
        raise ArgumentError(f"invalid permutation of C: {pAB}")
    for a, b in zip(pA[1], pB[0]):
        if A.size[a - 1] != B.size[b - 1]:
            raise DimensionMismatch(f"contracted axes {a} of A and {b} of B differ in length")
    open_sizes = [A.size[i - 1] for i in pA[0]] + [B.size[i - 1] for i in pB[1]]
    for c, p in enumerate(linearize(pAB)):
        if C.size[c] != open_sizes[p - 1]:
            raise DimensionMismatch(f"axis {c + 1} of C has length {C.size[c]}, expected {open_sizes[p - 1]}")


class QuasiStridedBackend:
    """TensorOperations backend running contractions on the gfx950 engine.
    `accumulator`: None, torch.float32 or torch.float64 (plan.jl:82-89)."""

    def __init__(self, accumulator=None):
        import torch
        if accumulator not in (None, torch.float32, torch.float64):
            raise ArgumentError(f"QuasiStridedBackend: accumulator must be nothing, Float32 or Float64, got "
                                f"{accumulator}")
        self.accumulator = accumulator

    def tensorcontract(self, C, A, pA, conjA, B, pB, conjB, pAB, alpha, beta, allocator=None):
        """TO.tensorcontract! (tensoroperations.jl:75-88); returns C."""
        Cv, Av, Bv = StridedView.of(C), StridedView.of(A), StridedView.of(B)
        compute_type(Av.dtype, Bv.dtype, Cv.dtype, self.accumulator)
        _argcheck(Cv, Av, pA, Bv, pB, pAB)
        indA, indB, indC = contraction_labels(pA, pB, pAB)
        checkpoint = allocator.checkpoint() if callable(getattr(allocator, "checkpoint", None)) else None
        plan = planned(Cv, Av, indA, Bv, indB, indC, conjA=conjA, conjB=conjB, accumulator=self.accumulator,
                       allocator=allocator, cache=True)
        try:
            execute(plan, alpha, beta)
        finally:
            plan.release()
            if checkpoint is not None:
                allocator.reset(checkpoint)
        return C

    def tensoradd(self, C, A, pA, conjA, alpha, beta, allocator=None):
        """TO.tensoradd! (tensoroperations.jl:91-98, forwarded to StridedNative
        there): C = beta*C + alpha*permutedims(op(A), linearize(pA)) on the
        device; returns C."""
        return _add.tensoradd(C, A, pA, conjA, alpha, beta)

    def tensortrace(self, C, A, p, q, conjA, alpha, beta, allocator=None):
        """TO.tensortrace! (tensoroperations.jl:100-107, forwarded to
        StridedNative there): C = beta*C + alpha*(partial trace of op(A) over
        the pairs q, permuted by linearize(p)) on the device; returns C."""
        return _add.tensortrace(C, A, p, q, conjA, alpha, beta)


def tensorcontract(C, A, pA, conjA, B, pB, conjB, pAB, alpha, beta, backend, allocator=None):
    return backend.tensorcontract(C, A, pA, conjA, B, pB, conjB, pAB, alpha, beta, allocator)
