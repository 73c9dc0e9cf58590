"""The engine API: `plan_contract`, `execute` (Julia's `execute!`), `contract`
(Julia's `contract!`), `ContractPlan` and `release` -- the host planner of
src/planning/plan.jl and src/execution/execute.jl over the gfx950 library.

Planning (labels -> M/N/K AxisGroups, conjugation folding, aliasing, M/N
swap) happens here on the host exactly as the reference does it; the groups,
storages and folded flags then go through the C ABI (`qs_plan_create`), and
`execute` is one `qs_execute` call on the current HIP stream.
"""
from __future__ import annotations

import ctypes as C
from collections import OrderedDict
from dataclasses import dataclass

import torch

from . import _lib
from . import allocators as _alloc
from .errors import ArgumentError, raise_status
from .labels import (AxisGroup, axis_length, classify_labels, group_from_labels, leading_unit_run,
                     order_contract_labels, order_free_labels, prefer_swap)
from .strided import DTYPE_CODE, StridedView

SUPPORTED_ELTYPES = (torch.float32, torch.float64, torch.complex64, torch.complex128)

# The M extent of one MFMA register fragment: the unit-stride run in C that
# lets the epilogue store a fragment's rows contiguously.  Plays the role of
# the CPU kernel's `m_tile` in the swap rule (labels.jl:269-270).
DEFAULT_M_TILE = 16


@dataclass(frozen=True)
class PathModes:
    """Internal, for tests and benchmarks (plan.jl:60-64): "never" rules out
    the dot / outer / read-in-place path; "always" forces read-in-place."""
    dot: str = "auto"
    outer: str = "auto"
    unpacked_b: str = "auto"


@dataclass(frozen=True)
class MFMAKernel:
    """The kernel argument of `plan_contract`: the register-tile height that
    the swap rule compares C's unit-stride run against (`tile_size(kernel, 1)`
    of the reference), and an optional workgroup-tile override (TileId)."""
    m_tile: int = DEFAULT_M_TILE
    tile: int = -1
    k_splits: int = 0

    def tile_size(self, i: int = 1) -> int:
        return self.m_tile if i == 1 else 16


# ----------------------------------------------------------------- types

def _is_complex(t: torch.dtype) -> bool:
    return t.is_complex


def _real_bits(t: torch.dtype) -> int:
    return 64 if t in (torch.float64, torch.complex128) else 32


def compute_type(TA, TB, TC, accumulator=None) -> torch.dtype:
    """`compute_type` (plan.jl:278-285): promote_type of the three, or the
    accumulator's precision in the promoted domain."""
    for t in (TA, TB, TC):
        if t not in SUPPORTED_ELTYPES:
            raise ArgumentError(f"plan_contract: eltypes (A, B, C) = ({TA}, {TB}, {TC}); each must be one of "
                                "Float32, Float64, ComplexF32, ComplexF64")
    if not _is_complex(TC) and (_is_complex(TA) or _is_complex(TB)):
        raise ArgumentError(f"plan_contract: a complex operand (A: {TA}, B: {TB}) needs a complex C, got {TC}")
    cplx = _is_complex(TA) or _is_complex(TB) or _is_complex(TC)
    bits = max(_real_bits(TA), _real_bits(TB), _real_bits(TC))
    if accumulator is not None:
        if accumulator not in (torch.float32, torch.float64):
            raise ArgumentError(f"accumulator must be None, float32 or float64, got {accumulator}")
        bits = _real_bits(accumulator)
    if cplx:
        return torch.complex128 if bits == 64 else torch.complex64
    return torch.float64 if bits == 64 else torch.float32


def op_conjugates(op: str) -> bool:
    """A total table with a throwing fallback (plan.jl:176-186)."""
    table = {"identity": False, "conj": True, "transpose": False, "adjoint": True}
    if op not in table:
        raise ArgumentError(f"unsupported StridedView.op {op}: QuasiStrided folds a view's `op` into the packing "
                            "transform and recognizes only identity/conj/transpose/adjoint")
    return table[op]


def isconj(v: StridedView, flag: bool) -> bool:
    """XOR of the flag and the view's op; always False for real data (plan.jl:191-192)."""
    return v.is_complex and (bool(flag) ^ op_conjugates(v.op))


def _mightalias(x: StridedView, y: StridedView) -> bool:
    a0, a1 = x.data_range()
    b0, b1 = y.data_range()
    return a0 < b1 and b0 < a1 and a1 > a0 and b1 > b0


# ------------------------------------------------------------------ plan

class ContractPlan:
    """Reusable plan (plan.jl:16-45): the M/N/K AxisGroups in the planner's
    order, the operand storages and bases (after an M/N swap the `A*` fields
    describe the original B), the folded transforms, and the device-side plan
    (offset tables, tile choice, scratch).  Use from one thread at a time."""

    def __init__(self, handle, kernel, mgroup, ngroup, kgroup, Astorage, Abase, Bstorage, Bbase, Cstorage, Cbase,
                 atransform, btransform, T, swapped, views):
        self._handle = handle
        self.kernel = kernel
        self.mgroup, self.ngroup, self.kgroup = mgroup, ngroup, kgroup
        self.Astorage, self.Abase = Astorage, Abase
        self.Bstorage, self.Bbase = Bstorage, Bbase
        self.Cstorage, self.Cbase = Cstorage, Cbase
        self.atransform, self.btransform = atransform, btransform
        self.T = T
        self.swapped = swapped
        self._views = views  # keeps every operand alive while the plan lives
        self._al_keep = None  # the allocator bridge, kept until the plan is destroyed
        self._cached = False  # owned by the layout cache
        self.desc = None  # the qs_contract_desc the device plan was built from
        info = _lib.QSPlanInfo()
        st = _lib.lib().qs_plan_get_info(handle, C.byref(info))
        if st != 0:
            raise_status(st, _lib.last_error())
        self.info = {f: getattr(info, f) for f, _ in _lib.QSPlanInfo._fields_}

    @property
    def path(self) -> str:
        return _lib.PATH_NAMES.get(self.info["path"], "?")

    @property
    def handle(self):
        if self._handle is None:
            raise ArgumentError("ContractPlan: used after release")
        return self._handle

    def release(self):
        """Free the plan (its workspace goes back to its allocator first).  A
        plan from the layout cache stays cached; this drops the reference."""
        if self._handle is not None and not self._cached:
            _lib.lib().qs_plan_destroy(self._handle)
        self._handle = None
        self._al_keep = None

    def release_workspace(self):
        """`release!(plan, allocator)` (workspace.jl:118-123): hand the device
        workspace back to the allocator now; the plan cannot execute after."""
        st = _lib.lib().qs_plan_release(self.handle)
        if st != 0:
            raise_status(st, _lib.last_error())
        _refresh_info(self)

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def __repr__(self):
        return (f"ContractPlan(T={self.T}, path={self.path}, M={self.info['m']}, N={self.info['n']}, "
                f"K={self.info['k']}, tile={self.info['tile_m']}x{self.info['tile_n']}x{self.info['tile_k']}, "
                f"k_splits={self.info['k_splits']}, swapped={self.swapped})")


def _refresh_info(plan: ContractPlan):
    info = _lib.QSPlanInfo()
    st = _lib.lib().qs_plan_get_info(plan.handle, C.byref(info))
    if st != 0:
        raise_status(st, _lib.last_error())
    plan.info = {f: getattr(info, f) for f, _ in _lib.QSPlanInfo._fields_}


def autotune(plan: ContractPlan, reps: int = 5, stream=None) -> dict:
    """Time every (workgroup tile, split-K) candidate on the plan's own
    operands (C is not modified), adopt the fastest, and remember it for the
    plan's shape class; later plans of that class start from it."""
    r = _lib.QSTuneResult()
    s = C.c_void_p(stream.cuda_stream) if stream is not None else _stream_of(plan)
    st = _lib.lib().qs_autotune(plan.handle, s, int(reps), C.byref(r))
    if st != 0:
        raise_status(st, _lib.last_error())
    _refresh_info(plan)
    return {f: getattr(r, f) for f, _ in _lib.QSTuneResult._fields_ if not f.startswith("_")}


def set_error_report(plan: ContractPlan, on=True):
    """Split-precision plans: sample-check every execute against a Dot2
    evaluation of the same entries (costs a stream synchronization).
    `on="guard"`: an execute whose sampled error exceeds the target is redone
    with more moduli, then with the plain MFMA kernel."""
    mode = 2 if on == "guard" else (1 if on else 0)
    st = _lib.lib().qs_plan_set_error_report(plan.handle, mode)
    if st != 0:
        raise_status(st, _lib.last_error())
    _refresh_info(plan)


def error_report(plan: ContractPlan) -> dict:
    r = _lib.QSErrorReport()
    st = _lib.lib().qs_plan_get_error_report(plan.handle, C.byref(r))
    if st != 0:
        raise_status(st, _lib.last_error())
    return {f: getattr(r, f) for f, _ in _lib.QSErrorReport._fields_ if not f.startswith("_")}


def autotune_save(path: str):
    st = _lib.lib().qs_autotune_save(str(path).encode())
    if st != 0:
        raise_status(st, _lib.last_error())


def autotune_load(path: str):
    st = _lib.lib().qs_autotune_load(str(path).encode())
    if st != 0:
        raise_status(st, _lib.last_error())


def autotune_clear():
    _lib.lib().qs_autotune_clear()


def release(plan: ContractPlan):
    """`release!` (workspace.jl:118-123): free the plan's device memory."""
    plan.release()


def _check_block(name, v):
    if v is not None and int(v) < 1:
        raise ArgumentError(f"Blocking requires {name} >= 1, got {name} = {v}")


def _fill_group(g: AxisGroup) -> _lib.QSGroup:
    return _lib.make_group(g.lengths, g.strides[0], g.strides[1])


def planned(C_, A_, indA, B_, indB, indC, *, kernel=None, conjA=False, conjB=False, m_block=None,
            k_block=None, n_block=None, accumulator=None, path_modes=None, precision="plain",
            target_error=0.0, allocator=None, cache=False) -> ContractPlan:
    """`planned` (plan.jl:197-262) up to the device plan.  `cache`: take the
    device plan from the layout cache (device memory, the library's
    allocator), rebinding its storages; its `release` leaves it cached."""
    Cv, Av, Bv = StridedView.of(C_), StridedView.of(A_), StridedView.of(B_)
    indA, indB, indC = tuple(indA), tuple(indB), tuple(indC)
    for name, ind, v in (("indA", indA, Av), ("indB", indB, Bv), ("indC", indC, Cv)):
        if len(ind) != v.ndim:
            raise ArgumentError(f"{name} has {len(ind)} labels for an operand of {v.ndim} axes")
    _check_block("m_block", m_block)
    _check_block("k_block", k_block)
    _check_block("n_block", n_block)
    T = compute_type(Av.dtype, Bv.dtype, Cv.dtype, accumulator)
    if isconj(Cv, False):
        raise ArgumentError(f"plan_contract: cannot write into a conjugated view (C has op {Cv.op}); writing a "
                            "conjugated output is not supported")
    if _mightalias(Cv, Av) or _mightalias(Cv, Bv):
        raise ArgumentError("plan_contract: C must not share memory with A or B")
    dev = {Av.parent.device.type, Bv.parent.device.type, Cv.parent.device.type}
    if len(dev) != 1:
        raise ArgumentError("plan_contract: A, B and C must all be in device memory or all in host memory")
    memory = _lib.QS_MEM_DEVICE if dev == {"cuda"} else _lib.QS_MEM_HOST
    atransform = "conj" if isconj(Av, conjA) else "identity"
    btransform = "conj" if isconj(Bv, conjB) else "identity"

    mlabels, nlabels, klabels = classify_labels(indA, indB, indC)
    morder = order_free_labels(mlabels, indC, Cv)
    norder = order_free_labels(nlabels, indC, Cv)
    real_T = not T.is_complex
    run_m = leading_unit_run(morder, indC, Cv)
    run_n = leading_unit_run(norder, indC, Cv) if real_T else 0
    mgroup = group_from_labels(morder, (indA, Av), (indC, Cv))
    ngroup = group_from_labels(norder, (indB, Bv), (indC, Cv))
    m_length, n_length = axis_length(mgroup), axis_length(ngroup)
    korder = order_contract_labels(klabels, indA, Av, morder, indB, Bv, norder, m_length, n_length)
    kgroup = group_from_labels(korder, (indA, Av), (indB, Bv))

    kern = kernel if kernel is not None else MFMAKernel()
    m_tile = kern.tile_size(1)
    swapped = real_T and prefer_swap(run_m, run_n, m_tile, m_tile)
    if swapped:
        kgroup = group_from_labels(korder, (indB, Bv), (indA, Av))
        mgroup, ngroup = ngroup, mgroup
        Av, Bv = Bv, Av
        atransform, btransform = btransform, atransform

    modes = path_modes or PathModes()
    d = _lib.QSContractDesc()
    d.a_type, d.b_type, d.c_type = DTYPE_CODE[Av.dtype], DTYPE_CODE[Bv.dtype], DTYPE_CODE[Cv.dtype]
    d.compute_type = DTYPE_CODE[T]
    d.conj_a = 1 if atransform == "conj" else 0
    d.conj_b = 1 if btransform == "conj" else 0
    d.memory = memory
    d.precision = {"plain": _lib.QS_PLAIN, "split": _lib.QS_SPLIT_INT8}[precision]
    d.m, d.n, d.k = _fill_group(mgroup), _fill_group(ngroup), _fill_group(kgroup)
    d.a, d.a_base, d.a_len = Av.parent.data_ptr(), Av.offset, Av.parent.numel()
    d.b, d.b_base, d.b_len = Bv.parent.data_ptr(), Bv.offset, Bv.parent.numel()
    d.c, d.c_base, d.c_len = Cv.parent.data_ptr(), Cv.offset, Cv.parent.numel()
    d.target_rel_err = float(target_error)
    d.tile = int(getattr(kern, "tile", -1))
    d.k_splits = int(getattr(kern, "k_splits", 0))
    d.path_mode = 1 if (modes.dot == "never" and modes.outer == "never") else 0
    views = (Av, Bv, Cv)
    if cache and memory == _lib.QS_MEM_DEVICE and _alloc.is_default(allocator):
        handle = _cached_handle(d)
        plan = ContractPlan(handle, kern, mgroup, ngroup, kgroup, Av.parent, Av.offset, Bv.parent, Bv.offset,
                            Cv.parent, Cv.offset, atransform, btransform, T, swapped, views)
        plan._cached = True
        plan.desc = d
        return plan
    handle = C.c_void_p()
    keep = None
    if _alloc.is_default(allocator):
        st = _lib.lib().qs_plan_create(C.byref(d), C.byref(handle))
    else:
        al, keep = _alloc.bridge(allocator)
        st = _lib.lib().qs_plan_create_with_allocator(C.byref(d), C.byref(al), C.byref(handle))
    if st != 0:
        raise_status(st, _lib.last_error())
    plan = ContractPlan(handle, kern, mgroup, ngroup, kgroup, Av.parent, Av.offset, Bv.parent, Bv.offset,
                        Cv.parent, Cv.offset, atransform, btransform, T, swapped, (Av, Bv, Cv))
    plan._al_keep = keep
    plan.desc = d   # the qs_contract_desc the device plan was built from
    return plan


# Device plans by layout for the TensorOperations backend, whose every call
# plans afresh: the descriptor without its storages (types, conj, precision,
# path and tile requests, the three groups with both stride maps) on one
# device.  A hit rebinds the storages (qs_plan_rebind, which re-checks the
# spans) and reuses the device tables and workspace: no device allocation per
# call.  The oldest plan is freed past PLAN_CACHE_SIZE.
PLAN_CACHE_SIZE = 64
_plan_cache: "OrderedDict[tuple, C.c_void_p]" = OrderedDict()


def _gkey(g):
    n = g.ndim
    return (n, tuple(g.lengths[:n]), tuple(g.strides[0][:n]), tuple(g.strides[1][:n]))


def _cached_handle(d) -> C.c_void_p:
    key = (torch.cuda.current_device(), d.a_type, d.b_type, d.c_type, d.compute_type, d.conj_a, d.conj_b,
           d.precision, d.target_rel_err, d.tile, d.k_splits, d.path_mode, _gkey(d.m), _gkey(d.n), _gkey(d.k))
    lib = _lib.lib()
    h = _plan_cache.get(key)
    if h is not None:
        _plan_cache.move_to_end(key)
        st = lib.qs_plan_rebind(h, C.c_void_p(d.a), d.a_base, d.a_len, C.c_void_p(d.b), d.b_base, d.b_len,
                                C.c_void_p(d.c), d.c_base, d.c_len)
        if st != 0:
            raise_status(st, _lib.last_error())
        return h
    h = C.c_void_p()
    st = lib.qs_plan_create(C.byref(d), C.byref(h))
    if st != 0:
        raise_status(st, _lib.last_error())
    _plan_cache[key] = h
    while len(_plan_cache) > PLAN_CACHE_SIZE:
        _, old = _plan_cache.popitem(last=False)
        lib.qs_plan_destroy(old)
    return h


def clear_plan_cache():
    """Free every cached contraction plan."""
    while _plan_cache:
        _, h = _plan_cache.popitem(last=False)
        _lib.lib().qs_plan_destroy(h)


def plan_contract(C_, A_, indA, B_, indB, indC, *, kernel=None, conjA=False, conjB=False, m_block=None,
                  k_block=None, n_block=None, allocator=None, accumulator=None, path_modes=None,
                  precision="plain", target_error=0.0) -> ContractPlan:
    """Plan `C[indC] = A[indA] * B[indB]` (plan.jl:115-133).  Operands are
    StridedViews or torch tensors (a lazily conjugated tensor is a `conj` op).
    `m_block`/`k_block`/`n_block` are validated as the reference does; the
    device picks its own workgroup tile.  `allocator` supplies the plan's
    device workspace (see `allocators`); None or DefaultAllocator: the
    library's."""
    return planned(C_, A_, indA, B_, indB, indC, kernel=kernel, conjA=conjA, conjB=conjB, m_block=m_block,
                   k_block=k_block, n_block=n_block, accumulator=accumulator, path_modes=path_modes,
                   precision=precision, target_error=target_error, allocator=allocator)


def _scalar_buffer(x, T: torch.dtype):
    """`convert(T, x)` into a C buffer of the compute type."""
    z = complex(x)
    if not T.is_complex:
        if z.imag != 0:
            raise ArgumentError(f"InexactError: cannot convert {x} to the real compute type {T}")
        return (C.c_double * 1)(z.real) if T == torch.float64 else (C.c_float * 1)(z.real)
    return (C.c_double * 2)(z.real, z.imag) if T == torch.complex128 else (C.c_float * 2)(z.real, z.imag)


def _stream_of(plan: ContractPlan):
    dev = plan.Cstorage.device
    if dev.type == "cuda":
        return C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    return C.c_void_p(0)


def execute(plan: ContractPlan, alpha, beta, stream=None):
    """`execute!` (execute.jl:11-20): C = alpha*A*B + beta*C on the current
    HIP stream (or `stream`, a torch.cuda.Stream); returns the C storage.
    beta == 0 never reads C; alpha == 0 or K == 0 never reads A or B; operand
    spans are checked before anything is written (BoundsError)."""
    a = _scalar_buffer(alpha, plan.T)
    b = _scalar_buffer(beta, plan.T)
    s = C.c_void_p(stream.cuda_stream) if stream is not None else _stream_of(plan)
    st = _lib.lib().qs_execute(plan.handle, a, b, s)
    if st != 0:
        raise_status(st, _lib.last_error())
    return plan.Cstorage


execute_ = execute  # spelled like the reference's `execute!`


def contract(C_, alpha, A_, indA, B_, indB, beta, indC, *, accumulator=None):
    """`contract!` (execute.jl:84-95): plan, run and release in one call; returns C."""
    plan = planned(C_, A_, indA, B_, indB, indC, accumulator=accumulator)
    try:
        execute(plan, alpha, beta)
    finally:
        plan.release()
    return C_


contract_ = contract
