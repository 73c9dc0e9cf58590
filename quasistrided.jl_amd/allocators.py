"""Workspace allocators: plan_contract's `allocator` (src/planning/plan.jl:125-131)
bridged to device memory pools.

In the reference, a plan's packed panels and panel of C come from a
TensorOperations allocator (`tensoralloc`, src/execution/workspace.jl:83-106)
and go back to it in reverse order (`release!`, workspace.jl:109-123).  Here
the plan's device workspace -- split-K partial sums and the split-precision
residues, scales and samples -- comes from the allocator when the plan is
built (and when autotuning grows it) and goes back to it, split-precision
buffer first, when the plan is released.  The offset tables stay the
library's, as the reference's offset vectors are plain Vectors.

* `DefaultAllocator()` (or None): the library allocates (hipMalloc / hipFree).
* `TorchAllocator(device)`: torch's caching allocator, the stream-ordered pool
  of HBM that the rest of a torch program draws from.
* `PoolAllocator(device, release_threshold)`: a device memory pool inside the
  library (`hipMemPool_t`), for callers with no pool of their own; the
  library calls it directly, with no Python callback per block.
* Any object with `alloc(nbytes) -> int` (a device address aligned to 256
  bytes, or 0 when it has none) and `free(ptr, nbytes)`.  Optional
  `checkpoint()` / `reset(checkpoint)` are called around a backend
  contraction, as TensorOperations' `allocator_checkpoint!` /
  `allocator_reset!` are (src/integrations/tensoroperations.jl:84-86).
"""
from __future__ import annotations

import ctypes as C
import sys

from . import _lib


class DefaultAllocator:
    """The library's own device allocation (TensorOperations.DefaultAllocator)."""


class TorchAllocator:
    """Workspace from torch's caching allocator on `device`.  Counts what it
    handed out and took back (`allocs`, `frees`, `live_bytes`)."""

    def __init__(self, device=None):
        import torch
        self.device = torch.cuda.current_device() if device is None else torch.device(device).index or 0
        self.allocs = 0
        self.frees = 0
        self.live_bytes = 0
        self.log = []   # ("alloc" | "free", ptr, nbytes), in call order

    def alloc(self, nbytes: int) -> int:
        import torch
        ptr = int(torch.cuda.caching_allocator_alloc(int(nbytes), self.device))
        self.allocs += 1
        self.live_bytes += int(nbytes)
        self.log.append(("alloc", ptr, int(nbytes)))
        return ptr

    def free(self, ptr: int, nbytes: int) -> None:
        import torch
        torch.cuda.caching_allocator_delete(int(ptr))
        self.frees += 1
        self.live_bytes -= int(nbytes)
        self.log.append(("free", int(ptr), int(nbytes)))


class PoolAllocator:
    """Workspace from a device memory pool held by the library
    (`qs_pool_allocator_create`).  The pool keeps up to `release_threshold`
    bytes reserved once blocks come back (None: everything it has grown to),
    so plans built and destroyed in a loop reuse its memory.  `close()` (or
    garbage collection) destroys the pool; it refuses while plans hold
    blocks."""

    def __init__(self, device=None, release_threshold=None):
        import torch
        from .errors import raise_status
        self.device = torch.cuda.current_device() if device is None else torch.device(device).index or 0
        thr = (1 << 64) - 1 if release_threshold is None else int(release_threshold)
        if thr < 0:
            raise ValueError("PoolAllocator: release_threshold must be >= 0")
        self._c_allocator = _lib.QSAllocator()
        st = _lib.lib().qs_pool_allocator_create(int(self.device), thr, C.byref(self._c_allocator))
        if st != 0:
            self._c_allocator = None
            raise_status(st, _lib.last_error())

    def stats(self) -> dict:
        """`reserved_bytes` held from the device, `used_bytes` handed out now,
        `live_blocks` held by plans."""
        from .errors import raise_status
        if self._c_allocator is None:
            raise ValueError("PoolAllocator: closed")
        r, u, n = C.c_uint64(), C.c_uint64(), C.c_int64()
        st = _lib.lib().qs_pool_allocator_stats(C.byref(self._c_allocator), C.byref(r), C.byref(u), C.byref(n))
        if st != 0:
            raise_status(st, _lib.last_error())
        return {"reserved_bytes": r.value, "used_bytes": u.value, "live_blocks": n.value}

    def close(self) -> None:
        from .errors import raise_status
        if self._c_allocator is None:
            return
        st = _lib.lib().qs_pool_allocator_destroy(C.byref(self._c_allocator))
        if st != 0:
            raise_status(st, _lib.last_error())
        self._c_allocator = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def is_default(allocator) -> bool:
    return allocator is None or isinstance(allocator, DefaultAllocator)


def bridge(allocator):
    """The C `qs_allocator` for a Python allocator, and the objects that must
    outlive every plan built with it (the ctypes callbacks)."""
    c_al = getattr(allocator, "_c_allocator", None)
    if isinstance(allocator, PoolAllocator):
        if c_al is None:
            raise ValueError("PoolAllocator: closed")
        return c_al, (c_al, allocator)
    if not (callable(getattr(allocator, "alloc", None)) and callable(getattr(allocator, "free", None))):
        raise TypeError(f"allocator {allocator!r}: needs alloc(nbytes) and free(ptr, nbytes)")

    def _alloc(ctx, nbytes):
        try:
            return int(allocator.alloc(int(nbytes)) or 0)
        except Exception as e:   # reported as an allocation failure
            print(f"quasistrided_amd: allocator.alloc({nbytes}) raised {e!r}", file=sys.stderr)
            return 0

    def _free(ctx, ptr, nbytes):
        try:
            allocator.free(int(ptr), int(nbytes))
        except Exception as e:
            print(f"quasistrided_amd: allocator.free raised {e!r}", file=sys.stderr)

    fa = _lib.QS_ALLOC_FN(_alloc)
    ff = _lib.QS_FREE_FN(_free)
    al = _lib.QSAllocator(fa, ff, None)
    return al, (al, fa, ff, allocator)
