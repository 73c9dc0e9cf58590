"""Grouped execution of many plans in one persistent launch, and HIP-graph
capture of repeated sweeps (qs_group_* / qs_graph_* of include/qs_amd.h)."""
from __future__ import annotations

import ctypes as C

from . import _lib
from .errors import raise_status
from .plan import ContractPlan, _scalar_buffer


def _check(st):
    if st != 0:
        raise_status(st, _lib.last_error())


def _coef(values, n):
    if values is None:
        return None
    vals = list(values) if not isinstance(values, (int, float, complex)) else [values] * n
    if len(vals) != n:
        raise ValueError(f"expected {n} coefficients, got {len(vals)}")
    buf = (C.c_double * (2 * n))()
    for i, v in enumerate(vals):
        z = complex(v)
        buf[2 * i], buf[2 * i + 1] = z.real, z.imag
    return buf


class Graph:
    """A captured sweep; `launch()` replays it on the current stream."""

    def __init__(self, handle, group):
        self._h = handle
        self._group = group   # keeps the group (and its plans) alive

    def launch(self, stream=None):
        import torch
        s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
        _check(_lib.lib().qs_graph_launch(self._h, C.c_void_p(s)))

    def release(self):
        if self._h is not None:
            _lib.lib().qs_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class GroupPlan:
    """Many device-memory plans run as one grouped persistent launch (per
    compute type, gather direction and domain).  `tiles`: optional per-plan TileId
    (-1 = the grouped heuristic)."""

    def __init__(self, plans, tiles=None):
        self.plans = list(plans)
        n = len(self.plans)
        arr = (C.c_void_p * max(n, 1))(*[p.handle for p in self.plans])
        t = None
        if tiles is not None:
            t = (C.c_int32 * max(n, 1))(*[int(x) for x in tiles])
        h = C.c_void_p()
        _check(_lib.lib().qs_group_create(arr, n, t, C.byref(h)))
        self._h = h
        f = C.c_int64()
        _check(_lib.lib().qs_group_flops(h, C.byref(f)))
        self.flops = f.value

    def execute(self, alphas=1.0, betas=0.0, stream=None):
        import torch
        n = len(self.plans)
        s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
        _check(_lib.lib().qs_group_execute(self._h, _coef(alphas, n), _coef(betas, n), C.c_void_p(s)))

    def stats(self) -> dict:
        """Grouped persistent launches per execute, and the members that run
        on their own (qs_group_stats)."""
        n_l, n_s = C.c_int32(), C.c_int32()
        _check(_lib.lib().qs_group_stats(self._h, C.byref(n_l), C.byref(n_s)))
        return {"launches": n_l.value, "standalone": n_s.value}

    def autotune(self, reps=3, stream=None) -> dict:
        """Time every grouped tile class per shape class of the members (as
        grouped launches writing a scratch C), keep the fastest in the autotune
        cache and rebuild the group with it."""
        import torch
        s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
        r = _lib.QSGroupTuneResult()
        _check(_lib.lib().qs_group_autotune(self._h, C.c_void_p(s), int(reps), C.byref(r)))
        return {f: getattr(r, f) for f, _ in _lib.QSGroupTuneResult._fields_ if not f.startswith("_")}

    def capture(self, repeats=1, alphas=1.0, betas=0.0, stream=None) -> Graph:
        import torch
        n = len(self.plans)
        s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
        g = C.c_void_p()
        _check(_lib.lib().qs_group_capture(self._h, _coef(alphas, n), _coef(betas, n), int(repeats),
                                           C.c_void_p(s), C.byref(g)))
        return Graph(g, self)

    def release(self):
        if self._h is not None:
            _lib.lib().qs_group_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass
