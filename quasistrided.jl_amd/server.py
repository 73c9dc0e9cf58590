"""The continuous-batching request loop (qs_server_* of include/qs_amd.h)."""
from __future__ import annotations

import ctypes as C

from . import _lib
from .errors import raise_status
from .strided import StridedView


def _check(st):
    if st != 0:
        raise_status(st, _lib.last_error())


def _pair(x):
    z = complex(x)
    return (C.c_double * 2)(z.real, z.imag)


class Server:
    """Serve contractions over resident operands.  `budget_bytes` bounds the
    operand bytes held by admitted requests; `grouped=False` launches every
    request alone, in arrival order (the per-request baseline)."""

    def __init__(self, budget_bytes: int, max_batch: int = 64, grouped: bool = True):
        h = C.c_void_p()
        _check(_lib.lib().qs_server_create(int(budget_bytes), int(max_batch), 1 if grouped else 0, C.byref(h)))
        self._h = h
        self._keep = {}

    def submit(self, plan, alpha=1.0, beta=0.0, A=None, B=None, Cout=None) -> int:
        """Run `plan` (optionally against other storages of the same layout,
        given as StridedViews whose parents and offsets are used).  Returns a
        ticket."""
        def st(v):
            if v is None:
                return None, 0, 0
            v = StridedView.of(v)
            return v.parent.data_ptr(), v.offset, v.parent.numel()
        a, b, c = st(A), st(B), st(Cout)
        if plan.swapped and (A is not None or B is not None):
            a, b = b, a
        t = C.c_uint64()
        _check(_lib.lib().qs_server_submit(self._h, plan.handle, a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2],
                                           _pair(alpha), _pair(beta), C.byref(t)))
        self._keep[t.value] = (plan, A, B, Cout)
        return t.value

    def poll(self, ticket: int):
        done, lat = C.c_int32(), C.c_double()
        _check(_lib.lib().qs_server_poll(self._h, ticket, C.byref(done), C.byref(lat)))
        if done.value:
            self._keep.pop(ticket, None)
            return lat.value
        return None

    def wait(self, ticket: int) -> float:
        lat = C.c_double()
        _check(_lib.lib().qs_server_wait(self._h, ticket, C.byref(lat)))
        self._keep.pop(ticket, None)
        return lat.value

    def drain(self):
        _check(_lib.lib().qs_server_drain(self._h))

    def stats(self) -> dict:
        s = _lib.QSServerStats()
        _check(_lib.lib().qs_server_stats_get(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in _lib.QSServerStats._fields_}

    def reset_stats(self):
        _check(_lib.lib().qs_server_reset_stats(self._h))

    def close(self):
        if self._h is not None:
            _lib.lib().qs_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
