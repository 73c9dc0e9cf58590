"""The allocator bridge on the host: a Python allocator becomes the C
`qs_allocator` whose function pointers call it (as the library does when a
plan acquires and releases its workspace); the default needs no bridge."""
import ctypes as C

import pytest

from helpers import qs

alloc_mod = qs.allocators
_lib = qs._lib


class Arena:
    """A bump allocator over a fake address range (never dereferenced)."""

    def __init__(self, base=1 << 40):
        self.top = base
        self.calls = []

    def alloc(self, nbytes):
        p = self.top
        self.top += (nbytes + 255) // 256 * 256
        self.calls.append(("alloc", p, nbytes))
        return p

    def free(self, ptr, nbytes):
        self.calls.append(("free", ptr, nbytes))


def test_default_needs_no_bridge():
    assert alloc_mod.is_default(None)
    assert alloc_mod.is_default(qs.DefaultAllocator())
    assert not alloc_mod.is_default(Arena())


def test_bridge_calls_the_python_allocator():
    a = Arena()
    al, keep = alloc_mod.bridge(a)
    assert isinstance(al, _lib.QSAllocator)
    p1 = al.alloc(None, 1000)
    p2 = al.alloc(None, 24)
    assert p1 == 1 << 40 and p2 == (1 << 40) + 1024
    al.free(None, p2, 24)
    al.free(None, p1, 1000)
    assert a.calls == [("alloc", p1, 1000), ("alloc", p2, 24), ("free", p2, 24), ("free", p1, 1000)]


def test_bridge_reports_a_raising_allocator_as_no_memory(capsys):
    class Broken(Arena):
        def alloc(self, nbytes):
            raise RuntimeError("pool exhausted")

    al, keep = alloc_mod.bridge(Broken())
    assert al.alloc(None, 64) is None   # NULL: the library's QS_ERR_MEMORY / unsplit fallback
    assert "pool exhausted" in capsys.readouterr().err


def test_bridge_rejects_objects_without_alloc_and_free():
    with pytest.raises(TypeError):
        alloc_mod.bridge(object())


def test_struct_layout_matches_the_header():
    # qs_allocator: two function pointers and a context pointer.
    assert C.sizeof(_lib.QSAllocator) == 3 * C.sizeof(C.c_void_p)
    assert [f for f, _ in _lib.QSAllocator._fields_] == ["alloc", "free", "ctx"]
