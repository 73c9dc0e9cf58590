"""The C-ABI library loads and exports every symbol include/qs_amd.h declares;
host-only entry points (no GPU compute) behave like the reference.  CPU only."""
import ctypes as C
import os
import re

import pytest

from helpers import qs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(qs_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(os.path.join(ROOT, "quasistrided.jl_amd", "libqs_amd.so"))
    syms = declared_symbols()
    assert "qs_plan_create" in syms and "qs_execute" in syms
    for s in sorted(syms):
        assert hasattr(lib, s), s
    # and the Python binding types every one of them
    assert syms <= set(qs._lib.SIGNATURES), syms - set(qs._lib.SIGNATURES)


def test_version_and_fill_offsets_host_side(oracle):
    assert b"gfx950" in qs._lib.lib().qs_version()
    g = qs.AxisGroup((3, 2), ((1, 15), (1, 12)))
    assert g.fill_offsets(0, 6) == ([0, 1, 2, 15, 16, 17], [0, 1, 2, 12, 13, 14])
    assert g.fill_offsets(2, 2) == ([2, 15], [2, 12])
    with pytest.raises(qs.BoundsError):
        g.fill_offsets(6, 1)
    with pytest.raises(qs.ArgumentError):
        g.fill_offsets(0, -1)
    # against the oracle on random groups
    import numpy as np
    rng = np.random.default_rng(3)
    for _ in range(100):
        D = int(rng.integers(0, 4))
        lengths = tuple(int(x) for x in rng.integers(1, 5, D))
        s = [tuple(int(x) for x in rng.integers(-8, 9, D)) for _ in range(2)]
        Q = int(np.prod(lengths)) if D else 1
        first = int(rng.integers(0, Q + 1))
        count = int(rng.integers(0, Q - first + 1))
        r, want = oracle.fill_offsets(lengths, s, first, count)
        assert r == 0
        assert list(qs.AxisGroup(lengths, tuple(s)).fill_offsets(first, count)) == want


def _desc(**kw):
    d = qs._lib.QSContractDesc()
    d.a_type = d.b_type = d.c_type = d.compute_type = 1
    d.m = qs._lib.make_group((4,), (1,), (1,))
    d.n = qs._lib.make_group((4,), (4,), (4,))
    d.k = qs._lib.make_group((4,), (4,), (1,))
    d.a_len = d.b_len = d.c_len = 16
    d.tile = -1
    for k, v in kw.items():
        setattr(d, k, v)
    return d


@pytest.mark.parametrize("kw,status", [
    (dict(a_type=9), 1),
    (dict(a_type=3, c_type=1, compute_type=3), 1),        # complex A into real C
    (dict(memory=5), 1),
    (dict(a_len=-1), 1),
])
def test_plan_create_rejects_invalid_input_before_touching_the_device(kw, status):
    lib = qs._lib.lib()
    h = C.c_void_p()
    d = _desc(**kw)
    assert lib.qs_plan_create(C.byref(d), C.byref(h)) == status
    assert h.value is None
    assert qs._lib.last_error()


def test_plan_create_overflow_policy():
    lib = qs._lib.lib()
    big = (1 << 63) - 1
    d = _desc()
    d.m = qs._lib.make_group((big // 2 + 2, big // 2 + 2), (1, 1), (1, 1))
    h = C.c_void_p()
    assert lib.qs_plan_create(C.byref(d), C.byref(h)) == 4
    d = _desc()
    d.k = qs._lib.make_group((-1,), (1,), (1,))
    assert lib.qs_plan_create(C.byref(d), C.byref(h)) == 1


def test_destroy_null_is_noop():
    assert qs._lib.lib().qs_plan_destroy(None) == 0


def test_fault_inject_arms_and_rejects_unknown_kinds():
    lib = qs._lib.lib()
    assert lib.qs_fault_inject(7, 0) == qs._lib.QS_ERR_ARGUMENT
    for kind in (qs._lib.QS_FAULT_DEVICE_ALLOC, qs._lib.QS_FAULT_PINNED_ALLOC):
        assert lib.qs_fault_inject(kind, 3) == 0
        assert lib.qs_fault_inject(kind, -1) == 0   # disarmed again


def test_autotune_cache_file_roundtrip(tmp_path):
    qs.autotune_clear()
    p = tmp_path / "tune.txt"
    qs.autotune_save(str(p))
    assert p.read_text().startswith("#")
    p.write_text("1,1,1,1,0,1,20,20,20 1 2 0.5\n")
    qs.autotune_load(str(p))
    q = tmp_path / "again.txt"
    qs.autotune_save(str(q))
    assert "1,1,1,1,0,1,20,20,20 1 2" in q.read_text()
    bad = tmp_path / "bad.txt"
    bad.write_text("garbage\n")
    with pytest.raises(qs.ArgumentError):
        qs.autotune_load(str(bad))
    with pytest.raises(qs.ArgumentError):
        qs.autotune_load(str(tmp_path / "missing.txt"))
    qs.autotune_clear()


def test_bench_refuses_more_than_one_gpu():
    # One GPU is the scope: `bench.py --gpus N` with N > 1 exits with a
    # message before it touches a device.
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2, (r.returncode, r.stderr)
    assert "one GPU" in r.stderr and r.stdout == ""
