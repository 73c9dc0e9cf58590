"""ORACLE -- test infrastructure only.  ctypes wrapper of liboracle.so (the C
restatement of QuasiStrided.jl's contraction path, see qs_oracle.h) for
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline.  The product
never imports this module.

Storages are numpy arrays in host memory; groups are (lengths, strides0,
strides1) triples with the same meaning as include/qs_amd.h's qs_group.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
MAX_DIMS = 16

F32, F64, C32, C64 = 0, 1, 2, 3
NP_CODE = {np.dtype(np.float32): F32, np.dtype(np.float64): F64, np.dtype(np.complex64): C32,
           np.dtype(np.complex128): C64}


class _Group(C.Structure):
    _fields_ = [("ndim", C.c_int32), ("_pad", C.c_int32), ("lengths", C.c_int64 * MAX_DIMS),
                ("strides", (C.c_int64 * MAX_DIMS) * 2)]


class _Desc(C.Structure):
    _fields_ = [
        ("a_type", C.c_int32), ("b_type", C.c_int32), ("c_type", C.c_int32), ("compute_type", C.c_int32),
        ("conj_a", C.c_int32), ("conj_b", C.c_int32), ("memory", C.c_int32), ("precision", C.c_int32),
        ("m", _Group), ("n", _Group), ("k", _Group),
        ("a", C.c_void_p), ("a_base", C.c_int64), ("a_len", C.c_int64),
        ("b", C.c_void_p), ("b_base", C.c_int64), ("b_len", C.c_int64),
        ("c", C.c_void_p), ("c_base", C.c_int64), ("c_len", C.c_int64),
        ("target_rel_err", C.c_double), ("tile", C.c_int32), ("path_mode", C.c_int32),
        ("k_splits", C.c_int32), ("_reserved", C.c_int32),
    ]


_lib = None


def build():
    """Compile liboracle.so with the Makefile beside this file."""
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stderr[-4000:])
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = C.CDLL(LIB_PATH)
        I64P = C.POINTER(C.c_int64)
        h.qso_fill_offsets.argtypes = [C.c_int, I64P, C.c_int, I64P, C.c_int64, C.c_int64, I64P]
        h.qso_validate_group.argtypes = [C.c_int, I64P, C.c_int, I64P]
        h.qso_affine_ramp.argtypes = [C.c_int, I64P, C.c_int, I64P, I64P]
        h.qso_describe_block.argtypes = [I64P, C.c_int64, C.c_int64, C.c_int64, I64P]
        h.qso_contract.argtypes = [C.POINTER(_Desc), C.c_void_p, C.c_void_p, C.c_int64]
        h.qso_contract_blis.argtypes = [C.POINTER(_Desc), C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_int64]
        h.qso_blis_kernel.restype = C.c_char_p
        h.qso_blis_kernel.argtypes = [C.c_int]
        h.qso_cache_shares.argtypes = [I64P, I64P, I64P]
        _lib = h
    return _lib


def _i64(xs):
    xs = list(xs) or [0]
    return (C.c_int64 * len(xs))(*xs)


def fill_offsets(lengths, strides, first, count):
    """fill_offsets! over P maps; strides is a list of P stride tuples.
    Returns (status, [P lists])."""
    D, P = len(lengths), len(strides)
    flat = [s for st in strides for s in st]
    out = (C.c_int64 * max(P * max(count, 0), 1))(*([-7] * max(P * max(count, 0), 1)))
    r = lib().qso_fill_offsets(D, _i64(lengths), P, _i64(flat), first, count, out)
    n = max(count, 0)
    return r, [list(out[p * n:(p + 1) * n]) for p in range(P)]


def validate_group(lengths, strides):
    flat = [s for st in strides for s in st]
    return lib().qso_validate_group(len(lengths), _i64(lengths), len(strides), _i64(flat))


def affine_ramp(lengths, strides):
    P = len(strides)
    flat = [s for st in strides for s in st]
    steps = (C.c_int64 * max(P, 1))()
    r = lib().qso_affine_ramp(len(lengths), _i64(lengths), P, _i64(flat), steps)
    return bool(r), tuple(steps[:P])


def describe_block(buf, first, count):
    out = (C.c_int64 * 4)()
    r = lib().qso_describe_block(_i64(buf), len(buf), first, count, out)
    return r, (out[0], out[1], out[2], bool(out[3]))


def _group(lengths, s0, s1):
    g = _Group()
    g.ndim = len(lengths)
    for d in range(len(lengths)):
        g.lengths[d], g.strides[0][d], g.strides[1][d] = int(lengths[d]), int(s0[d]), int(s1[d])
    return g


def _scalar(x, code):
    z = complex(x)
    if code == F64:
        return (C.c_double * 2)(z.real, 0.0)
    if code == F32:
        return (C.c_float * 2)(z.real, 0.0)
    if code == C64:
        return (C.c_double * 2)(z.real, z.imag)
    return (C.c_float * 2)(z.real, z.imag)


def make_desc(A, a_base, B, b_base, Cst, c_base, mgroup, ngroup, kgroup, compute, conj_a=False, conj_b=False):
    """A qs_contract_desc over host numpy storages (1-D arrays)."""
    for x in (A, B, Cst):
        assert x.ndim == 1 and x.flags["C_CONTIGUOUS"]
    d = _Desc()
    d.a_type, d.b_type, d.c_type = NP_CODE[A.dtype], NP_CODE[B.dtype], NP_CODE[Cst.dtype]
    d.compute_type = compute
    d.conj_a, d.conj_b = int(conj_a), int(conj_b)
    d.memory = 1
    d.m, d.n, d.k = _group(*mgroup), _group(*ngroup), _group(*kgroup)
    d.a, d.a_base, d.a_len = A.ctypes.data, a_base, A.size
    d.b, d.b_base, d.b_len = B.ctypes.data, b_base, B.size
    d.c, d.c_base, d.c_len = Cst.ctypes.data, c_base, Cst.size
    d.tile = -1
    return d


def contract(A, a_base, B, b_base, Cst, c_base, mgroup, ngroup, kgroup, compute, alpha, beta,
             conj_a=False, conj_b=False, k_block=0, blis=False, blocking=(0, 0, 0)):
    """Run the ScalarKernel restatement (or the BLIS baseline) in place on
    `Cst`.  Returns the status code (0 = success)."""
    d = make_desc(A, a_base, B, b_base, Cst, c_base, mgroup, ngroup, kgroup, compute, conj_a, conj_b)
    al, be = _scalar(alpha, compute), _scalar(beta, compute)
    if blis:
        return lib().qso_contract_blis(C.byref(d), al, be, *blocking)
    return lib().qso_contract(C.byref(d), al, be, k_block)


def blis_kernel(code) -> str:
    return lib().qso_blis_kernel(code).decode()


def cache_shares():
    a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
    lib().qso_cache_shares(C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value
