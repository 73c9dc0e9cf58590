"""qs_contract_noalloc: a contraction that allocates nothing (its offset
tables live in a fixed device array), against qs_plan_create + qs_execute on
the same descriptor: two calls in a row, the second on a smaller input, for
the GEMM, skinny, dot and outer paths and complex types; bitwise equal where
the plan runs K unsplit, within 1e-13 otherwise.  The library's allocation
counts do not move across the calls."""
import ctypes as C

import numpy as np
import pytest
import torch

from helpers import SV, qs, rel_fro, storage_np

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def case(msg):
    print(f"[case] {msg}", flush=True)


def lib():
    return qs._lib.lib()


def counts():
    s = qs._lib.QSAllocStats()
    assert qs._lib.lib().qs_alloc_stats_get(C.byref(s)) == 0
    return {f: getattr(s, f) for f, _ in qs._lib.QSAllocStats._fields_}


def rnd(rng, dtype, n):
    if np.dtype(dtype).kind == "c":
        return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(dtype)
    return rng.standard_normal(n).astype(dtype)


def colmajor(size):
    out, s = [], 1
    for n in size:
        out.append(s)
        s *= n
    return tuple(out)


def make(rng, kind, dtype, scale):
    """(C view, A, B, indA, indB, indC) of one workload; `scale` shrinks it."""
    if kind == "permuted":          # C[a,b,c,d] := A[a,e,b,f] * conj(B[f,c,e,d]), A a stride-2 slice
        d = 32 // scale
        P = rnd(rng, dtype, d * d * 2 * d * d)
        pst = colmajor((d, d, 2 * d, d))
        A = SV(torch.from_numpy(P).to(DEV), (d,) * 4, (pst[0], pst[1], 2 * pst[2], pst[3]))
        B = SV(torch.from_numpy(rnd(rng, dtype, d ** 4)).to(DEV), (d,) * 4, colmajor((d,) * 4))
        Cv = SV(torch.from_numpy(rnd(rng, dtype, d ** 4)).to(DEV), (d,) * 4, colmajor((d,) * 4))
        return Cv, A, B, (1, 2, 3, 4), (4, 5, 2, 6), (1, 3, 5, 6)
    if kind == "c_anb":              # C[a,n,b] := A[a,k,b] * B[k,n], 48 / 96
        a, k, b, n = 48 // scale, 96 // scale, 48 // scale, 96 // scale
        A = SV(torch.from_numpy(rnd(rng, dtype, a * k * b)).to(DEV), (a, k, b), colmajor((a, k, b)))
        B = SV(torch.from_numpy(rnd(rng, dtype, k * n)).to(DEV), (k, n), colmajor((k, n)))
        Cv = SV(torch.from_numpy(rnd(rng, dtype, a * n * b)).to(DEV), (a, n, b), colmajor((a, n, b)))
        return Cv, A, B, (1, 2, 3), (2, 4), (1, 4, 3)
    if kind == "skinny":             # M = 16: B read in place
        M, K, N = 16, 512 // scale, 2048 // scale
    elif kind == "dot":              # N = 1
        M, K, N = 4096 // scale, 512 // scale, 1
    else:                            # outer: K = 1
        M, K, N = 1024 // scale, 1, 768 // scale
    A = SV(torch.from_numpy(rnd(rng, dtype, M * K)).to(DEV), (M, K), (1, M))
    B = SV(torch.from_numpy(rnd(rng, dtype, K * N)).to(DEV), (K, N), (1, K))
    Cv = SV(torch.from_numpy(rnd(rng, dtype, M * N)).to(DEV), (M, N), (1, M))
    return Cv, A, B, (1, 2), (2, 3), (1, 3)


@pytest.mark.parametrize("kind,dtype", [("permuted", np.float64), ("permuted", np.complex128),
                                        ("c_anb", np.float64), ("c_anb", np.float32), ("skinny", np.float64),
                                        ("dot", np.float64), ("outer", np.complex64)])
def test_noalloc_matches_planned_execute(kind, dtype):
    L = lib()
    rng = np.random.default_rng(81)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for call, scale in enumerate((1, 2)):     # the second call on a smaller input
        case(f"qs_contract_noalloc {kind} {np.dtype(dtype).name} call {call} (1/{scale} size)")
        Cv, A, B, iA, iB, iC = make(rng, kind, dtype, scale)
        plan = qs.plan_contract(Cv, A, iA, B, iB, iC, conjB=kind == "permuted")
        c0 = Cv.parent.clone()
        al = qs.plan._scalar_buffer(1.25 - 0.5j if plan.T.is_complex else 1.25, plan.T)
        be = qs.plan._scalar_buffer(0.5 + 0.25j if plan.T.is_complex else 0.5, plan.T)
        st = qs._lib.lib().qs_execute(plan.handle, al, be, stream)
        assert st == 0, qs._lib.last_error()
        torch.cuda.synchronize()
        want = storage_np(Cv.parent)
        Cv.parent.copy_(c0)
        torch.cuda.synchronize()
        before = counts()
        st = L.qs_contract_noalloc(C.byref(plan.desc), al, be, stream)
        assert st == 0, qs._lib.last_error()
        torch.cuda.synchronize()
        assert counts() == before
        got = storage_np(Cv.parent)
        tol = 1e-13 if plan.T in (torch.float64, torch.complex128) else 1e-5
        assert rel_fro(got, want) <= tol, (kind, call)
        if plan.info["k_splits"] == 1:
            assert np.array_equal(got, want), (kind, call, plan)
        plan.release()


def test_noalloc_rejections():
    case("qs_contract_noalloc: host operands, split precision and oversized tables are rejected, C untouched")
    L = lib()
    rng = np.random.default_rng(82)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    one = (C.c_double * 1)(1.0)
    zero = (C.c_double * 1)(0.0)
    # tables past the fixed array: M + N + K = 600004
    M, K, N = 600000, 2, 2
    A = SV(torch.zeros(M * K, dtype=torch.float64, device=DEV), (M, K), (1, M))
    B = SV(torch.zeros(K * N, dtype=torch.float64, device=DEV), (K, N), (1, K))
    Cv = SV(torch.full((M * N,), 3.0, dtype=torch.float64, device=DEV), (M, N), (1, M))
    plan = qs.plan_contract(Cv, A, (1, 2), B, (2, 3), (1, 3))
    st = L.qs_contract_noalloc(C.byref(plan.desc), one, zero, stream)
    assert st == qs._lib.QS_ERR_UNSUPPORTED and "scratch" in qs._lib.last_error(), qs._lib.last_error()
    torch.cuda.synchronize()
    assert torch.all(Cv.parent == 3.0).item()
    plan.release()
    # split precision
    Av = SV(torch.from_numpy(rng.standard_normal(64 * 64)).to(DEV), (64, 64), (1, 64))
    Bv = SV(torch.from_numpy(rng.standard_normal(64 * 64)).to(DEV), (64, 64), (1, 64))
    Cw = SV(torch.zeros(64 * 64, dtype=torch.float64, device=DEV), (64, 64), (1, 64))
    sp = qs.plan_contract(Cw, Av, (1, 2), Bv, (2, 3), (1, 3), precision="split")
    assert L.qs_contract_noalloc(C.byref(sp.desc), one, zero, stream) == qs._lib.QS_ERR_UNSUPPORTED
    sp.release()
    # host operands
    Ah = SV(torch.from_numpy(rng.standard_normal(64 * 64)), (64, 64), (1, 64))
    Bh = SV(torch.from_numpy(rng.standard_normal(64 * 64)), (64, 64), (1, 64))
    Ch = SV(torch.zeros(64 * 64, dtype=torch.float64), (64, 64), (1, 64))
    hp = qs.plan_contract(Ch, Ah, (1, 2), Bh, (2, 3), (1, 3))
    assert L.qs_contract_noalloc(C.byref(hp.desc), one, zero, stream) == qs._lib.QS_ERR_UNSUPPORTED
    hp.release()
    # an out-of-bounds span: BoundsError before anything is launched
    Cs = SV(torch.full((64 * 64,), 5.0, dtype=torch.float64, device=DEV), (64, 64), (1, 64))
    p2 = qs.plan_contract(Cs, Av, (1, 2), Bv, (2, 3), (1, 3))
    d = p2.desc
    d.c_len = 64 * 64 - 1
    assert L.qs_contract_noalloc(C.byref(d), one, zero, stream) == qs._lib.QS_ERR_BOUNDS
    torch.cuda.synchronize()
    assert torch.all(Cs.parent == 5.0).item()
    p2.release()
