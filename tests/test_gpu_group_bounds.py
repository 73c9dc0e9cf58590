"""Groups against their members' storages: every member's spans are checked
exactly before an execute or a capture launches anything, a batched group's
coordinate offsets are checked when it is built, and a member rebound after
the group was built runs on its new storages."""
import ctypes as C

import numpy as np
import pytest
import torch

from helpers import SV, oracle_of_plan, qs, rel_fro, storage_np

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def case(msg):
    print(f"[case] {msg}", flush=True)


def gemm_views(rng, M, K, N, c_extra=0):
    A = SV(torch.from_numpy(rng.standard_normal(M * K)).to(DEV), (M, K), (1, M))
    B = SV(torch.from_numpy(rng.standard_normal(K * N)).to(DEV), (K, N), (1, K))
    Cv = SV(torch.from_numpy(rng.standard_normal(M * N + c_extra)).to(DEV), (M, N), (1, M))
    return Cv, A, B


def rebind(plan, A, B, Cst, c_len=None):
    lib = qs._lib.lib()
    st = lib.qs_plan_rebind(plan.handle, C.c_void_p(A.data_ptr()), 0, A.numel(), C.c_void_p(B.data_ptr()), 0,
                            B.numel(), C.c_void_p(Cst.data_ptr()), 0, Cst.numel() if c_len is None else c_len)
    assert st == 0, qs._lib.last_error()


def test_capture_checks_member_spans_before_any_launch():
    case("group capture: a member rebound to a short C is rejected with BoundsError, nothing written")
    rng = np.random.default_rng(61)
    views = [gemm_views(rng, 96, 40, 80) for _ in range(3)]
    plans = [qs.plan_contract(Cv, A, (1, 2), B, (2, 3), (1, 3)) for (Cv, A, B) in views]
    grp = qs.GroupPlan(plans)
    starts = [storage_np(v[0].parent) for v in views]
    # plan 2 now writes into a storage one element too short for its C span
    short = torch.full((96 * 80 - 1,), 7.0, dtype=torch.float64, device=DEV)
    rebind(plans[2], views[2][1].parent, views[2][2].parent, short)
    with pytest.raises(qs.BoundsError):
        grp.capture(repeats=1, alphas=1.0, betas=0.0)
    with pytest.raises(qs.BoundsError):
        grp.execute(1.0, 0.0)
    torch.cuda.synchronize()
    for v, c0 in zip(views, starts):
        assert np.array_equal(storage_np(v[0].parent), c0)
    assert torch.all(short == 7.0).item()
    grp.release()


def test_batched_group_rejects_out_of_range_offsets():
    case("batched group: a coordinate offset past C's storage is rejected at create")
    rng = np.random.default_rng(62)
    Cv, A, B = gemm_views(rng, 64, 32, 48)
    plan = qs.plan_contract(Cv, A, (1, 2), B, (2, 3), (1, 3))
    lib = qs._lib.lib()
    n = 3
    zeros = (C.c_int64 * n)(0, 0, 0)
    c_off = (C.c_int64 * n)(0, 0, 1)     # coordinate 2 ends one element past C
    h = C.c_void_p()
    st = lib.qs_group_create_batched(plan.handle, n, zeros, zeros, c_off, -1, C.byref(h))
    assert st == qs._lib.QS_ERR_BOUNDS, (st, qs._lib.last_error())
    assert not h.value
    c_ok = (C.c_int64 * n)(0, 0, 0)
    st = lib.qs_group_create_batched(plan.handle, n, zeros, zeros, c_ok, -1, C.byref(h))
    assert st == 0 and h.value
    lib.qs_group_destroy(h)


def test_member_rebound_after_grouping_runs_on_new_storages(oracle):
    case("group: a member rebound after the group was built writes its new C, old C untouched")
    rng = np.random.default_rng(63)
    views = [gemm_views(rng, 100 + 8 * i, 48, 72) for i in range(3)]
    plans = [qs.plan_contract(Cv, A, (1, 2), B, (2, 3), (1, 3)) for (Cv, A, B) in views]
    grp = qs.GroupPlan(plans)
    grp.execute(1.0, 0.5)
    torch.cuda.synchronize()
    # new storages for member 1, same layout
    M, K, N = 108, 48, 72
    A2 = torch.from_numpy(rng.standard_normal(M * K)).to(DEV)
    B2 = torch.from_numpy(rng.standard_normal(K * N)).to(DEV)
    C2 = torch.from_numpy(rng.standard_normal(M * N)).to(DEV)
    c2_start = storage_np(C2)
    old_c1 = storage_np(views[1][0].parent)
    rebind(plans[1], A2, B2, C2)
    graph = None
    for it in range(2):
        if it == 0:
            grp.execute(1.0, 0.5)
        else:
            graph = grp.capture(repeats=1, alphas=1.0, betas=0.5)
            graph.launch()
        torch.cuda.synchronize()
        assert np.array_equal(storage_np(views[1][0].parent), old_c1), it   # the old C is not written
        want = c2_start.reshape(N, M).T
        want = storage_np(A2).reshape(K, M).T @ storage_np(B2).reshape(N, K).T + 0.5 * want
        got = storage_np(C2).reshape(N, M).T
        assert rel_fro(got, want) <= 1e-13, it
        c2_start = storage_np(C2)
    graph.release()
    grp.release()
