"""Steady-state executes allocate nothing, as the reference's `execute!` does
(test/execution/test_execute.jl:72-94): after a plan's first execute, 100
more executes of the same plan leave the library's device, pinned-host and
allocator-bridge counts unchanged, for device operands and for host
operands (whose staging buffers are kept by the plan)."""
import ctypes as C

import numpy as np
import pytest
import torch

from helpers import SV, oracle_of_plan, qs, rel_fro, storage_np

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def case(msg):
    print(f"[case] {msg}", flush=True)


def counts():
    s = qs._lib.QSAllocStats()
    assert qs._lib.lib().qs_alloc_stats_get(C.byref(s)) == 0
    return {f: getattr(s, f) for f, _ in qs._lib.QSAllocStats._fields_}


def views(rng, device, M=19, K=23, N=17, split_k=False):
    if split_k:
        M, K, N = 64, 4096, 64    # few tiles, long K: the plan splits K (partials workspace)
    A = SV(torch.from_numpy(rng.standard_normal(M * K)).to(device), (M, K), (1, M))
    B = SV(torch.from_numpy(rng.standard_normal(K * N)).to(device), (K, N), (1, K))
    Cv = SV(torch.from_numpy(np.zeros(M * N)).to(device), (M, N), (1, M))
    return Cv, A, B


@pytest.mark.parametrize("where", ["device", "host"])
@pytest.mark.parametrize("shape", ["small", "split_k", "permuted"])
def test_steady_state_executes_allocate_nothing(oracle, where, shape):
    case(f"100 executes of one plan, operands in {where} memory, {shape}: allocation counts unchanged")
    rng = np.random.default_rng(71)
    device = DEV if where == "device" else "cpu"
    if shape == "permuted":
        d = 12
        st = (1, d, d * d, d ** 3)
        A = SV(torch.from_numpy(rng.standard_normal(d ** 4)).to(device), (d,) * 4, st)
        B = SV(torch.from_numpy(rng.standard_normal(d ** 4)).to(device), (d,) * 4, st)
        Cv = SV(torch.from_numpy(np.zeros(d ** 4)).to(device), (d,) * 4, st)
        plan = qs.plan_contract(Cv, A, (1, 2, 3, 4), B, (4, 5, 2, 6), (1, 3, 5, 6), conjB=True)
    else:
        Cv, A, B = views(rng, device, split_k=shape == "split_k")
        plan = qs.plan_contract(Cv, A, (1, 2), B, (2, 3), (1, 3))
    if shape == "split_k" and where == "device":
        assert plan.info["k_splits"] > 1, plan
    qs.execute(plan, 1.0, 0.0)            # the first execute may size staging buffers
    torch.cuda.synchronize()
    before = counts()
    for i in range(100):
        qs.execute(plan, 1.0, 0.5 if i % 2 else 0.0)
    torch.cuda.synchronize()
    after = counts()
    assert after == before, (before, after)
    # and the result is still right
    c0 = storage_np(Cv.parent)
    qs.execute(plan, 1.0, 0.5)
    torch.cuda.synchronize()
    want = oracle_of_plan(oracle, plan, 1.0, 0.5, storage_np(plan.Astorage), storage_np(plan.Bstorage), c0)
    assert rel_fro(storage_np(Cv.parent), want) <= 1e-13
    plan.release()
