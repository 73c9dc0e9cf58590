"""Allocation failures inside the library, forced from the host with
qs_fault_inject: each device or pinned-host allocation that a scenario makes
(plan, execute, staging, split workspace, group, graph, batched group,
tensoradd, autotune) is failed in turn.  Every time, the scenario either
completes with the right result (where the library has a fallback: split-K
partial sums run unsplit) or raises MemoryError with every C untouched;
nothing is left allocated once its objects are freed; and the next run, with
nothing armed, succeeds.  The reference raises every error before any write
(src/execution/interface.jl:164-181); these tests hold the library's own
failure paths to that rule."""
import ctypes as C

import numpy as np
import pytest
import torch

from helpers import SV, qs, rel_fro, storage_np

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
DEVICE_ALLOC, PINNED_ALLOC = qs._lib.QS_FAULT_DEVICE_ALLOC, qs._lib.QS_FAULT_PINNED_ALLOC


def case(msg):
    print(f"[case] {msg}", flush=True)


def counts():
    s = qs._lib.QSAllocStats()
    assert qs._lib.lib().qs_alloc_stats_get(C.byref(s)) == 0
    return s


def live():
    """Device and pinned allocations not yet freed."""
    s = counts()
    return s.device_allocs - s.device_frees, s.pinned_allocs - s.pinned_frees


def arm(kind, after):
    assert qs._lib.lib().qs_fault_inject(kind, after) == 0


def gemm_views(rng, M, K, N, device=DEV):
    A = SV(torch.from_numpy(rng.standard_normal(M * K)).to(device), (M, K), (1, M))
    B = SV(torch.from_numpy(rng.standard_normal(K * N)).to(device), (K, N), (1, K))
    Cv = SV(torch.from_numpy(rng.standard_normal(M * N)).to(device), (M, N), (1, M))
    return Cv, A, B


# Each scenario returns (run, outs): `run(keep)` makes the library calls,
# appending what it creates to `keep` (released afterwards, last first), with
# its one writing call last; `outs` are the storages it writes.

def sc_plain(rng):
    Cv, A, B = gemm_views(rng, 200, 96, 150)

    def run(keep):
        p = qs.plan_contract(Cv, A, (1, 2), B, (2, 3), (1, 3))
        keep.append(p)
        qs.execute(p, 1.0, 0.5)
    return run, [Cv.parent]


def sc_host(rng):
    Cv, A, B = gemm_views(rng, 120, 64, 90, device="cpu")

    def run(keep):
        p = qs.plan_contract(Cv, A, (1, 2), B, (2, 3), (1, 3))
        keep.append(p)
        qs.execute(p, 1.0, 0.5)
    return run, [Cv.parent]


def sc_split_k(rng):
    Cv, A, B = gemm_views(rng, 64, 4096, 64)

    def run(keep):
        p = qs.plan_contract(Cv, A, (1, 2), B, (2, 3), (1, 3))
        keep.append(p)
        qs.execute(p, 1.0, 0.5)
    return run, [Cv.parent]


def sc_split_guarded(rng):
    Cv, A, B = gemm_views(rng, 256, 256, 256)

    def run(keep):
        p = qs.plan_contract(Cv, A, (1, 2), B, (2, 3), (1, 3), precision="split")
        keep.append(p)
        qs.set_error_report(p, "guard")
        qs.execute(p, 1.0, 0.5)
    return run, [Cv.parent]


def sc_group_graph(rng):
    views = [gemm_views(rng, 64 + 40 * i, 48 + 16 * i, 72) for i in range(3)]

    def run(keep):
        plans = [qs.plan_contract(Cv, A, (1, 2), B, (2, 3), (1, 3)) for (Cv, A, B) in views]
        keep.extend(plans)
        grp = qs.GroupPlan(plans)
        keep.append(grp)
        g = grp.capture(repeats=1, alphas=1.0, betas=0.5)
        keep.append(g)
        g.launch()
    return run, [v[0].parent for v in views]


def sc_batched(rng):
    nb, M, K, N = 6, 40, 24, 32
    A = torch.from_numpy(rng.standard_normal(M * K * nb)).to(DEV)
    B = torch.from_numpy(rng.standard_normal(K * N * nb)).to(DEV)
    Cst = torch.from_numpy(rng.standard_normal(M * N * nb)).to(DEV)
    Av = SV(A, (M, K, nb), (1, M, M * K))
    Bv = SV(B, (K, N, nb), (1, K, K * N))
    Cv = SV(Cst, (M, N, nb), (1, M, M * N))

    def run(keep):
        bp = qs.plan_batched(Cv, Av, (1, 2, 9), Bv, (2, 3, 9), (1, 3, 9))
        keep.append(bp)
        bp.execute(1.0, 0.5)
    return run, [Cst]


def sc_tensoradd(rng):
    a = torch.from_numpy(rng.standard_normal(70 * 90)).to(DEV)
    c = torch.from_numpy(rng.standard_normal(90 * 70)).to(DEV)
    Av = SV(a, (70, 90), (1, 70))
    Cv = SV(c, (90, 70), (1, 90))

    def run(keep):
        qs.tensoradd(Cv, Av, ((2, 1), ()), False, 1.0, 0.5)
    return run, [c]


def sc_autotune(rng):
    Cv, A, B = gemm_views(rng, 160, 64, 96)

    def run(keep):
        p = qs.plan_contract(Cv, A, (1, 2), B, (2, 3), (1, 3))
        keep.append(p)
        qs.autotune(p, reps=1)
        qs.execute(p, 1.0, 0.5)
    return run, [Cv.parent]


def sc_group_autotune(rng):
    views = [gemm_views(rng, 80 + 48 * i, 40, 64 + 16 * i) for i in range(3)]

    def run(keep):
        plans = [qs.plan_contract(Cv, A, (1, 2), B, (2, 3), (1, 3)) for (Cv, A, B) in views]
        keep.extend(plans)
        grp = qs.GroupPlan(plans)
        keep.append(grp)
        grp.autotune(reps=1)
        grp.execute(1.0, 0.5)
    return run, [v[0].parent for v in views]


SCENARIOS = {"plain": sc_plain, "host": sc_host, "split_k": sc_split_k, "split_guarded": sc_split_guarded,
             "group_graph": sc_group_graph, "batched": sc_batched, "tensoradd": sc_tensoradd,
             "autotune": sc_autotune, "group_autotune": sc_group_autotune}


def cleanup(keep):
    for obj in reversed(keep):
        obj.release()
    keep.clear()
    qs.add.clear_plan_cache()
    torch.cuda.synchronize()


def attempt(run, outs, init, kind=None, after=-1):
    """Run once from `init` with a fault armed; returns (error or None, outputs)."""
    for o, i in zip(outs, init):
        o.copy_(i)
    torch.cuda.synchronize()
    keep = []
    err = None
    if kind is not None:
        arm(kind, after)
    try:
        run(keep)
        torch.cuda.synchronize()
    except MemoryError as e:
        err = e
    finally:
        if kind is not None:
            arm(kind, -1)
    got = [storage_np(o) for o in outs]
    cleanup(keep)
    return err, got


@pytest.mark.parametrize("name", list(SCENARIOS))
def test_every_allocation_failure_is_clean(name):
    case(f"fault injection over every allocation of scenario {name}")
    qs.autotune_clear()
    rng = np.random.default_rng(91)
    run, outs = SCENARIOS[name](rng)
    init = [o.clone() for o in outs]
    start = live()
    s0 = counts()
    err, want = attempt(run, outs, init)
    assert err is None
    s1 = counts()
    assert live() == start, "a clean run leaves nothing allocated"
    n_dev = s1.device_allocs - s0.device_allocs
    n_pin = s1.pinned_allocs - s0.pinned_allocs
    assert n_dev + n_pin > 0, "the scenario allocates"
    failed = 0
    for kind, n in ((DEVICE_ALLOC, n_dev), (PINNED_ALLOC, n_pin)):
        for after in range(n):
            print(f"[case] {name}: kind {kind}, allocation {after + 1} of {n} fails", flush=True)
            qs.autotune_clear()
            err, got = attempt(run, outs, init, kind, after)
            assert live() == start, (name, kind, after, "leak")
            if err is None:
                # A fallback path: the result is still right.
                for g, w in zip(got, want):
                    assert rel_fro(g, w) <= 1e-13, (name, kind, after)
            else:
                failed += 1
                for g, i in zip(got, init):
                    assert np.array_equal(g, storage_np(i)), (name, kind, after, "C written before the error")
    assert failed > 0, "at least one injected failure surfaces as an error"
    # Disarmed: the scenario runs again, with the same result.
    err, again = attempt(run, outs, init)
    assert err is None
    for g, w in zip(again, want):
        assert rel_fro(g, w) <= 1e-13
    assert live() == start
