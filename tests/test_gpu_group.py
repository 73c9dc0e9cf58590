"""Grouped persistent execution and graph capture on the GPU: every member
against the oracle, the same bits alone and in the mix, graph replays equal
to eager runs."""
import numpy as np
import pytest
import torch

from helpers import RTOL, SV, oracle_of_plan, qs, rel_fro, storage_np

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def case(msg):
    print(f"[case] {msg}", flush=True)


def mps_env(rng, dtype, a, ap, b, sigma, w=5):
    """L[a,w,a'] * psi[a',s,b] -> C[a,w,s,b] (column-major, device)."""
    def rnd(n):
        if np.dtype(dtype).kind == "c":
            return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(dtype)
        return rng.standard_normal(n).astype(dtype)
    L = SV(torch.from_numpy(rnd(a * w * ap)).to(DEV), (a, w, ap), (1, a, a * w))
    P = SV(torch.from_numpy(rnd(ap * sigma * b)).to(DEV), (ap, sigma, b), (1, ap, ap * sigma))
    Cv = SV(torch.from_numpy(rnd(a * w * sigma * b)).to(DEV), (a, w, sigma, b), (1, a, a * w, a * w * sigma))
    return Cv, L, P


def dims(rng, n, lo=16, hi=200):
    out = []
    for _ in range(n):
        a, ap, b = (int(round(np.exp(rng.uniform(np.log(lo), np.log(hi))))) for _ in range(3))
        out.append((a, ap, b, int(rng.integers(2, 5))))
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.float32])
def test_group_members_match_oracle_and_standalone(oracle, dtype):
    case(f"group of 24 MPS env contractions {np.dtype(dtype).name}")
    rng = np.random.default_rng(3)
    shapes = dims(rng, 24)
    views, plans, starts = [], [], []
    for (a, ap, b, s) in shapes:
        Cv, L, P = mps_env(rng, dtype, a, ap, b, s)
        views.append((Cv, L, P))
        starts.append(storage_np(Cv.parent))
        plans.append(qs.plan_contract(Cv, L, (1, 2, 3), P, (3, 4, 5), (1, 2, 4, 5)))
    cplx = np.dtype(dtype).kind == "c"
    alphas = [(1.0 + 0.25j * i) if cplx else 1.0 + 0.1 * i for i in range(len(plans))]
    betas = [(0.5 - 0.5j) if cplx else 0.5 for _ in plans]
    grp = qs.GroupPlan(plans)
    grp.execute(alphas, betas)
    torch.cuda.synchronize()
    mixed = [storage_np(v[0].parent) for v in views]
    for i, (plan, (Cv, L, P)) in enumerate(zip(plans, views)):
        want = oracle_of_plan(oracle, plan, alphas[i], betas[i], storage_np(plan.Astorage), storage_np(plan.Bstorage),
                              starts[i])
        assert rel_fro(mixed[i], want) <= RTOL[plan.T], (i, plan)
    # alone vs in the mix: the same bits
    for i in (0, 5, 17):
        Cv = views[i][0]
        Cv.parent.copy_(torch.from_numpy(starts[i]).to(DEV))
        alone = qs.GroupPlan([plans[i]])
        alone.execute([alphas[i]], [betas[i]])
        torch.cuda.synchronize()
        assert np.array_equal(storage_np(Cv.parent), mixed[i]), i
        alone.release()
    grp.release()


def test_group_graph_replays_equal_eager(oracle):
    case("group graph capture, 3 replays with eager work between")
    rng = np.random.default_rng(9)
    shapes = dims(rng, 16)
    views, plans = [], []
    for (a, ap, b, s) in shapes:
        Cv, L, P = mps_env(rng, np.float64, a, ap, b, s)
        views.append(Cv)
        plans.append(qs.plan_contract(Cv, L, (1, 2, 3), P, (3, 4, 5), (1, 2, 4, 5)))
    grp = qs.GroupPlan(plans)
    grp.execute(1.0, 0.0)
    torch.cuda.synchronize()
    eager = [storage_np(v.parent) for v in views]
    graph = grp.capture(repeats=2, alphas=1.0, betas=0.0)
    for r in range(3):
        for v in views:
            v.parent.fill_(float("nan"))
        graph.launch()
        torch.cuda.synchronize()
        x = torch.randn(1000, device=DEV).sum().item()   # eager work between replays
        assert np.isfinite(x)
        for i, v in enumerate(views):
            assert np.array_equal(storage_np(v.parent), eager[i]), (r, i)
    graph.release()
    grp.release()


@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.float32, np.complex64])
def test_group_every_tile_class_matches_oracle(oracle, dtype):
    # Tile ids 0..5 forced on every member; the grouped kernel runs 128x64,
    # 64x64, 32x64 and (Float64 only) 32x32 bodies, the others map onto them,
    # and a launch whose members all take 128x64 runs the one-body kernel.
    # K order inside a tile is the same for every class, so all classes give
    # the same bits.
    case(f"group with every tile class forced {np.dtype(dtype).name}")
    rng = np.random.default_rng(21)
    shapes = dims(rng, 10, lo=6, hi=120)
    views, plans, starts = [], [], []
    for (a, ap, b, s) in shapes:
        Cv, L, P = mps_env(rng, dtype, a, ap, b, s)
        views.append(Cv)
        starts.append(storage_np(Cv.parent))
        plans.append(qs.plan_contract(Cv, L, (1, 2, 3), P, (3, 4, 5), (1, 2, 4, 5)))
    cplx = np.dtype(dtype).kind == "c"
    alpha, beta = ((0.75 + 0.5j), (0.5 - 0.25j)) if cplx else (0.75, 0.5)
    want = [oracle_of_plan(oracle, p, alpha, beta, storage_np(p.Astorage), storage_np(p.Bstorage), starts[i])
            for i, p in enumerate(plans)]
    first = None
    for tile in range(6):
        for i, v in enumerate(views):
            v.parent.copy_(torch.from_numpy(starts[i]).to(DEV))
        grp = qs.GroupPlan(plans, tiles=[tile] * len(plans))
        grp.execute(alpha, beta)
        torch.cuda.synchronize()
        got = [storage_np(v.parent) for v in views]
        for i, p in enumerate(plans):
            assert rel_fro(got[i], want[i]) <= RTOL[p.T], (tile, i, shapes[i])
        if first is None:
            first = got
        for i in range(len(plans)):
            assert np.array_equal(got[i], first[i]), (tile, i)
        grp.release()


@pytest.mark.parametrize("cdt,rdt", [(np.complex128, np.float64), (np.complex64, np.float32)])
def test_group_mixed_domain_members_run_grouped(oracle, cdt, rdt):
    # Complex L times real psi and real L times complex psi (C complex, every
    # storage type native) run in grouped launches of the mixed-domain tiles,
    # beside plain complex members: one launch per domain, none standalone.
    # Every tile class forced in turn gives the same bits.
    case(f"group with mixed-domain members {np.dtype(cdt).name} / {np.dtype(rdt).name}")
    rng = np.random.default_rng(31)
    shapes = dims(rng, 12, lo=8, hi=150)

    def rnd(dt, n):
        if np.dtype(dt).kind == "c":
            return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(dt)
        return rng.standard_normal(n).astype(dt)

    views, plans, starts = [], [], []
    for i, (a, ap, b, s) in enumerate(shapes):
        ta, tb = [(cdt, rdt), (rdt, cdt), (cdt, cdt)][i % 3]
        L = SV(torch.from_numpy(rnd(ta, a * 5 * ap)).to(DEV), (a, 5, ap), (1, a, a * 5))
        P = SV(torch.from_numpy(rnd(tb, ap * s * b)).to(DEV), (ap, s, b), (1, ap, ap * s))
        Cv = SV(torch.from_numpy(rnd(cdt, a * 5 * s * b)).to(DEV), (a, 5, s, b), (1, a, a * 5, a * 5 * s))
        views.append(Cv)
        starts.append(storage_np(Cv.parent))
        plans.append(qs.plan_contract(Cv, L, (1, 2, 3), P, (3, 4, 5), (1, 2, 4, 5), conjA=(i % 2 == 1)))
    doms = {p.info["mixed_domain"] for p in plans}
    assert doms == {0, 1, 2}, doms
    alphas = [0.75 + 0.25j * i for i in range(len(plans))]
    betas = [0.5 - 0.5j for _ in plans]
    want = [oracle_of_plan(oracle, p, alphas[i], betas[i], storage_np(p.Astorage), storage_np(p.Bstorage), starts[i])
            for i, p in enumerate(plans)]
    first = None
    for tiles in (None, [2] * len(plans), [1] * len(plans), [4] * len(plans)):
        for i, v in enumerate(views):
            v.parent.copy_(torch.from_numpy(starts[i]).to(DEV))
        grp = qs.GroupPlan(plans, tiles=tiles)
        st = grp.stats()
        assert st == {"launches": 3, "standalone": 0}, st
        grp.execute(alphas, betas)
        torch.cuda.synchronize()
        got = [storage_np(v.parent) for v in views]
        for i, p in enumerate(plans):
            assert rel_fro(got[i], want[i]) <= RTOL[p.T], (tiles, i, shapes[i], p.info["mixed_domain"])
        if first is None:
            first = got
        for i in range(len(plans)):
            assert np.array_equal(got[i], first[i]), (tiles, i)
        grp.release()
    # alone vs in the mix: the same bits
    for i in (0, 1, 2):
        views[i].parent.copy_(torch.from_numpy(starts[i]).to(DEV))
        alone = qs.GroupPlan([plans[i]])
        alone.execute([alphas[i]], [betas[i]])
        torch.cuda.synchronize()
        assert np.array_equal(storage_np(views[i].parent), first[i]), i
        alone.release()


def test_group_mixed_types_run_standalone(oracle):
    case("group with mixed-storage members")
    rng = np.random.default_rng(5)
    plans, views, starts = [], [], []
    for k in range(4):
        M, K, N = 70 + k, 50, 40
        A = SV(torch.from_numpy(rng.standard_normal(M * K).astype(np.float32)).to(DEV), (M, K), (1, M))
        B = SV(torch.from_numpy(rng.standard_normal(K * N)).to(DEV), (K, N), (1, K))
        Cv = SV(torch.from_numpy(rng.standard_normal(M * N)).to(DEV), (M, N), (1, M))
        starts.append(storage_np(Cv.parent))
        views.append(Cv)
        plans.append(qs.plan_contract(Cv, A, (1, 2), B, (2, 3), (1, 3)))
    grp = qs.GroupPlan(plans)
    assert grp.stats() == {"launches": 0, "standalone": 4}
    grp.execute(1.5, 0.5)
    torch.cuda.synchronize()
    for i, p in enumerate(plans):
        want = oracle_of_plan(oracle, p, 1.5, 0.5, storage_np(p.Astorage), storage_np(p.Bstorage), starts[i])
        assert rel_fro(storage_np(views[i].parent), want) <= 1e-13


def test_group_autotune_keeps_results_and_c_and_persists(tmp_path):
    case("group autotune: members' C untouched, results unchanged, classes saved and reloaded")
    rng = np.random.default_rng(17)
    shapes = dims(rng, 20, lo=16, hi=300)
    views, plans = [], []
    for (a, ap, b, s) in shapes:
        Cv, L, P = mps_env(rng, np.float64, a, ap, b, s)
        views.append(Cv)
        plans.append(qs.plan_contract(Cv, L, (1, 2, 3), P, (3, 4, 5), (1, 2, 4, 5)))
    qs.autotune_clear()
    grp = qs.GroupPlan(plans)
    grp.execute(1.0, 0.0)
    torch.cuda.synchronize()
    first = [storage_np(v.parent) for v in views]
    for v in views:
        v.parent.fill_(3.0)
    res = grp.autotune(reps=2)
    torch.cuda.synchronize()
    print("[group tune]", res, flush=True)
    assert res["classes"] >= 1 and res["candidates"] == 4 and res["after_ms"] <= res["before_ms"]
    assert all(torch.all(v.parent == 3.0).item() for v in views)       # tuning writes a scratch C
    grp.execute(1.0, 0.0)
    torch.cuda.synchronize()
    for i, v in enumerate(views):
        assert np.array_equal(storage_np(v.parent), first[i]), i        # same bits with the tuned tiles
    path = str(tmp_path / "tune.txt")
    qs.autotune_save(path)
    text = open(path).read()
    n_g = sum(line.startswith("g:") for line in text.splitlines())
    # the tuned classes are kept only when the whole group ran faster with them
    assert n_g == (res["classes"] if res["after_ms"] < res["before_ms"] else 0), text
    qs.autotune_clear()
    qs.autotune_load(path)
    again = qs.GroupPlan(plans)          # built from the reloaded classes
    for v in views:
        v.parent.fill_(float("nan"))
    again.execute(1.0, 0.0)
    torch.cuda.synchronize()
    for i, v in enumerate(views):
        assert np.array_equal(storage_np(v.parent), first[i]), i
    again.release()
    grp.release()
    qs.autotune_clear()
