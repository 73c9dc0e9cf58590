"""The serving loop on the GPU: each result is the same alone and in the
mix (bitwise), matches the oracle, and the memory budget gates admission."""
import numpy as np
import pytest
import torch

from helpers import RTOL, SV, oracle_of_plan, qs, rel_fro, storage_np

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def case(msg):
    print(f"[case] {msg}", flush=True)


def colmajor(size):
    out, s = [], 1
    for n in size:
        out.append(s)
        s *= n
    return tuple(out)


def rnd(rng, dtype, n):
    if np.dtype(dtype).kind == "c":
        return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(dtype)
    return rng.standard_normal(n).astype(dtype)


def permuted(rng, dtype, d):
    """C[a,b,c,d] := A[a,e,b,f] * B[f,c,e,d], all extents d."""
    A = SV(torch.from_numpy(rnd(rng, dtype, d ** 4)).to(DEV), (d,) * 4, colmajor((d,) * 4))
    B = SV(torch.from_numpy(rnd(rng, dtype, d ** 4)).to(DEV), (d,) * 4, colmajor((d,) * 4))
    Cv = SV(torch.from_numpy(rnd(rng, dtype, d ** 4)).to(DEV), (d,) * 4, colmajor((d,) * 4))
    return Cv, A, B, (1, 2, 3, 4), (4, 5, 2, 6), (1, 3, 5, 6)


def permuted_mixed_domain(rng, d):
    """C[a,b,c,d] := A[a,e,b,f] * B[f,c,e,d], A and C complex, B real."""
    A = SV(torch.from_numpy(rnd(rng, np.complex128, d ** 4)).to(DEV), (d,) * 4, colmajor((d,) * 4))
    B = SV(torch.from_numpy(rnd(rng, np.float64, d ** 4)).to(DEV), (d,) * 4, colmajor((d,) * 4))
    Cv = SV(torch.from_numpy(rnd(rng, np.complex128, d ** 4)).to(DEV), (d,) * 4, colmajor((d,) * 4))
    return Cv, A, B, (1, 2, 3, 4), (4, 5, 2, 6), (1, 3, 5, 6)


def requests(rng, n):
    out = []
    for i in range(n):
        kind = i % 5
        if kind == 0:
            out.append(permuted(rng, np.float64, 8))
        elif kind == 1:
            out.append(permuted(rng, np.complex128, 12))
        elif kind == 2:
            out.append(permuted(rng, np.float32, 16))
        elif kind == 4:
            out.append(permuted_mixed_domain(rng, 10))
        else:
            a, k, b, nn = 24, 200, 20, 150
            A = SV(torch.from_numpy(rnd(rng, np.float64, a * k * b)).to(DEV), (a, k, b), colmajor((a, k, b)))
            B = SV(torch.from_numpy(rnd(rng, np.float64, k * nn)).to(DEV), (k, nn), colmajor((k, nn)))
            Cv = SV(torch.from_numpy(rnd(rng, np.float64, a * nn * b)).to(DEV), (a, nn, b), colmajor((a, nn, b)))
            out.append((Cv, A, B, (1, 2, 3), (2, 4), (1, 4, 3)))
    return out


def test_server_results_same_alone_and_in_mix(oracle):
    case("server: 25 mixed requests (mixed domains among them), alone vs in the mix, grouped vs per-request")
    rng = np.random.default_rng(21)
    reqs = requests(rng, 25)
    plans = [qs.plan_contract(Cv, A, iA, B, iB, iC) for (Cv, A, B, iA, iB, iC) in reqs]
    assert plans[4].info["mixed_domain"] == 1
    starts = [storage_np(r[0].parent) for r in reqs]
    alphas = [1.0 + 0.1 * i for i in range(len(reqs))]
    betas = [0.5 if i % 2 else 0.0 for i in range(len(reqs))]
    srv = qs.Server(budget_bytes=1 << 34, max_batch=16)
    tickets = [srv.submit(p, alphas[i], betas[i]) for i, p in enumerate(plans)]
    lats = [srv.wait(t) for t in tickets]
    assert all(l >= 0 for l in lats)
    torch.cuda.synchronize()
    mixed = [storage_np(r[0].parent) for r in reqs]
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["batches"] >= 1
    for i, p in enumerate(plans):
        want = oracle_of_plan(oracle, p, alphas[i], betas[i], storage_np(p.Astorage), storage_np(p.Bstorage),
                              starts[i])
        assert rel_fro(mixed[i], want) <= RTOL[p.T], (i, p)
    # alone, through the same server
    for i in (0, 1, 2, 3, 4, 9):
        reqs[i][0].parent.copy_(torch.from_numpy(starts[i]).to(DEV))
        srv.wait(srv.submit(plans[i], alphas[i], betas[i]))
        torch.cuda.synchronize()
        assert np.array_equal(storage_np(reqs[i][0].parent), mixed[i]), i
    srv.close()
    # per-request launches (no grouping): the same bits
    solo = qs.Server(budget_bytes=1 << 34, grouped=False)
    for i in (0, 1, 3, 4):
        reqs[i][0].parent.copy_(torch.from_numpy(starts[i]).to(DEV))
        solo.wait(solo.submit(plans[i], alphas[i], betas[i]))
        torch.cuda.synchronize()
        assert np.array_equal(storage_np(reqs[i][0].parent), mixed[i]), i
    solo.close()


def test_server_budget_gates_admission():
    case("server: memory budget admission and rejection")
    rng = np.random.default_rng(22)
    reqs = [permuted(rng, np.float64, 8) for _ in range(10)]
    plans = [qs.plan_contract(Cv, A, iA, B, iB, iC) for (Cv, A, B, iA, iB, iC) in reqs]
    per = 3 * 8 ** 4 * 8
    srv = qs.Server(budget_bytes=per + 100, max_batch=8)   # one request at a time
    ts = [srv.submit(p, 1.0, 0.0) for p in plans]
    for t in ts:
        srv.wait(t)
    st = srv.stats()
    assert st["completed"] == 10 and st["peak_bytes"] <= per + 100 and st["batches"] == 10
    with pytest.raises(MemoryError):
        tiny = qs.Server(budget_bytes=per - 1)
        tiny.submit(plans[0], 1.0, 0.0)
    tiny.close()
    srv.close()


def test_server_other_storages_same_plan(oracle):
    case("server: one plan, many resident operand sets")
    rng = np.random.default_rng(23)
    Cv, A, B, iA, iB, iC = permuted(rng, np.float64, 10)
    plan = qs.plan_contract(Cv, A, iA, B, iB, iC)
    srv = qs.Server(budget_bytes=1 << 30)
    sets = [permuted(rng, np.float64, 10) for _ in range(6)]
    starts = [storage_np(s[0].parent) for s in sets]
    ts = [srv.submit(plan, 1.0, 0.5, A=s[1], B=s[2], Cout=s[0]) for s in sets]
    for t in ts:
        srv.wait(t)
    torch.cuda.synchronize()
    for s, c0 in zip(sets, starts):
        p2 = qs.plan_contract(s[0], s[1], s[3], s[2], s[4], s[5])
        want = oracle_of_plan(oracle, p2, 1.0, 0.5, storage_np(p2.Astorage), storage_np(p2.Bstorage), c0)
        assert rel_fro(storage_np(s[0].parent), want) <= 1e-13
    srv.close()
