"""The benchmarked workloads at their full sizes, checked through properties
that do not depend on size (as the reference checks randomized contractions
against an independent reference, test/execution/test_nest.jl:50-74):

* the grouped sweep of 512 MPS environment contractions (bonds 16..512,
  Float64 and ComplexF64) as a captured HIP graph: every member's checksum
  u^T C v against (u^T A)(B v), bounded by |u||A||B||v|; 8 sampled members
  against the oracle; every replay bitwise equal to the eager group;
* split precision at M = N = K = 4096 with 2^u row / column scaling: the
  sampled Dot2 report at or below 1e-13, the checksum, and the whole result
  against the plain FP64 MFMA result;
* the served stream: large requests (M = N = K = 2048) on the device while
  24 small ones are submitted; each result bitwise equal to its result alone.
"""
import os
import sys

import numpy as np
import pytest
import torch

from helpers import SV, oracle_of_plan, qs, rel_fro, storage_np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402  (the benchmark's own workload generators)

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


def offsets(group, p):
    """Map p's offsets of every coordinate of an axis group (first dim fastest)."""
    off = np.zeros(1, dtype=np.int64)
    for L, s in zip(group.lengths, group.strides[p]):
        off = (off[None, :] + s * np.arange(L, dtype=np.int64)[:, None]).reshape(-1)
    return torch.from_numpy(off).to(DEV)


def mat_of(plan, which):
    """The plan's operand as a matrix over its groups: A[m, k], B[k, n] or
    C[m, n], gathered through the groups' offsets."""
    if which == "A":
        return plan.Astorage[plan.Abase + offsets(plan.mgroup, 0)[:, None] + offsets(plan.kgroup, 0)[None, :]]
    if which == "B":
        return plan.Bstorage[plan.Bbase + offsets(plan.kgroup, 1)[:, None] + offsets(plan.ngroup, 0)[None, :]]
    return plan.Cstorage[plan.Cbase + offsets(plan.mgroup, 1)[:, None] + offsets(plan.ngroup, 1)[None, :]]


def checksum_ok(A, B, Cm, alpha, gen, tol=1e-12):
    """u^T C v == alpha (u^T A)(B v) within tol |u||A||B||v| (on the device, in
    the operands' own precision: float64 / complex128)."""
    dt = Cm.dtype
    u = torch.randn(Cm.shape[0], dtype=torch.float64, device=DEV, generator=gen).to(dt)
    v = torch.randn(Cm.shape[1], dtype=torch.float64, device=DEV, generator=gen).to(dt)
    lhs = u @ (Cm @ v)
    rhs = alpha * ((u @ A) @ (B @ v))
    scale = abs(alpha) * (u.abs() @ (A.abs() @ (B.abs() @ v.abs())))
    return abs(complex(lhs - rhs)) <= tol * float(scale), (complex(lhs), complex(rhs), float(scale))


@pytest.mark.parametrize("cplx", [False, True])
def test_grouped_512_mps_sweep_full_size_graph(oracle, cplx):
    case(f"grouped sweep: 512 MPS env contractions, bonds 16..512, {'ComplexF64' if cplx else 'Float64'}, graph")
    plans, keep, flops = bench.config3(qs, torch, torch.device(DEV), cplx)
    assert len(plans) == 512
    ks = [p.info["k"] for p in plans]
    assert min(ks) >= 16 and max(ks) <= 512
    grp = qs.GroupPlan(plans)
    grp.execute(1.0, 0.0)
    torch.cuda.synchronize()
    eager = [c.clone() for (_, _, c) in keep]
    graph = grp.capture(repeats=1, alphas=1.0, betas=0.0)
    for r in range(3):
        for (_, _, c) in keep:
            c.fill_(float("nan"))
        graph.launch()
        torch.cuda.synchronize()
        x = torch.randn(4096, device=DEV).sum().item()    # eager work between replays
        assert np.isfinite(x)
        bad = [i for i, (_, _, c) in enumerate(keep) if not torch.equal(c, eager[i])]
        assert not bad, (r, bad[:10])
    case("grouped sweep: checksums of all 512 members")
    gen = torch.Generator(device=DEV)
    gen.manual_seed(31)
    for i, p in enumerate(plans):
        ok, info = checksum_ok(mat_of(p, "A"), mat_of(p, "B"), mat_of(p, "C"), 1.0, gen)
        assert ok, (i, p, info)
    case("grouped sweep: 8 sampled members against the oracle")
    rng = np.random.default_rng(32)
    picked = []
    for i in rng.permutation(len(plans)):
        p = plans[int(i)]
        if (8 if cplx else 2) * p.info["m"] * p.info["n"] * p.info["k"] <= 4e8:
            picked.append(int(i))
        if len(picked) == 8:
            break
    for i in picked:
        p = plans[i]
        want = oracle_of_plan(oracle, p, 1.0, 0.0, storage_np(p.Astorage), storage_np(p.Bstorage),
                              np.zeros(p.Cstorage.numel(), dtype=storage_np(p.Cstorage).dtype))
        assert rel_fro(storage_np(p.Cstorage), want) <= 1e-13, (i, p)
    graph.release()
    grp.release()


def test_split_rank4_4096_full_size(oracle):
    case("split precision, rank-4 all 64 (M = N = K = 4096), 2^u scaling, F64: report, checksum, vs plain")
    rng = np.random.default_rng(4)
    d = 64
    ua = rng.integers(-20, 21, size=(d, d))
    ub = rng.integers(-20, 21, size=(d, d))
    A = rng.standard_normal((d, d, d, d)) * np.exp2(ua)[:, None, :, None]   # A[i,m,j,n], per (i, j)
    B = rng.standard_normal((d, d, d, d)) * np.exp2(ub)[None, :, None, :]   # B[m,k,n,l], per (k, l)
    st = colmajor((d,) * 4)
    Av = SV(torch.from_numpy(A.reshape(-1, order="F")).to(DEV), (d,) * 4, st)
    Bv = SV(torch.from_numpy(B.reshape(-1, order="F")).to(DEV), (d,) * 4, st)
    Cs = SV(torch.full((d ** 4,), float("nan"), dtype=torch.float64, device=DEV), (d,) * 4, st)
    Cp = SV(torch.full((d ** 4,), float("nan"), dtype=torch.float64, device=DEV), (d,) * 4, st)
    split = qs.plan_contract(Cs, Av, (1, 5, 2, 6), Bv, (5, 3, 6, 4), (1, 2, 3, 4), precision="split")
    plain = qs.plan_contract(Cp, Av, (1, 5, 2, 6), Bv, (5, 3, 6, 4), (1, 2, 3, 4))
    assert split.path == "split" and plain.path == "gemm"
    qs.set_error_report(split, True)
    qs.execute(split, 1.0, 0.0)
    qs.execute(plain, 1.0, 0.0)
    torch.cuda.synchronize()
    rep = qs.error_report(split)
    print(f"[split4096] moduli={split.info['moduli']} report={rep}", flush=True)
    assert rep["samples"] == 256 and rep["rel_fro"] <= 1e-13, rep
    assert torch.isfinite(Cs.parent).all().item()
    diff = float(torch.linalg.vector_norm(Cs.parent - Cp.parent) / torch.linalg.vector_norm(Cp.parent))
    print(f"[split4096] rel. Frobenius vs plain FP64 MFMA: {diff:.3e}", flush=True)
    assert diff <= 1e-13, diff
    # checksum on the host: C as a (i,j) x (k,l) matrix
    Cm = storage_np(Cs.parent).reshape((d,) * 4, order="F").reshape(d * d, d * d, order="F")
    Am = A.transpose(0, 2, 1, 3).reshape(d * d, d * d, order="F")
    Bm = B.transpose(0, 2, 1, 3).reshape(d * d, d * d, order="F")
    for s in range(3):
        r = np.random.default_rng(200 + s)
        u = r.standard_normal(d * d)
        v = r.standard_normal(d * d)
        lhs = u @ Cm @ v
        rhs = (u @ Am) @ (Bm @ v)
        scale = np.abs(u) @ np.abs(Am) @ np.abs(Bm) @ np.abs(v)
        assert abs(lhs - rhs) <= 1e-12 * scale, (s, lhs, rhs, scale)
    split.release()
    plain.release()


def large_request(gen, dt, n=2048):
    """Config-1 pattern C[a,n,b] := A[a,k,b] * B[k,n], M = N = K = n."""
    a = 32
    b = n // a
    A = torch.randn(a * n * b, dtype=dt, device=DEV, generator=gen)
    B = torch.randn(n * n, dtype=dt, device=DEV, generator=gen)
    Cst = torch.randn(a * n * b, dtype=dt, device=DEV, generator=gen)
    Cv = SV(Cst, (a, n, b), colmajor((a, n, b)))
    plan = qs.plan_contract(Cv, SV(A, (a, n, b), colmajor((a, n, b))), (1, 2, 3), SV(B, (n, n), (1, n)), (2, 4),
                            (1, 4, 3))
    return plan, Cv


def small_request(gen, dt, d):
    """Config-2 pattern C[a,b,c,d] := A[a,e,b,f] * B[f,c,e,d], all extents d."""
    st = colmajor((d,) * 4)
    A = SV(torch.randn(d ** 4, dtype=dt, device=DEV, generator=gen), (d,) * 4, st)
    B = SV(torch.randn(d ** 4, dtype=dt, device=DEV, generator=gen), (d,) * 4, st)
    Cv = SV(torch.randn(d ** 4, dtype=dt, device=DEV, generator=gen), (d,) * 4, st)
    return qs.plan_contract(Cv, A, (1, 2, 3, 4), B, (4, 5, 2, 6), (1, 3, 5, 6)), Cv


def test_served_large_requests_in_flight_while_small_ones_arrive(oracle):
    case("served: two M=N=K=2048 requests (F64, C64) on the device while 24 small ones are submitted")
    gen = torch.Generator(device=DEV)
    gen.manual_seed(55)
    reqs = [large_request(gen, torch.complex128), large_request(gen, torch.float64)]
    for i in range(24):
        dt = torch.float64 if i % 2 == 0 else torch.complex128
        reqs.append(small_request(gen, dt, (8, 12, 16)[i % 3]))
    starts = [c.parent.clone() for (_, c) in reqs]
    alphas = [1.0 + 0.05 * i for i in range(len(reqs))]
    betas = [0.5 if i % 3 else 0.0 for i in range(len(reqs))]
    srv = qs.Server(budget_bytes=200 * 10 ** 9, max_batch=64)
    torch.cuda.synchronize()
    tickets = [srv.submit(reqs[0][0], alphas[0], betas[0]), srv.submit(reqs[1][0], alphas[1], betas[1])]
    for i in range(2, len(reqs)):
        tickets.append(srv.submit(reqs[i][0], alphas[i], betas[i]))
    large_done_at_submit_end = [srv.poll(t) is not None for t in tickets[:2]]
    lats = [srv.wait(t) for t in tickets]
    torch.cuda.synchronize()
    st = srv.stats()
    print(f"[served] large done when the last small one was submitted: {large_done_at_submit_end}; "
          f"latencies ms large {lats[:2]} small p50 {np.median(lats[2:]):.3f}; batches {st['batches']}", flush=True)
    assert st["completed"] == len(reqs)
    mixed = [c.parent.clone() for (_, c) in reqs]
    # each alone, through a fresh server, from the same starting C
    solo = qs.Server(budget_bytes=200 * 10 ** 9, max_batch=64)
    for i, (plan, Cv) in enumerate(reqs):
        Cv.parent.copy_(starts[i])
        torch.cuda.synchronize()
        solo.wait(solo.submit(plan, alphas[i], betas[i]))
        torch.cuda.synchronize()
        assert torch.equal(Cv.parent, mixed[i]), i
    solo.close()
    srv.close()
    # the results themselves: large ones by checksum, small ones against the oracle
    g2 = torch.Generator(device=DEV)
    g2.manual_seed(56)
    for i in (0, 1):
        plan, Cv = reqs[i]
        Am, Bm, Cm = mat_of(plan, "A"), mat_of(plan, "B"), mat_of(plan, "C")
        C0 = starts[i][plan.Cbase + offsets(plan.mgroup, 1)[:, None] + offsets(plan.ngroup, 1)[None, :]]
        ok, info = checksum_ok(Am, Bm, Cm - betas[i] * C0, alphas[i], g2)
        assert ok, (i, info)
    for i in range(2, len(reqs), 5):
        plan, Cv = reqs[i]
        want = oracle_of_plan(oracle, plan, alphas[i], betas[i], storage_np(plan.Astorage), storage_np(plan.Bstorage),
                              starts[i].cpu().numpy())
        assert rel_fro(storage_np(Cv.parent), want) <= 1e-13, i
