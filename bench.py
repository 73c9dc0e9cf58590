"""Benchmark of the block-scatter contraction on one MI355X.

Headline workload (one "step" = one `execute` of the plan, operands resident
in HBM): the permuted strided contraction
    C[a,b,c,d] := A[a,e,b,f] * conj(B[f,c,e,d]),  all extents 32 (M = N = K = 1024),
Float64, A a stride-2 slice of a 64-wide parent along b, beta = 0.5 with C
prefilled N(0,1).  Metric: GFLOP/s (2 flops per real MAC), plus the dominant
kernel's fraction of the FP64 MFMA peak and a single-thread CPU baseline (the
oracle's BLIS restatement of the reference's SIMDKernel nest).

Extra lines in `extra` cover the other plain shapes (rank-4 all-64 M=N=K=4096,
the 48/96 shape, the ComplexF64 variant).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-extra] [--no-cpu]
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PEAK_FP64 = 78.6e12
PEAK_FP32 = 157.3e12
PEAK_HBM = 8.0e12


def _load():
    pkg = os.path.join(ROOT, "quasistrided.jl_amd")
    spec = importlib.util.spec_from_file_location("quasistrided_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["quasistrided_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def colmajor(size):
    out, s = [], 1
    for n in size:
        out.append(s)
        s *= n
    return tuple(out)


def randn(rng, dtype, n):
    if np.dtype(dtype).kind == "c":
        return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(dtype)
    return rng.standard_normal(n).astype(dtype)


def config2(qs, torch, dtype, dev):
    """C[a,b,c,d] := A[a,e,b,f] * conj(B[f,c,e,d]), all 32; A = P[:, :, 1:2:64, :]."""
    rng = np.random.default_rng(2)
    d = 32
    P = randn(rng, dtype, d * d * 2 * d * d)
    pst = colmajor((d, d, 2 * d, d))
    SV = qs.StridedView
    A = SV(torch.from_numpy(P).to(dev), (d,) * 4, (pst[0], pst[1], 2 * pst[2], pst[3]))
    B = SV(torch.from_numpy(randn(rng, dtype, d ** 4)).to(dev), (d,) * 4, colmajor((d,) * 4))
    C = SV(torch.from_numpy(randn(rng, dtype, d ** 4)).to(dev), (d,) * 4, colmajor((d,) * 4))
    plan = qs.plan_contract(C, A, (1, 2, 3, 4), B, (4, 5, 2, 6), (1, 3, 5, 6), conjB=True)
    cplx = np.dtype(dtype).kind == "c"
    flops = (8 if cplx else 2) * 1024 ** 3
    # compulsory bytes: A (only the selected planes), B, C read (beta != 0) and written
    es = np.dtype(dtype).itemsize
    bytes_ = 4 * (d ** 4) * es
    return plan, flops, bytes_, (P, A, B, C), 0.5


def config4(qs, torch, dev):
    """C[i,j,k,l] := A[i,m,j,n] * B[m,k,n,l], all 64 (M = N = K = 4096), F64."""
    rng = np.random.default_rng(4)
    d = 64
    SV = qs.StridedView
    A = SV(torch.from_numpy(rng.standard_normal(d ** 4)).to(dev), (d,) * 4, colmajor((d,) * 4))
    B = SV(torch.from_numpy(rng.standard_normal(d ** 4)).to(dev), (d,) * 4, colmajor((d,) * 4))
    C = SV(torch.zeros(d ** 4, dtype=torch.float64, device=dev), (d,) * 4, colmajor((d,) * 4))
    plan = qs.plan_contract(C, A, (1, 5, 2, 6), B, (5, 3, 6, 4), (1, 2, 3, 4))
    return plan, 2 * 4096 ** 3, 3 * 8 * 4096 ** 2, (A, B, C), 0.0


def bench_split(qs, torch, dev, d=64, reps=10, dtype=None):
    """Rank-4 all-64 with entries N(0,1) * 2^u (u integer in [-20, 20] per
    row of A and per column of B), Float64 (or Float32): autotuned plain MFMA
    of the type vs split precision (INT8 residue GEMMs; target 1e-13 for
    Float64, 1e-6 for Float32), both timed per execute, sustained and
    interleaved; the split error from the sampled Dot2 report."""
    dtype = dtype or torch.float64
    rng = np.random.default_rng(4)
    SV = qs.StridedView
    ua = rng.integers(-20, 21, size=(d, d))
    ub = rng.integers(-20, 21, size=(d, d))
    g = torch.Generator(device=dev)
    g.manual_seed(4)
    A = torch.randn((d, d, d, d), dtype=torch.float64, device=dev, generator=g)
    B = torch.randn((d, d, d, d), dtype=torch.float64, device=dev, generator=g)
    # column-major A[i,m,j,n]: storage index i + d m + d^2 j + d^3 n -> torch dims (n, j, m, i)
    A *= torch.from_numpy(np.exp2(ua)).to(dev).t()[None, :, None, :]
    B *= torch.from_numpy(np.exp2(ub)).to(dev).t()[:, None, :, None]
    A, B = A.reshape(-1).to(dtype), B.reshape(-1).to(dtype)
    C1 = torch.zeros(d ** 4, dtype=dtype, device=dev)
    C2 = torch.zeros(d ** 4, dtype=dtype, device=dev)
    st = colmajor((d,) * 4)
    mk = lambda Cst, prec: qs.plan_contract(SV(Cst, (d,) * 4, st), SV(A, (d,) * 4, st), (1, 5, 2, 6),
                                            SV(B, (d,) * 4, st), (5, 3, 6, 4), (1, 2, 3, 4), precision=prec)
    plain = mk(C1, "plain")
    qs.autotune(plain, reps=3)   # the plain side at its best tile, as in the rank-4 bench
    split = mk(C2, "split")
    flops = 2 * d ** 6
    out = {}
    # Sustained: each path back to back after about 50 ms of its own warm-up
    # (long INT8 and FP64 MFMA runs move the clocks by 10-20 %, and the INT8
    # GEMMs draw the most power).  Interleaved: executes alternate plain /
    # split, both under the same mixed load.
    for name, p, w in (("plain", plain, 20), ("split", split, 40)):
        _, km, kmed = time_plan(qs, torch, p, 0.0, reps, w)
        out[name + "_ms"] = round(km * 1e3, 4)
        out[name + "_ms_median"] = round(kmed * 1e3, 4)
        out[name + ("_tflops_fp64_equiv" if dtype == torch.float64 else "_tflops_fp32_equiv")] = round(flops / km / 1e12, 2)
    times = time_plans_alternating(qs, torch, (plain, split), 0.0, reps, 2)
    for name, (km, _) in zip(("plain", "split"), times):
        out[name + "_ms_interleaved"] = round(km * 1e3, 4)
    qs.set_error_report(split, True)
    qs.execute(split, 1.0, 0.0)
    torch.cuda.synchronize()
    rep = qs.error_report(split)
    diff = (C2.double() - C1.double()).norm().item() / C1.double().norm().item()
    out.update({"moduli": rep["moduli"], "split_vs_plain_speedup": round(out["plain_ms"] / out["split_ms"], 3),
                "split_vs_plain_speedup_interleaved": round(out["plain_ms_interleaved"] / out["split_ms_interleaved"], 3),
                "split_rel_fro_err_sampled_dot2": rep["rel_fro"], "split_max_rel_err_sampled": rep["max_rel"],
                "split_vs_plain_rel_fro": diff,
                "int8_pops_equiv": round(flops * rep["moduli"] / (out["split_ms"] * 1e-3)
                                         / 1e15, 3)})
    return out


def bench_fp32(qs, torch, dev):
    """Float32 on the FP32 MFMA (peak 157.3 TFLOP/s): the permuted all-32
    workload and the rank-4 all-64 contraction, autotuned, launch period
    from events at both ends of the timed region."""
    out = {}
    plan, flops, _, views, beta = config2(qs, torch, np.float32, dev)
    qs.autotune(plan, reps=5)
    _, period, _, _ = time_plan_region(qs, torch, plan, beta, 100, 10)
    out["permuted_all32_f32"] = {"period_us": round(period * 1e6, 2), "tflops": round(flops / period / 1e12, 2),
                                 "mfma_frac": round(flops / period / PEAK_FP32, 4), "plan": repr(plan)}
    del plan, views
    rng = np.random.default_rng(4)
    d = 64
    SV = qs.StridedView
    mk = lambda: SV(torch.from_numpy(rng.standard_normal(d ** 4).astype(np.float32)).to(dev), (d,) * 4,
                    colmajor((d,) * 4))
    A, B = mk(), mk()
    C = SV(torch.zeros(d ** 4, dtype=torch.float32, device=dev), (d,) * 4, colmajor((d,) * 4))
    plan = qs.plan_contract(C, A, (1, 5, 2, 6), B, (5, 3, 6, 4), (1, 2, 3, 4))
    qs.autotune(plan, reps=3)
    _, period, _, _ = time_plan_region(qs, torch, plan, 0.0, 10, 3)
    fl = 2 * 4096 ** 3
    out["rank4_all64_f32"] = {"period_ms": round(period * 1e3, 4), "tflops": round(fl / period / 1e12, 2),
                              "mfma_frac": round(fl / period / PEAK_FP32, 4), "plan": repr(plan)}
    return out


def bench_mixed_domain(qs, torch, dev, reps=50):
    """Complex A times real B (ComplexRealKernel's case) and real A times
    complex B, on the permuted all-32 workload (ComplexF64 / Float64 storages,
    ComplexF64 C), against the all-complex contraction of the same shapes.
    Useful flops: 4 M N K for a complex-real product (8 M N K for complex)."""
    out = {}
    rng = np.random.default_rng(9)
    d = 32
    SV = qs.StridedView
    pst = colmajor((d, d, 2 * d, d))
    ast = (pst[0], pst[1], 2 * pst[2], pst[3])
    for name, ta, tb in (("c64_x_f64", np.complex128, np.float64), ("f64_x_c64", np.float64, np.complex128),
                         ("c64_x_c64", np.complex128, np.complex128)):
        A = SV(torch.from_numpy(randn(rng, ta, d * d * 2 * d * d)).to(dev), (d,) * 4, ast)
        B = SV(torch.from_numpy(randn(rng, tb, d ** 4)).to(dev), (d,) * 4, colmajor((d,) * 4))
        C = SV(torch.from_numpy(randn(rng, np.complex128, d ** 4)).to(dev), (d,) * 4, colmajor((d,) * 4))
        plan = qs.plan_contract(C, A, (1, 2, 3, 4), B, (4, 5, 2, 6), (1, 3, 5, 6))
        qs.autotune(plan, reps=3)
        _, km, kmed = time_plan(qs, torch, plan, 0.5, reps, 5)
        fl = (8 if name == "c64_x_c64" else 4) * 1024 ** 3
        out[name] = {"kernel_us": round(kmed * 1e6, 1), "useful_tflops": round(fl / kmed / 1e12, 2),
                     "mfma_frac": round(fl / kmed / PEAK_FP64, 4), "mixed_domain": plan.info["mixed_domain"],
                     "plan": repr(plan)}
        plan.release()
    out["c64_x_f64_vs_c64_x_c64_time"] = round(out["c64_x_c64"]["kernel_us"] / out["c64_x_f64"]["kernel_us"], 3)
    return out


def bench_batched(qs, torch, dev, reps=20):
    """Batch labels: C[i,j,b] := A[i,k,b] * B[k,j,b], F64, column-major, as one
    grouped launch (and its HIP graph), against torch.bmm on the same data
    (the vendor's strided-batched FP64 GEMM, operands laid out batch-major
    for it).  Two shapes: 2048 batches of 64^3 and 256 of 256^3."""
    SV = qs.StridedView
    g = torch.Generator(device=dev)
    g.manual_seed(11)
    out = {}
    for (n, b) in ((64, 2048), (256, 256)):
        i = j = k = n
        A = torch.randn(i * k * b, dtype=torch.float64, device=dev, generator=g)
        B = torch.randn(k * j * b, dtype=torch.float64, device=dev, generator=g)
        Cst = torch.zeros(i * j * b, dtype=torch.float64, device=dev)
        cm = lambda sz: tuple(int(np.prod(sz[:d])) for d in range(len(sz)))
        bp = qs.plan_batched(SV(Cst, (i, j, b), cm((i, j, b))), SV(A, (i, k, b), cm((i, k, b))), (1, 2, 3),
                             SV(B, (k, j, b), cm((k, j, b))), (2, 4, 3), (1, 4, 3))
        flops = 2 * i * j * k * b
        tune = bp.autotune(reps=5)
        graph = bp.capture(repeats=1, alpha=1.0, beta=0.0)
        graph.launch()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            graph.launch()
        e1.record()
        torch.cuda.synchronize()
        t_graph = e0.elapsed_time(e1) * 1e-3 / reps
        # torch.bmm over [b, i, k] x [b, k, j] (the column-major slices, transposed views made contiguous)
        At = A.view(b, k, i).transpose(1, 2).contiguous()
        Bt = B.view(b, j, k).transpose(1, 2).contiguous()
        torch.bmm(At, Bt)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            Ct = torch.bmm(At, Bt)
        e1.record()
        torch.cuda.synchronize()
        t_bmm = e0.elapsed_time(e1) * 1e-3 / reps
        err = (Ct.transpose(1, 2).reshape(-1) - Cst).norm().item() / Ct.norm().item()
        out[f"{b}x{n}^3"] = {"graph_us": round(t_graph * 1e6, 1), "graph_tflops": round(flops / t_graph / 1e12, 2),
                             "graph_mfma_frac": round(flops / t_graph / PEAK_FP64, 4),
                             "torch_bmm_us": round(t_bmm * 1e6, 1), "torch_bmm_tflops": round(flops / t_bmm / 1e12, 2),
                             "vs_torch_bmm": round(t_bmm / t_graph, 3), "rel_fro_vs_bmm": err,
                             "tile": bp.tile, "tile_us": {str(t): round(v * 1e6, 1) for t, v in tune.items()}}
        graph.release()
        bp.release()
        del A, B, Cst, At, Bt, Ct
    return out


def bench_bandwidth(qs, torch, dev, reps=20):
    """The HBM-bound paths at sizes well past the 4 MB L2 of an XCD (and the
    MALL for the large operand), F64, operands N(0,1): algorithmic bytes per
    execute over the HIP-event kernel time on the launch stream, against
    8 TB/s.
      dot_kfast   C[m] := A[m,k] * b[k], A row-major (K contiguous), M=65536, K=2048
      dot_mfast   the same with A column-major (M contiguous)
      outer       C[m,n] := a[m] * b[n], M=N=8192, beta=0 (C written, never read)
      skinny      C[m,n] := A[m,k] * B[k,n], M=16, K=4096, N=32768, B column-major
                  (read in place, streamed once)
      transpose_add  tensoradd! C = permutedims(A, (2,1)), 8192 x 8192, beta=0
                  (timed per call; the layout's add plan comes from the cache)"""
    SV = qs.StridedView
    g = torch.Generator(device=dev)
    g.manual_seed(9)
    rnd = lambda n: torch.randn(n, dtype=torch.float64, device=dev, generator=g)
    out = {}

    def run(name, plan, nbytes, flops, beta=0.0):
        _, km, _ = time_plan(qs, torch, plan, beta, reps, 3)
        out[name] = {"path": plan.path, "kernel_us": round(km * 1e6, 2), "gb_per_s": round(nbytes / km / 1e9, 1),
                     "hbm_frac": round(nbytes / km / PEAK_HBM, 4), "bytes": nbytes,
                     "gflops": round(flops / km / 1e9, 1)}

    M, K = 65536, 2048
    A, b, c = rnd(M * K), rnd(K), torch.zeros(M, dtype=torch.float64, device=dev)
    for name, strides in (("dot_kfast", (K, 1)), ("dot_mfast", (1, M))):
        plan = qs.plan_contract(SV(c, (M,), (1,)), SV(A, (M, K), strides), (1, 2), SV(b, (K,), (1,)), (2,), (1,))
        run(name, plan, 8 * (M * K + K + M), 2 * M * K)
    del A, b, c
    M = N = 8192
    a, b, C = rnd(M), rnd(N), torch.zeros(M * N, dtype=torch.float64, device=dev)
    plan = qs.plan_contract(SV(C, (M, N), (1, M)), SV(a, (M,), (1,)), (1,), SV(b, (N,), (1,)), (2,), (1, 2))
    run("outer", plan, 8 * (M + N + M * N), 2 * M * N)
    del a, b, C
    M, K, N = 16, 4096, 32768
    A, B, C = rnd(M * K), rnd(K * N), torch.zeros(M * N, dtype=torch.float64, device=dev)
    plan = qs.plan_contract(SV(C, (M, N), (1, M)), SV(A, (M, K), (1, M)), (1, 2), SV(B, (K, N), (1, K)), (2, 3),
                            (1, 3))
    qs.autotune(plan, reps=3)
    run("skinny", plan, 8 * (M * K + K * N + M * N), 2 * M * K * N)
    del A, B, C
    # tensoradd! transposing 8192 x 8192 (beta = 0: A read once, C written once).
    n = 8192
    A, Cst = rnd(n * n), torch.empty(n * n, dtype=torch.float64, device=dev)
    Av, Cv = SV(A, (n, n), (1, n)), SV(Cst, (n, n), (1, n))
    for _ in range(3):
        qs.tensoradd(Cv, Av, ((2, 1), ()), False, 1.0, 0.0)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for e0, e1 in ev:
        e0.record()
        qs.tensoradd(Cv, Av, ((2, 1), ()), False, 1.0, 0.0)
        e1.record()
    torch.cuda.synchronize()
    t = float(np.mean([e0.elapsed_time(e1) for e0, e1 in ev])) * 1e-3
    nbytes = 2 * 8 * n * n
    out["transpose_add"] = {"path": "tensoradd", "call_us": round(t * 1e6, 2), "gb_per_s": round(nbytes / t / 1e9, 1),
                            "hbm_frac": round(nbytes / t / PEAK_HBM, 4), "bytes": nbytes,
                            "note": "per tensoradd call: plan-cache lookup, storage rebind, span checks and launch"}
    del A, Cst
    return out


def config1(qs, torch, dev):
    """C[a,n,b] := A[a,k,b] * B[k,n], a=b=48, k=n=96, F64."""
    rng = np.random.default_rng(1)
    a, k, b, n = 48, 96, 48, 96
    SV = qs.StridedView
    A = SV(torch.from_numpy(rng.standard_normal(a * k * b)).to(dev), (a, k, b), colmajor((a, k, b)))
    B = SV(torch.from_numpy(rng.standard_normal(k * n)).to(dev), (k, n), colmajor((k, n)))
    C = SV(torch.zeros(a * n * b, dtype=torch.float64, device=dev), (a, n, b), colmajor((a, n, b)))
    plan = qs.plan_contract(C, A, (1, 2, 3), B, (2, 4), (1, 4, 3))
    return plan, 2 * 2304 * 96 * 96, 8 * (2304 * 96 + 96 * 96 + 2304 * 96), (A, B, C), 0.0


def config3(qs, torch, dev, cplx, n=512, seed=3):
    """512 MPS environment contractions L[a,w,a'] * psi[a',s,b] -> C[a,w,s,b],
    w = 5, s uniform in {2,3,4}, a, a', b log-uniform in [16, 512]."""
    rng = np.random.default_rng(seed)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    dt = torch.complex128 if cplx else torch.float64
    SV = qs.StridedView
    plans, keep, flops = [], [], 0
    for _ in range(n):
        a, ap, b = (int(round(np.exp(rng.uniform(np.log(16), np.log(512))))) for _ in range(3))
        s = int(rng.integers(2, 5))
        w = 5
        L = torch.randn(a * w * ap, dtype=dt, device=dev, generator=g)
        P = torch.randn(ap * s * b, dtype=dt, device=dev, generator=g)
        Cst = torch.zeros(a * w * s * b, dtype=dt, device=dev)
        Lv = SV(L, (a, w, ap), (1, a, a * w))
        Pv = SV(P, (ap, s, b), (1, ap, ap * s))
        Cv = SV(Cst, (a, w, s, b), (1, a, a * w, a * w * s))
        plans.append(qs.plan_contract(Cv, Lv, (1, 2, 3), Pv, (3, 4, 5), (1, 2, 4, 5)))
        keep.append((L, P, Cst))
        flops += (8 if cplx else 2) * a * w * ap * s * b
    return plans, keep, flops


def bench_grouped(qs, torch, dev, cplx, reps=5):
    """Config-3 sweep: per-launch plain execution of the 512 plans vs one
    grouped persistent launch vs its captured HIP graph."""
    plans, keep, flops = config3(qs, torch, dev, cplx)
    n = len(plans)
    for p in plans:
        qs.execute(p, 1.0, 0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for p in plans:
            qs.execute(p, 1.0, 0.0)
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t0) / reps
    grp = qs.GroupPlan(plans)
    tune = grp.autotune(reps=3)
    grp.execute(1.0, 0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        grp.execute(1.0, 0.0)
    torch.cuda.synchronize()
    t_group = (time.perf_counter() - t0) / reps
    graph = grp.capture(repeats=1, alphas=1.0, betas=0.0)
    # Replays settle over the first few (the clocks move with the load):
    # three untimed, then 2 * reps timed back to back.
    for _ in range(3):
        graph.launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(2 * reps):
        graph.launch()
    e1.record()
    torch.cuda.synchronize()
    t_graph = e0.elapsed_time(e1) * 1e-3 / (2 * reps)
    # The batch's roofline: each member at min(FP64 MFMA peak, its bytes at
    # 8 TB/s) -- sum over members of max(flops / peak, bytes / bandwidth),
    # bytes = A + B + C (write only, beta = 0).
    esz = 16 if cplx else 8
    roof = sum(max((8 if cplx else 2) * p.info["m"] * p.info["n"] * p.info["k"] / PEAK_FP64,
                   (p.info["m"] * p.info["k"] + p.info["k"] * p.info["n"] + p.info["m"] * p.info["n"]) * esz / PEAK_HBM)
               for p in plans)
    out = {
        "contractions": n,
        "gflop_per_sweep": round(flops / 1e9, 2),
        "roofline_ms": round(roof * 1e3, 4),
        "graph_roofline_frac": round(roof / t_graph, 4),
        "per_launch_contractions_per_s": round(n / t_plain, 1),
        "grouped_contractions_per_s": round(n / t_group, 1),
        "graph_contractions_per_s": round(n / t_graph, 1),
        "graph_vs_per_launch": round(t_plain / t_graph, 2),
        "graph_tflops": round(flops / t_graph / 1e12, 2),
        "graph_mfma_frac": round(flops / t_graph / PEAK_FP64, 4),
        "autotune": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in tune.items()},
    }
    graph.release()
    grp.release()
    del plans, keep
    return out


def served_pools(qs, torch, dev):
    """Resident operands and plans of the served stream: small permuted
    C[a,b,c,d] := A[a,e,b,f] * B[f,c,e,d] (all extents d in {8, 12, 16}) and
    large C[a,n,b] := A[a,k,b] * B[k,n] with M = N = K in {1024, 2048, 3072,
    4096}; each in Float64 and ComplexF64.  Every shape has one plan and a
    ring of output buffers."""
    SV = qs.StridedView
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    pools = {"small": [], "large": []}
    for dt in (torch.float64, torch.complex128):
        for d in (8, 12, 16):
            A = torch.randn(d ** 4, dtype=dt, device=dev, generator=g)
            B = torch.randn(d ** 4, dtype=dt, device=dev, generator=g)
            Cs = [torch.zeros(d ** 4, dtype=dt, device=dev) for _ in range(64)]
            st = colmajor((d,) * 4)
            plan = qs.plan_contract(SV(Cs[0], (d,) * 4, st), SV(A, (d,) * 4, st), (1, 2, 3, 4), SV(B, (d,) * 4, st),
                                    (4, 5, 2, 6), (1, 3, 5, 6))
            fl = (8 if dt.is_complex else 2) * d ** 6
            pools["small"].append((plan, [SV(c, (d,) * 4, st) for c in Cs], fl))
        for n in (1024, 2048, 3072, 4096):
            a = b = 32 if n <= 2048 else 64
            a = int(np.sqrt(n)) if int(np.sqrt(n)) ** 2 == n else a
            b = n // a
            A = torch.randn(a * n * b, dtype=dt, device=dev, generator=g)
            B = torch.randn(n * n, dtype=dt, device=dev, generator=g)
            Cs = [torch.zeros(a * n * b, dtype=dt, device=dev) for _ in range(4)]
            plan = qs.plan_contract(SV(Cs[0], (a, n, b), colmajor((a, n, b))), SV(A, (a, n, b), colmajor((a, n, b))),
                                    (1, 2, 3), SV(B, (n, n), (1, n)), (2, 4), (1, 4, 3))
            fl = (8 if dt.is_complex else 2) * n ** 3
            pools["large"].append((plan, [SV(c, (a, n, b), colmajor((a, n, b))) for c in Cs], fl))
    return pools


def bench_served(qs, torch, dev, pools, grouped, rate=2000.0, duration=2.0, seed=5):
    """Poisson arrivals at `rate` req/s for `duration` s: 90 % small, 10 %
    large, F64 and C64; 200 GB admission budget.  Returns latency and
    throughput figures."""
    rng = np.random.default_rng(seed)
    n = int(rate * duration)
    gaps = rng.exponential(1.0 / rate, n)
    arrivals = np.cumsum(gaps)
    kinds = rng.random(n) < 0.9
    srv = qs.Server(budget_bytes=200 * 10 ** 9, max_batch=128, grouped=grouped)
    ring = {}
    tickets, flops = [], 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lag = []
    for i in range(n):
        pool = pools["small"] if kinds[i] else pools["large"]
        plan, outs, fl = pool[int(rng.integers(len(pool)))]
        k = ring.get(id(plan), 0)
        ring[id(plan)] = k + 1
        target = t0 + arrivals[i]
        while time.perf_counter() < target:
            pass
        lag.append(time.perf_counter() - target)
        tickets.append(srv.submit(plan, 1.0, 0.0, Cout=outs[k % len(outs)]))
        flops += fl
    srv.drain()
    st = srv.stats()
    for t in tickets:
        srv.poll(t)
    srv.close()
    return {
        "requests": n, "offered_req_per_s": rate, "mode": "grouped" if grouped else "per-request launch",
        "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3), "max_ms": round(st["max_ms"], 3),
        "throughput_req_per_s": round(st["throughput_rps"], 1), "tflops": round(st["tflops"], 2),
        "mean_batch": round(st["mean_batch"], 2), "peak_gb": round(st["peak_bytes"] / 1e9, 3),
        "submit_lag_p99_ms": round(float(np.percentile(lag, 99)) * 1e3, 3),
    }


def bench_saturated(qs, torch, dev, pools, grouped, n=3000, seed=6, small_share=0.9):
    """All requests submitted at once (closed loop): the loop's capacity."""
    rng = np.random.default_rng(seed)
    kinds = rng.random(n) < small_share
    srv = qs.Server(budget_bytes=200 * 10 ** 9, max_batch=256, grouped=grouped)
    ring = {}
    torch.cuda.synchronize()
    tickets = []
    for i in range(n):
        pool = pools["small"] if kinds[i] else pools["large"]
        plan, outs, fl = pool[int(rng.integers(len(pool)))]
        k = ring.get(id(plan), 0)
        ring[id(plan)] = k + 1
        tickets.append(srv.submit(plan, 1.0, 0.0, Cout=outs[k % len(outs)]))
    srv.drain()
    st = srv.stats()
    for t in tickets:
        srv.poll(t)
    srv.close()
    return {"mode": "grouped" if grouped else "per-request launch", "requests": n,
            "throughput_req_per_s": round(st["throughput_rps"], 1), "tflops": round(st["tflops"], 2),
            "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3), "mean_batch": round(st["mean_batch"], 2)}


def bench_plain_extra(qs, torch, dev, key):
    """The other plain shapes: rank-4 all-64 (M = N = K = 4096), the 48/96
    C[a,n,b] shape and the ComplexF64 permuted all-32 one, each autotuned."""
    maker = {"rank4_f64": lambda: config4(qs, torch, dev), "c_anb_f64": lambda: config1(qs, torch, dev),
             "permuted_c64": lambda: config2(qs, torch, np.complex128, dev)}[key]
    p, fl, by, vw, be = maker()
    qs.autotune(p, reps=3)
    steps = 20 if key.startswith("rank4") else 100
    w, km, _ = time_plan(qs, torch, p, be, steps, 3)
    out = {"gflops": round(fl * steps / w / 1e9, 1), "kernel_ms": round(km * 1e3, 4),
           "mfma_frac": round(fl / km / PEAK_FP64, 4), "plan": repr(p)}
    del p, vw
    return out


def time_plan(qs, torch, plan, beta, steps, warmup):
    """Warm up, then time `steps` executes back to back (wall clock between
    synchronizations) and each execute with HIP events on the launch stream."""
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        qs.execute(plan, 1.0, beta)
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        starts[i].record(stream)
        qs.execute(plan, 1.0, beta)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [s.elapsed_time(e) * 1e-3 for s, e in zip(starts, ends)]
    return wall, float(np.mean(per)), float(np.median(per))


def time_plan_region(qs, torch, plan, beta, steps, warmup):
    """The headline's timing.  Warm up; then the timed region: `steps`
    executes back to back between two synchronizations (wall clock), with one
    HIP event on the launch stream before the first and one after the last
    (their span over `steps` is the average launch period, gaps included).
    Nothing else is recorded inside the region: an event pair around every
    execute costs GPU time between the kernels.  A second, untimed pass then
    brackets each execute with events for the per-launch duration."""
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        qs.execute(plan, 1.0, beta)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        qs.execute(plan, 1.0, beta)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    period = e0.elapsed_time(e1) * 1e-3 / steps
    _, mean, median = time_plan(qs, torch, plan, beta, steps, 0)
    return wall, period, mean, median


def time_plans_alternating(qs, torch, plans, beta, reps, warmup):
    """Warm up every plan, then time `reps` rounds of one execute of each plan
    in turn with HIP events on the launch stream; (mean, median) seconds per plan."""
    stream = torch.cuda.current_stream()
    for p in plans:
        for _ in range(warmup):
            qs.execute(p, 1.0, beta)
    torch.cuda.synchronize()
    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
          for _ in plans]
    for i in range(reps):
        for j, p in enumerate(plans):
            ev[j][i][0].record(stream)
            qs.execute(p, 1.0, beta)
            ev[j][i][1].record(stream)
    torch.cuda.synchronize()
    out = []
    for j in range(len(plans)):
        per = [a.elapsed_time(b) * 1e-3 for a, b in ev[j]]
        out.append((float(np.mean(per)), float(np.median(per))))
    return out


def cpu_baseline(plan_views, seconds=10.0):
    """The oracle's BLIS restatement of the SIMDKernel nest, 1 thread, on the
    F64 headline workload: 1 warm-up, then the median of repetitions until
    about `seconds` of CPU time (at least 3, at most 50)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    plan, P, B, C, c0 = plan_views
    a = P
    b = B.parent.cpu().numpy()
    g = lambda G: (G.lengths, G.strides[0], G.strides[1])
    args = lambda c: (a, plan.Abase, b, plan.Bbase, c, plan.Cbase, g(plan.mgroup), g(plan.ngroup), g(plan.kgroup),
                      1, 1.0, 0.5)
    c = c0.copy()
    assert pyoracle.contract(*args(c), blis=True) == 0
    times = []
    t_start = time.perf_counter()
    while (len(times) < 3 or time.perf_counter() - t_start < seconds) and len(times) < 50:
        c = c0.copy()
        t0 = time.perf_counter()
        pyoracle.contract(*args(c), blis=True)
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), len(times), c, pyoracle.blis_kernel(1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    # About 20 ms of warm-up on the headline: right after autotuning, the
    # clocks take a few milliseconds to settle (a run timed then read 47 us
    # per launch, later runs of the same plan 40-42 us).
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune-cache", default=None,
                    help="autotune cache file: loaded if it exists (no tuning runs), else written after tuning")
    ap.add_argument("--only", default=None,
                    choices=["split", "split_f32", "grouped_f64", "grouped_c64", "served", "bandwidth", "batched", "mixed",
                             "rank4_f64", "c_anb_f64", "permuted_c64", "fp32", "peaks", "headline"],
                    help="only run one of the extra workloads and print its JSON (for profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 512 MiB device-to-device copy 5 times (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the contraction is measured on one MI355X; multi-GPU sharding is a design note only)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    qs = _load()
    dev = torch.device("cuda:0")

    if args.copy_calibration:
        n = (512 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"bytes_read_per_copy": 8 * n, "bytes_written_per_copy": 8 * n,
                                               "copies": 5}}), flush=True)
        return

    if args.only is not None:
        if args.only == "split":
            out = bench_split(qs, torch, dev)
        elif args.only == "split_f32":
            out = bench_split(qs, torch, dev, dtype=torch.float32)
        elif args.only == "bandwidth":
            out = bench_bandwidth(qs, torch, dev)
        elif args.only == "batched":
            out = bench_batched(qs, torch, dev)
        elif args.only == "mixed":
            out = bench_mixed_domain(qs, torch, dev)
        elif args.only.startswith("grouped"):
            out = bench_grouped(qs, torch, dev, args.only.endswith("c64"))
        elif args.only == "fp32":
            out = bench_fp32(qs, torch, dev)
        elif args.only == "headline":
            # The headline plan alone: autotuned, then the timed region three
            # times (--steps launches after --warmup), for counter runs.
            plan, flops, _, _, beta = config2(qs, torch, np.float64, dev)
            qs.autotune(plan, reps=10)
            out = []
            for _ in range(3):
                wall, period, _, _ = time_plan_region(qs, torch, plan, beta, args.steps, args.warmup)
                out.append({"plan": str(plan), "period_us": round(period * 1e6, 2),
                            "wall_us": round(wall / args.steps * 1e6, 2), "frac": round(flops / period / PEAK_FP64, 4)})
        elif args.only == "peaks":
            out = qs.measure_peaks()
        elif args.only in ("rank4_f64", "c_anb_f64", "permuted_c64"):
            out = bench_plain_extra(qs, torch, dev, args.only)
        else:
            pools = served_pools(qs, torch, dev)
            out = {"grouped": bench_served(qs, torch, dev, pools, True),
                   "per_request": bench_served(qs, torch, dev, pools, False)}
        print(json.dumps({args.only: out}), flush=True)
        return

    cached = args.tune_cache is not None and os.path.exists(args.tune_cache)
    if cached:
        qs.autotune_load(args.tune_cache)
    plan, flops, bytes_, views, beta = config2(qs, torch, np.float64, dev)
    c_init = views[3].parent.cpu().numpy().copy()
    if cached:
        tune = {"from_cache": args.tune_cache, "tile_m": plan.info["tile_m"], "tile_n": plan.info["tile_n"],
                "k_splits": plan.info["k_splits"]}
    else:
        tune = qs.autotune(plan, reps=10)
        if args.tune_cache:
            qs.autotune_save(args.tune_cache)
    wall, period, kern_mean, kern_med = time_plan_region(qs, torch, plan, beta, args.steps, args.warmup)
    value = flops * args.steps / wall / 1e9
    achieved = flops / period
    # HBM bytes per launch of the GEMM kernel, from the committed FETCH_SIZE /
    # WRITE_SIZE counter runs (scaled by a plain copy's reading).
    traffic_bytes, traffic_src = None, os.path.join(ROOT, "profiles", "headline_traffic.json")
    if os.path.exists(traffic_src):
        with open(traffic_src) as f:
            traffic_bytes = json.load(f).get("traffic_bytes_per_launch")

    # Measured peaks beside the datasheet ones (FP64 MFMA probe, device copy).
    peaks = qs.measure_peaks()
    result = {
        "metric": "GFLOP/s (plain FP64 block-scatter contraction)",
        "value": round(value, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded N(0,1) operands (numpy PCG64 seed 2), resident in HBM",
        "config": {
            "workload": "C[a,b,c,d] := A[a,e,b,f]*conj(B[f,c,e,d]), all extents 32 (M=N=K=1024), Float64, "
                        "A a stride-2 slice of a 64-wide parent on b, beta=0.5, C prefilled",
            "plan": repr(plan),
            "autotune": tune,
            "flops_per_step": flops,
        },
        "roofline": {
            "bound": "mfma",
            "achieved": round(achieved / 1e12, 3),
            "peak": 78.6,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64, 4),
            "frac_wall": round(flops / (wall / args.steps) / PEAK_FP64, 4),
            "peak_measured": round(peaks["fp64_mfma_tflops"], 2),
            "frac_of_measured_peak": round(achieved / 1e12 / peaks["fp64_mfma_tflops"], 4),
            "copy_gbs_measured": round(peaks["copy_gbs"], 1),
            "traffic": traffic_bytes,
            "launch_period_ms": round(period * 1e3, 5),
            "kernel_ms_mean": round(kern_mean * 1e3, 5),
            "kernel_ms_median": round(kern_med * 1e3, 5),
            "note": "achieved = flops per launch / launch_period_ms (HIP events at both ends of the timed "
                    "region, gaps included); frac_wall from ms_per_step (wall clock); kernel_ms_* from events "
                    "around each launch in a second pass; peak_measured: FP64 MFMA probe (qs_target_measure)",
        },
    }

    # correctness gate: one GPU execute from the initial C against the CPU baseline's result
    if not args.no_cpu:
        P, A, B, C = views
        med, reps, c_cpu, kname = cpu_baseline((plan, P, B, C, c_init))
        C.parent.copy_(torch.from_numpy(c_init).to(dev))
        qs.execute(plan, 1.0, beta)
        torch.cuda.synchronize()
        c_gpu = C.parent.cpu().numpy()
        result["config"]["rel_fro_err_vs_cpu"] = float(np.linalg.norm(c_gpu - c_cpu) / np.linalg.norm(c_cpu))
        result["cpu_baseline"] = {
            "value": round(flops / med / 1e9, 2),
            "unit": "GFLOP/s",
            "cores": 1,
            "kind": "port",
            "sample": f"the full headline contraction, median of {reps} reps after 1 warm-up "
                      f"(oracle BLIS nest, {kname}, 1 thread)",
        }
    if not args.no_extra:
        extra = {}
        for name, key in (("rank4_all64_f64", "rank4_f64"), ("c_anb_48x96_f64", "c_anb_f64"),
                          ("permuted_all32_c64", "permuted_c64")):
            extra[name] = bench_plain_extra(qs, torch, dev, key)
        extra["fp32"] = bench_fp32(qs, torch, dev)
        extra["split_rank4_all64_f64"] = bench_split(qs, torch, dev)
        extra["split_rank4_all64_f32"] = bench_split(qs, torch, dev, dtype=torch.float32)
        extra["bandwidth_f64"] = bench_bandwidth(qs, torch, dev)
        extra["batched_f64"] = bench_batched(qs, torch, dev)
        extra["mixed_domain_all32"] = bench_mixed_domain(qs, torch, dev)
        extra["grouped_512_mps_f64"] = bench_grouped(qs, torch, dev, False)
        extra["grouped_512_mps_c64"] = bench_grouped(qs, torch, dev, True)
        pools = served_pools(qs, torch, dev)
        extra["served_poisson_2k"] = {"grouped": bench_served(qs, torch, dev, pools, True),
                                      "per_request": bench_served(qs, torch, dev, pools, False)}
        extra["served_saturated"] = {"grouped": bench_saturated(qs, torch, dev, pools, True),
                                     "per_request": bench_saturated(qs, torch, dev, pools, False)}
        extra["served_saturated_small_only"] = {
            "grouped": bench_saturated(qs, torch, dev, pools, True, n=20000, small_share=1.0),
            "per_request": bench_saturated(qs, torch, dev, pools, False, n=20000, small_share=1.0)}
        del pools
        result["extra"] = extra
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
