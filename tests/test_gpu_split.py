"""Split precision (Ozaki scheme II: CRT over INT8 MFMA GEMMs) against the oracle, with the
delivered-error report.  Tolerances: relative Frobenius <= 1e-13 (FP64
target) and <= 1e-5 (FP32)."""
import numpy as np
import pytest
import torch

from helpers import SV, brute_ref, oracle_of_plan, qs, rel_fro, storage_np, values

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


def rank4(rng, d, dtype=np.float64, spread=20):
    """C[i,j,k,l] := A[i,m,j,n] * B[m,k,n,l]: entries N(0,1) * 2^u, u integer
    in [-spread, spread] per (i,j) of A and per (k,l) of B."""
    ua = rng.integers(-spread, spread + 1, size=(d, d))
    ub = rng.integers(-spread, spread + 1, size=(d, d))
    A = rng.standard_normal((d, d, d, d)) * np.exp2(ua)[:, None, :, None]
    B = rng.standard_normal((d, d, d, d)) * np.exp2(ub)[None, :, None, :]
    Av = SV(torch.from_numpy(A.reshape(-1, order="F").astype(dtype)).to(DEV), (d,) * 4, colmajor((d,) * 4))
    Bv = SV(torch.from_numpy(B.reshape(-1, order="F").astype(dtype)).to(DEV), (d,) * 4, colmajor((d,) * 4))
    Cv = SV(torch.from_numpy(rng.standard_normal(d ** 4).astype(dtype)).to(DEV), (d,) * 4, colmajor((d,) * 4))
    return Cv, Av, Bv


@pytest.mark.parametrize("d,beta", [(16, 0.0), (16, 0.5), (24, 0.0)])
def test_split_f64_rank4_meets_1e13(oracle, d, beta):
    case(f"split F64 rank-4 all {d}, 2^u row/col scaling, beta={beta}")
    rng = np.random.default_rng(40 + d)
    Cv, Av, Bv = rank4(rng, d)
    c0, a0, b0 = storage_np(Cv.parent), storage_np(Av.parent), storage_np(Bv.parent)
    plan = qs.plan_contract(Cv, Av, (1, 5, 2, 6), Bv, (5, 3, 6, 4), (1, 2, 3, 4), precision="split")
    assert plan.path == "split" and plan.info["moduli"] >= 10, plan.info
    qs.set_error_report(plan, True)
    qs.execute(plan, 1.25, beta)
    torch.cuda.synchronize()
    got = storage_np(Cv.parent)
    want = oracle_of_plan(oracle, plan, 1.25, beta, a0, b0, c0)
    err = rel_fro(got, want)
    rep = qs.error_report(plan)
    print(f"[split] d={d} moduli={plan.info['moduli']} err={err:.3e} report={rep}", flush=True)
    assert err <= 1e-13, err
    assert rep["samples"] > 0 and rep["rel_fro"] <= 1e-13, rep


def test_split_f64_ragged_and_strided(oracle):
    case("split F64 ragged extents (M, N, K not multiples of 64), strided operands")
    rng = np.random.default_rng(7)
    M, K, N = 70, 131, 95
    A = rng.standard_normal(2 * M * K)
    Av = SV(torch.from_numpy(A).to(DEV), (M, K), (2, 2 * M))            # every other row
    Bv = SV(torch.from_numpy(rng.standard_normal(K * N)).to(DEV), (K, N), (N, 1))   # K-fast
    Cv = SV(torch.from_numpy(rng.standard_normal(M * N + 5)).to(DEV), (M, N), (1, M), 5)
    c0 = storage_np(Cv.parent)
    plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), precision="split")
    assert plan.path == "split"
    qs.execute(plan, 1.0, -0.5)
    torch.cuda.synchronize()
    want = oracle_of_plan(oracle, plan, 1.0, -0.5, storage_np(plan.Astorage), storage_np(plan.Bstorage), c0)
    got = storage_np(Cv.parent)
    assert rel_fro(got, want) <= 1e-13
    assert np.array_equal(got[:5], c0[:5])


def test_split_f64_both_k_fast_ragged(oracle):
    """Both operands K fast (one row per wave in the row-norm pass) with row
    counts that are not multiples of the four rows a workgroup takes."""
    case("split F64 ragged, A and B both K fast")
    rng = np.random.default_rng(9)
    M, K, N = 70, 131, 95
    Av = SV(torch.from_numpy(rng.standard_normal(M * K)).to(DEV), (M, K), (K, 1))   # row-major: K fast
    Bv = SV(torch.from_numpy(rng.standard_normal(K * N)).to(DEV), (K, N), (1, K))   # K fast
    Cv = SV(torch.from_numpy(rng.standard_normal(M * N)).to(DEV), (M, N), (1, M))
    c0 = storage_np(Cv.parent)
    plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), precision="split")
    assert plan.path == "split" and plan.info["a_kfast"] == 1 and plan.info["b_kfast"] == 1, plan.info
    qs.execute(plan, 0.75, 0.5)
    torch.cuda.synchronize()
    want = oracle_of_plan(oracle, plan, 0.75, 0.5, storage_np(plan.Astorage), storage_np(plan.Bstorage), c0)
    assert rel_fro(storage_np(Cv.parent), want) <= 1e-13


def test_split_mixed_storage_types(oracle):
    """Float32 A and C, Float64 B: computed in Float64 (the promoted type),
    each operand's passes reading its own storage type; C rounded once into
    Float32, so the result is the oracle's to within an FP32 ulp."""
    case("split F64 compute, A and C Float32, B Float64")
    rng = np.random.default_rng(10)
    M, K, N = 90, 200, 70
    Av = SV(torch.from_numpy(rng.standard_normal(M * K).astype(np.float32)).to(DEV), (M, K), (1, M))
    Bv = SV(torch.from_numpy(rng.standard_normal(K * N)).to(DEV), (K, N), (1, K))
    Cv = SV(torch.from_numpy(rng.standard_normal(M * N).astype(np.float32)).to(DEV), (M, N), (1, M))
    c0 = storage_np(Cv.parent)
    plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), precision="split")
    assert plan.path == "split" and plan.T == torch.float64, plan
    qs.execute(plan, 1.0, 0.5)
    torch.cuda.synchronize()
    want = oracle_of_plan(oracle, plan, 1.0, 0.5, storage_np(plan.Astorage), storage_np(plan.Bstorage), c0)
    assert rel_fro(storage_np(Cv.parent).astype(np.float64), want.astype(np.float64)) <= 1e-7


def test_split_f32(oracle):
    case("split F32 (target 1e-6)")
    rng = np.random.default_rng(8)
    M, K, N = 200, 300, 150
    Av = SV(torch.from_numpy(rng.standard_normal(M * K).astype(np.float32)).to(DEV), (M, K), (1, M))
    Bv = SV(torch.from_numpy(rng.standard_normal(K * N).astype(np.float32)).to(DEV), (K, N), (1, K))
    Cv = SV(torch.from_numpy(np.zeros(M * N, np.float32)).to(DEV), (M, N), (1, M))
    plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), precision="split")
    qs.execute(plan, 1.0, 0.0)
    torch.cuda.synchronize()
    ref = brute_ref(Av, (1, 2), Bv, (2, 3), np.zeros((M, N)), (1, 3), 1.0, 0.0, dtype=np.float64)
    assert rel_fro(values(Cv), ref) <= 1e-5


def test_split_alpha_zero_never_reads_ab():
    case("split: alpha == 0 scales C only")
    M, K, N = 64, 64, 64
    Av = SV(torch.full((M * K,), float("nan"), dtype=torch.float64, device=DEV), (M, K), (1, M))
    Bv = SV(torch.full((K * N,), float("nan"), dtype=torch.float64, device=DEV), (K, N), (1, K))
    c0 = np.arange(M * N, dtype=np.float64)
    Cv = SV(torch.from_numpy(c0.copy()).to(DEV), (M, N), (1, M))
    plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), precision="split")
    qs.execute(plan, 0.0, 0.5)
    torch.cuda.synchronize()
    assert np.array_equal(storage_np(Cv.parent), 0.5 * c0)


def test_split_complex_is_rejected():
    Av = SV(torch.zeros(16, dtype=torch.complex128, device=DEV), (4, 4), (1, 4))
    Bv = SV(torch.zeros(16, dtype=torch.complex128, device=DEV), (4, 4), (1, 4))
    Cv = SV(torch.zeros(16, dtype=torch.complex128, device=DEV), (4, 4), (1, 4))
    with pytest.raises(qs.UnsupportedError):
        qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), precision="split")


def _plain_pair(rng, M, K, N, row_exp=None):
    A = rng.standard_normal((M, K))
    if row_exp is not None:
        A *= np.exp2(row_exp)[:, None]
    B = rng.standard_normal((K, N))
    Av = SV(torch.from_numpy(A.reshape(-1, order="F").copy()).to(DEV), (M, K), (1, M))
    Bv = SV(torch.from_numpy(B.reshape(-1, order="F").copy()).to(DEV), (K, N), (1, K))
    Cv = SV(torch.from_numpy(rng.standard_normal(M * N)).to(DEV), (M, N), (1, M))
    return A, B, Av, Bv, Cv


@pytest.mark.parametrize("M,K,N", [(70, 9000, 40), (24, 140000, 20)])
def test_split_long_k(oracle, M, K, N):
    """K = 9000: INT32 sums up to 2^27 on the four-wave kernel (no reduction
    inside its K loop up to a padded K of 130816).  K = 140000: past that, the
    eight-wave kernel, whose accumulators are reduced mod m every 4096 K."""
    case(f"split F64, M={M} K={K} N={N}")
    rng = np.random.default_rng(11)
    A, B, Av, Bv, Cv = _plain_pair(rng, M, K, N)
    c0 = storage_np(Cv.parent)
    plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), precision="split")
    qs.execute(plan, 1.0, 0.25)
    torch.cuda.synchronize()
    want = oracle_of_plan(oracle, plan, 1.0, 0.25, storage_np(Av.parent), storage_np(Bv.parent), c0)
    assert rel_fro(storage_np(Cv.parent), want) <= 1e-13


def test_split_extreme_row_scales_rowwise(oracle):
    case("split F64, rows of A scaled by 2^e, e in [-900, 900]: per-row error")
    rng = np.random.default_rng(12)
    M, K, N = 64, 200, 48
    e = rng.integers(-900, 901, size=M)
    A, B, Av, Bv, Cv = _plain_pair(rng, M, K, N, row_exp=e)
    plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), precision="split")
    qs.execute(plan, 1.0, 0.0)
    torch.cuda.synchronize()
    got = storage_np(Cv.parent).reshape(N, M).T
    want = oracle_of_plan(oracle, plan, 1.0, 0.0, storage_np(Av.parent), storage_np(Bv.parent),
                          np.zeros(M * N)).reshape(N, M).T
    for i in range(M):
        sc = np.abs(want[i]).max()   # norms of the rescaled rows: no overflow at 2^900
        err = np.linalg.norm((got[i] - want[i]) / sc) / np.linalg.norm(want[i] / sc)
        assert err <= 1e-13, (i, e[i], err)


def test_split_nan_row_and_zero_row(oracle):
    case("split F64: a NaN in row 3 of A poisons row 3 of C only; an all-zero row 7 gives beta * C")
    rng = np.random.default_rng(13)
    M, K, N = 40, 64, 24
    A, B, Av, Bv, Cv = _plain_pair(rng, M, K, N)
    A[3, 5] = np.nan
    A[7, :] = 0.0
    Av = SV(torch.from_numpy(A.reshape(-1, order="F").copy()).to(DEV), (M, K), (1, M))
    c0 = storage_np(Cv.parent).copy()
    plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), precision="split")
    qs.execute(plan, 1.0, 0.5)
    torch.cuda.synchronize()
    got = storage_np(Cv.parent).reshape(N, M).T
    assert np.isnan(got[3]).all()
    assert np.array_equal(got[7], 0.5 * c0.reshape(N, M).T[7])
    ok = np.ones(M, bool)
    ok[3] = False
    want = oracle_of_plan(oracle, plan, 1.0, 0.5, np.nan_to_num(storage_np(Av.parent)), storage_np(Bv.parent),
                          c0).reshape(N, M).T
    assert rel_fro(got[ok], want[ok]) <= 1e-13


def test_split_is_bitwise_reproducible():
    case("split F64: two executes with beta = 0 give identical bits")
    rng = np.random.default_rng(14)
    M, K, N = 300, 500, 260
    A, B, Av, Bv, Cv = _plain_pair(rng, M, K, N)
    plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), precision="split")
    qs.execute(plan, 1.0, 0.0)
    torch.cuda.synchronize()
    first = storage_np(Cv.parent).copy()
    qs.execute(plan, 1.0, 0.0)
    torch.cuda.synchronize()
    assert np.array_equal(storage_np(Cv.parent), first)


def _cancelling(rng, M, K, N, delta):
    """A = [X, X], B = [Y; -Y + delta E]: C = delta X E, with |C_ij| about
    (90 / delta)^-1 of ||A_i|| ||B_j|| at K = 4096 (cancellation)."""
    h = K // 2
    X = rng.standard_normal((M, h))
    Y = rng.standard_normal((h, N))
    E = rng.standard_normal((h, N))
    A = np.concatenate([X, X], axis=1)
    B = np.concatenate([Y, -Y + delta * E], axis=0)
    exact = delta * (X @ E)    # no cancellation in X E: float64 error ~1e-16 of it
    Av = SV(torch.from_numpy(A.reshape(-1, order="F").copy()).to(DEV), (M, K), (1, M))
    Bv = SV(torch.from_numpy(B.reshape(-1, order="F").copy()).to(DEV), (K, N), (1, K))
    return A, B, Av, Bv, exact


def test_split_under_cancellation_normwise_bound_and_guard():
    case("split F64 under cancellation (|C| << |A||B|): normwise bound, report flags it, guarded mode meets 1e-13")
    rng = np.random.default_rng(15)
    M, K, N = 192, 4096, 160
    A, B, Av, Bv, exact = _cancelling(rng, M, K, N, 0.02)
    norms = np.linalg.norm(A, axis=1)[:, None] * np.linalg.norm(B, axis=0)[None, :]
    print(f"[cancel] median ||A_i|| ||B_j|| / |C_ij| = {np.median(norms / np.abs(exact)):.0f}", flush=True)

    def run(mode, beta=0.0, c0=None):
        c0 = np.zeros(M * N) if c0 is None else c0
        Cv = SV(torch.from_numpy(c0.copy()).to(DEV), (M, N), (1, M))
        plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), precision="split" if mode is not None else "plain")
        if mode is not None:
            qs.set_error_report(plan, mode)
        qs.execute(plan, 1.0, beta)
        torch.cuda.synchronize()
        got = storage_np(Cv.parent).reshape(N, M).T
        rep = qs.error_report(plan) if mode is not None else None
        return got, rep, plan

    plain, _, _ = run(None)
    plain_err = rel_fro(plain, exact)
    got, rep, plan = run(True)
    err = rel_fro(got, exact)
    print(f"[cancel] plain FP64 rel err {plain_err:.3e}; split moduli={plan.info['moduli']} rel err {err:.3e} "
          f"report {rep}", flush=True)
    # the normwise bound holds entry by entry
    assert np.all(np.abs(got - exact) <= 1e-13 * norms)
    # relative to C the target is missed, and the report says so (its sampled error tracks the true one)
    assert err > 1e-13 and rep["exceeded"] == 1, (err, rep)
    assert 0.3 * err <= rep["rel_fro"] <= 3.0 * err, (err, rep)
    # guarded: redone with more moduli, from C as it was (beta != 0 here)
    c0 = rng.standard_normal(M * N)
    want = exact + 0.5 * c0.reshape(N, M).T
    got, rep, plan = run("guard", beta=0.5, c0=c0)
    gerr = rel_fro(got, want)
    print(f"[cancel] guarded moduli={plan.info['moduli']} err {gerr:.3e} report {rep}", flush=True)
    assert gerr <= 1e-13 and rep["exceeded"] == 0 and rep["escalations"] >= 1 and rep["fallback"] == 0, (gerr, rep)
    assert rep["moduli"] > 13                                # the raised count stays with the plan


def test_split_guard_falls_back_to_plain_past_15_moduli():
    case("split F64 guarded, cancellation beyond what 15 moduli resolve: the plain kernel finishes the execute")
    rng = np.random.default_rng(16)
    M, K, N = 96, 4096, 80
    A, B, Av, Bv, exact = _cancelling(rng, M, K, N, 2e-3)
    Cv = SV(torch.from_numpy(np.zeros(M * N)).to(DEV), (M, N), (1, M))
    plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), precision="split", target_error=1e-15)
    qs.set_error_report(plan, "guard")
    qs.execute(plan, 1.0, 0.0)
    torch.cuda.synchronize()
    rep = qs.error_report(plan)
    got = storage_np(Cv.parent).reshape(N, M).T
    Cp = SV(torch.from_numpy(np.zeros(M * N)).to(DEV), (M, N), (1, M))
    pp = qs.plan_contract(Cp, Av, (1, 2), Bv, (2, 3), (1, 3), kernel=qs.MFMAKernel(k_splits=1))
    assert pp.info["k_splits"] == 1
    qs.execute(pp, 1.0, 0.0)
    torch.cuda.synchronize()
    print(f"[cancel] fallback report {rep} moduli {plan.info['moduli']}", flush=True)
    assert rep["fallback"] == 1 and plan.info["moduli"] == 15, rep
    assert np.array_equal(got, storage_np(Cp.parent).reshape(N, M).T)   # the plain kernel's bits
