"""GPU parity: the gfx950 library through the C ABI against the oracle (same
groups, same storages, same seeds) and against numpy.einsum, at the
reference's test shapes and the benchmark configurations.

Tolerance: relative Frobenius error <= 1e-13 (FP64 / ComplexF64) and <= 1e-5
(FP32 / ComplexF32) against the oracle; integer-valued inputs are exact.
"""
import itertools

import numpy as np
import pytest
import torch

from helpers import RTOL, SV, TORCH_OF, brute_ref, oracle_of_plan, qs, randn, rel_fro, storage_np, values

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def case(msg):
    print(f"[case] {msg}", flush=True)


def dev_view(st_np, size, strides, offset=0, op="identity"):
    return SV(torch.from_numpy(np.ascontiguousarray(st_np)).to(DEV), size, strides, offset, op)


def colmajor_strides(size):
    out, s = [], 1
    for n in size:
        out.append(s)
        s *= n
    return tuple(out)


def run_and_check(oracle, Cv, Av, indA, Bv, indB, indC, alpha, beta, conjA=False, conjB=False, rtol=None,
                  expect_path=None, **kw):
    """Plan + execute on the GPU; oracle on the plan's groups; einsum reference.
    Returns the plan."""
    a_np, b_np, c_np = storage_np(Av.parent), storage_np(Bv.parent), storage_np(Cv.parent)
    cstart = values(Cv)
    plan = qs.plan_contract(Cv, Av, indA, Bv, indB, indC, conjA=conjA, conjB=conjB, **kw)
    if expect_path is not None:
        assert plan.path == expect_path, plan
    # oracle on the storages as the plan sees them (A/B may be swapped)
    a_o = b_np if plan.swapped else a_np
    b_o = a_np if plan.swapped else b_np
    qs.execute(plan, alpha, beta)
    torch.cuda.synchronize()
    got_storage = storage_np(Cv.parent)
    want_storage = oracle_of_plan(oracle, plan, alpha, beta, a_o, b_o, c_np)
    T = plan.T
    tol = RTOL[T] if rtol is None else rtol
    gv = values(SV(torch.from_numpy(got_storage), Cv.size, Cv.strides, Cv.offset))
    wv = values(SV(torch.from_numpy(want_storage), Cv.size, Cv.strides, Cv.offset))
    err = rel_fro(gv, wv)
    assert err <= tol, (err, plan)
    # nothing outside the view changed
    mask = np.ones(got_storage.size, bool)
    if Cv.numel():
        idx = np.full(Cv.size, Cv.offset, dtype=np.int64)
        for d, (n, s) in enumerate(zip(Cv.size, Cv.strides)):
            shp = [1] * len(Cv.size)
            shp[d] = n
            idx = idx + (np.arange(n) * s).reshape(shp)
        mask[idx.reshape(-1)] = False
    assert np.array_equal(got_storage[mask], c_np[mask], equal_nan=True)
    # independent reference (einsum in float64/complex128)
    ref = brute_ref(Av, indA, Bv, indB, cstart, indC, alpha, beta, conjA, conjB)
    if np.linalg.norm(ref) > 0:
        assert rel_fro(gv, ref) <= max(tol, 1e-12) * 10, rel_fro(gv, ref)
    return plan


# --------------------------------------------------------------- configs

def test_config1_plain_f64(oracle):
    case("config1 C[a,n,b] := A[a,k,b]*B[k,n], a=b=48, k=n=96, F64, seed 1")
    rng = np.random.default_rng(1)
    a, k, b, n = 48, 96, 48, 96
    A = dev_view(randn(rng, np.float64, a * k * b), (a, k, b), colmajor_strides((a, k, b)))
    B = dev_view(randn(rng, np.float64, k * n), (k, n), colmajor_strides((k, n)))
    C = dev_view(np.full(a * n * b, np.nan), (a, n, b), colmajor_strides((a, n, b)))
    plan = run_and_check(oracle, C, A, (1, 2, 3), B, (2, 4), (1, 4, 3), 1.0, 0.0, expect_path="gemm")
    assert plan.info["m"] == 2304 and plan.info["n"] == 96 and plan.info["k"] == 96


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_config2_permuted_strided(oracle, dtype):
    case(f"config2 C[a,b,c,d] := A[a,e,b,f]*conj(B[f,c,e,d]) all 32, stride-2 A slice, beta=0.5, {dtype.__name__}")
    rng = np.random.default_rng(2)
    d = 32
    P = randn(rng, dtype, d * d * 2 * d * d)                   # parent [32,32,64,32]
    pst = colmajor_strides((d, d, 2 * d, d))
    A = dev_view(P, (d, d, d, d), (pst[0], pst[1], 2 * pst[2], pst[3]))   # b = 1:2:64
    B = dev_view(randn(rng, dtype, d ** 4), (d,) * 4, colmajor_strides((d,) * 4))
    C = dev_view(randn(rng, dtype, d ** 4), (d,) * 4, colmajor_strides((d,) * 4))
    # labels a=1 e=2 b=3 f=4 c=5 d=6
    run_and_check(oracle, C, A, (1, 2, 3, 4), B, (4, 5, 2, 6), (1, 3, 5, 6), 1.0, 0.5, conjB=True,
                  expect_path="gemm")


def test_worked_fixture_exact(oracle):
    case("worked fixture 1:30 x 1:20, exact")
    A = np.arange(1.0, 31.0)
    B = np.arange(1.0, 21.0)
    Cv = dev_view(np.full(24, np.nan), (3, 4, 2), (1, 3, 12))
    Av = dev_view(A, (3, 5, 2), (1, 3, 15))
    Bv = dev_view(B, (5, 4), (1, 5))
    plan = qs.plan_contract(Cv, Av, (1, 2, 3), Bv, (2, 4), (1, 4, 3), path_modes=qs.PathModes("never", "never"))
    qs.execute(plan, 1.0, 0.0)
    got = values(Cv)
    want = np.einsum("akb,kn->anb", A.reshape(3, 5, 2, order="F"), B.reshape(5, 4, order="F"))
    assert np.array_equal(got, want)
    # the same through contract! with permuted operands and an offset view
    Cfull = dev_view(np.zeros(32), (4, 4, 2), (1, 4, 16))
    Cv2 = SV(Cfull.parent, (3, 4, 2), (1, 4, 16), 1)
    A4 = np.arange(1.0, 41.0)
    Av2 = dev_view(A4, (3, 5, 2), (1, 4, 20), 1)
    qs.contract(Cv2, 1.0, Av2, (1, 2, 3), Bv, (2, 4), 0.0, (1, 4, 3))
    A4r = A4.reshape(4, 5, 2, order="F")[1:]
    assert np.array_equal(values(Cv2), np.einsum("akb,kn->anb", A4r, B.reshape(5, 4, order="F")))
    assert np.all(values(Cfull)[0] == 0)


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.complex128, np.complex64])
def test_scattered_fixture(oracle, dtype):
    case(f"scattered fixture {np.dtype(dtype).name}: permuted A with a zero stride, reversed B, sliced C")
    rng = np.random.default_rng(11)
    an, kn, bn, nn = 32, 32, 8, 32
    A2 = randn(rng, dtype, an * kn)
    # A2 is [an, kn]; Aperm[k, b, a] = A2[a, k] with a zero-stride b axis
    Av = dev_view(A2, (kn, bn, an), (an, 0, 1))
    Bst = randn(rng, dtype, kn * nn)
    Bv = dev_view(Bst, (kn, nn), (-1, kn), kn - 1)
    big = (an + 2, nn + 3, bn + 1)
    Cbig = randn(rng, dtype, int(np.prod(big)))
    cst = colmajor_strides(big)
    Cv = dev_view(Cbig, (an, nn, bn), cst, 1 + cst[1])
    cplx = np.dtype(dtype).kind == "c"
    for alpha, beta in ((1.0, 0.0), ((1.3 - 0.4j) if cplx else 1.3, (0.7 + 0.2j) if cplx else 0.7)):
        run_and_check(oracle, Cv, Av, (2, 3, 1), Bv, (2, 4), (1, 4, 3), alpha, beta, conjA=cplx)


LO_CASES = (  # ccsd_t-shaped pairs: A[4], B[4] -> C[6] (test/planning/test_labels.jl LO_CASES pattern)
    ((1, 2, 3, 7), (7, 4, 5, 6), (1, 2, 3, 4, 5, 6)),
    ((2, 4, 5, 7), (1, 7, 3, 6), (1, 2, 3, 4, 5, 6)),
    ((3, 4, 7, 6), (7, 5, 1, 2), (1, 2, 3, 4, 5, 6)),
)


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_ccsd_shapes_permuted_sliced_c(oracle, dtype):
    rng = np.random.default_rng(0x1ABE10DE)
    d = 6
    perm = (2, 4, 0, 5, 3, 1)
    big = (d + 1, d + 2, d, d + 1, d, d + 3)
    bst = colmajor_strides(big)
    starts = (0, 1, 0, 1, 0, 2)
    cplx = np.dtype(dtype).kind == "c"
    for (indA, indB, indC), conj in itertools.product(LO_CASES, ((False, False), (True, True))):
        if conj[0] and not cplx:
            continue
        case(f"ccsd {indA} x {indB} -> {indC} conj={conj} {np.dtype(dtype).name}")
        Av = dev_view(randn(rng, dtype, d ** 4), (d,) * 4, colmajor_strides((d,) * 4))
        Bv = dev_view(randn(rng, dtype, d ** 4), (d,) * 4, colmajor_strides((d,) * 4))
        Cbig = randn(rng, dtype, int(np.prod(big)))
        off = sum(s * st for s, st in zip(starts, bst))
        Cv = SV(torch.from_numpy(Cbig).to(DEV), tuple(d for _ in perm), tuple(bst[p] for p in perm), off)
        alpha = (1.3 - 0.4j) if cplx else 1.3
        beta = (0.7 + 0.2j) if cplx else 0.7
        run_and_check(oracle, Cv, Av, indA, Bv, indB, indC, alpha, beta, conjA=conj[0], conjB=conj[1])


@pytest.mark.parametrize("TA,TB,TC,acc", [
    (np.float32, np.float64, np.float64, None),
    (np.float64, np.float32, np.float32, None),
    (np.float64, np.complex128, np.complex128, None),
    (np.complex64, np.float64, np.complex128, None),
    (np.complex64, np.complex128, np.complex128, None),
    (np.float64, np.float64, np.complex128, None),
    (np.complex128, np.float32, np.complex128, torch.float32),
    (np.float32, np.float32, np.float32, torch.float64),
])
def test_mixed_eltypes(oracle, TA, TB, TC, acc):
    case(f"mixed {np.dtype(TA).name} x {np.dtype(TB).name} -> {np.dtype(TC).name} acc={acc}")
    rng = np.random.default_rng(7)
    M, K, N = 45, 67, 39
    Av = dev_view(randn(rng, TA, M * K), (M, K), (1, M))
    Bv = dev_view(randn(rng, TB, K * N), (K, N), (1, K))
    Cv = dev_view(randn(rng, TC, M * N), (M, N), (1, M))
    T = qs.compute_type(TORCH_OF[np.dtype(TA)], TORCH_OF[np.dtype(TB)], TORCH_OF[np.dtype(TC)], acc)
    # one rounding into a narrower C: bound by the storage type too
    tol = max(RTOL[T], 1e-6 if np.dtype(TC) in (np.dtype(np.float32), np.dtype(np.complex64)) else 0)
    run_and_check(oracle, Cv, Av, (1, 2), Bv, (2, 3), (1, 3), 1.5, 0.5, accumulator=acc, rtol=tol)


@pytest.mark.parametrize("TA,TB,TC,acc,conjA,conjB,dom", [
    (np.float64, np.complex128, np.complex128, None, False, True, 2),
    (np.complex64, np.float64, np.complex128, None, True, False, 1),
    (np.complex128, np.float32, np.complex128, torch.float32, False, False, 1),
    (np.float32, np.complex64, np.complex64, None, False, True, 2),
    (np.float64, np.float64, np.complex128, None, False, False, 0),
])
def test_mixed_domain_kernels(oracle, TA, TB, TC, acc, conjA, conjB, dom):
    """ComplexRealKernel / RealComplexKernel (kernels.jl:125-160; cases after
    test/execution/test_mixed.jl:28-69): one real operand keeps one plane and
    two MFMAs per block; scattered strides, conj on the complex side, complex
    beta.  Two real operands into a complex C take the full complex path."""
    case(f"mixed domain {np.dtype(TA).name} x {np.dtype(TB).name} -> {np.dtype(TC).name} acc={acc}")
    rng = np.random.default_rng(95)
    d = 8
    A = randn(rng, TA, 2 * d ** 4)
    B = randn(rng, TB, d ** 4)
    Cst = randn(rng, TC, d ** 4)
    # C[a,b,c,d] = A[a,e,b,f] * B[f,c,e,d] with A a stride-2 slice along b
    Av = dev_view(A, (d,) * 4, (1, d, 2 * d * d, 2 * d ** 3))
    Bv = dev_view(B, (d,) * 4, (1, d, d * d, d ** 3))
    Cv = dev_view(Cst, (d,) * 4, (1, d, d * d, d ** 3))
    T = qs.compute_type(TORCH_OF[np.dtype(TA)], TORCH_OF[np.dtype(TB)], TORCH_OF[np.dtype(TC)], acc)
    tol = max(RTOL[T], 1e-6 if np.dtype(TC) in (np.dtype(np.float32), np.dtype(np.complex64)) else 0)
    beta = 0.3 - 0.7j
    plan = run_and_check(oracle, Cv, Av, (1, 2, 3, 4), Bv, (4, 5, 2, 6), (1, 3, 5, 6), 1.5, beta,
                         conjA=conjA, conjB=conjB, accumulator=acc, rtol=tol,
                         path_modes=qs.PathModes("never", "never"))
    want_dom = dom if not plan.swapped else {1: 2, 2: 1, 0: 0}[dom]
    assert plan.info["mixed_domain"] == want_dom, (plan, plan.info)


def test_one_rounding_into_narrow_c(oracle):
    # test/execution/test_c_panel.jl: 1, 2^-30, -1 along K survive only in the compute type
    case("one rounding: F64 x F64 -> F32, K = 3000")
    M, N, K = 13, 40, 3000
    A = np.zeros((M, K))
    A[:, 0] = 1
    A[:, K // 2] = 2.0 ** -30
    A[:, K - 1] = -1
    for beta in (0.5, 1.0, 0.0):
        Av = dev_view(A.reshape(-1, order="F"), (M, K), (1, M))
        Bv = dev_view(np.ones(K * N), (K, N), (1, K))
        Cv = dev_view(np.full(M * N, np.nan if beta == 0 else 2.0 ** -28, dtype=np.float32), (M, N), (1, M))
        plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3))
        qs.execute(plan, 1.0, beta)
        assert np.all(values(Cv) == np.float32(2.0 ** -30 + beta * 2.0 ** -28)), plan


# ------------------------------------------------------ degenerate paths

@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.float32])
@pytest.mark.parametrize("shape,layout", [((1, 700, 29), "dense"), ((37, 640, 1), "dense"),
                                          ((37, 640, 1), "transposed"), ((1, 70, 29), "transposed")])
def test_dot_path(oracle, dtype, shape, layout):
    M, K, N = shape
    case(f"dot M={M} K={K} N={N} {layout} {np.dtype(dtype).name}")
    rng = np.random.default_rng(3)
    if layout == "dense":
        Av = dev_view(randn(rng, dtype, M * K), (M, K), (1, M))
        Bv = dev_view(randn(rng, dtype, K * N), (K, N), (1, K))
    else:
        Av = dev_view(randn(rng, dtype, M * K), (M, K), (K, 1))
        Bv = dev_view(randn(rng, dtype, K * N), (K, N), (N, 1))
    Cv = dev_view(randn(rng, dtype, M * N), (M, N), (1, M))
    cplx = np.dtype(dtype).kind == "c"
    run_and_check(oracle, Cv, Av, (1, 2), Bv, (2, 3), (1, 3), 1.5, 0.5, conjA=cplx, expect_path="dot")


@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.float32])
@pytest.mark.parametrize("variant", ["plain", "cpermC", "bstrided"])
def test_outer_path(oracle, dtype, variant):
    case(f"outer {variant} {np.dtype(dtype).name}")
    rng = np.random.default_rng(4)
    M, N = 300, 257
    Av = dev_view(randn(rng, dtype, M), (M,), (1,))
    if variant == "bstrided":
        Bv = dev_view(randn(rng, dtype, 2 * N), (N,), (2,), 1)
    else:
        Bv = dev_view(randn(rng, dtype, N), (N,), (1,))
    if variant == "cpermC":
        Cv = dev_view(randn(rng, dtype, M * N), (M, N), (N, 1))
    else:
        Cv = dev_view(randn(rng, dtype, M * N), (M, N), (1, M))
    run_and_check(oracle, Cv, Av, (1,), Bv, (2,), (1, 2), 1.5, 0.5, expect_path="outer")


def test_short_circuits_and_empty(oracle):
    case("alpha == 0 / K == 0 never read A, B; beta == 0 never reads C; empty M/N")
    rng = np.random.default_rng(1)
    M, K, N = 5, 6, 4
    for Kx, alpha in ((0, 1.0), (K, 0.0)):
        for beta in (0.5, 0.0):
            Av = dev_view(np.full(max(M * Kx, 1), np.nan), (M, Kx), (1, M))
            Bv = dev_view(np.full(max(Kx * N, 1), np.nan), (Kx, N), (1, max(Kx, 1)))
            c0 = rng.standard_normal(M * N)
            Cv = dev_view(c0, (M, N), (1, M))
            plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3))
            if Kx == 0:
                assert plan.path == "scale"
            qs.execute(plan, alpha, beta)
            assert np.array_equal(values(Cv), beta * c0.reshape(M, N, order="F"))
    # beta == 0 never reads a NaN C, across tiles
    M, K, N = 300, 100, 190
    Av = dev_view(rng.standard_normal(M * K), (M, K), (1, M))
    Bv = dev_view(rng.standard_normal(K * N), (K, N), (1, K))
    Cv = dev_view(np.full(M * N, np.nan), (M, N), (1, M))
    run_and_check(oracle, Cv, Av, (1, 2), Bv, (2, 3), (1, 3), 2.5, 0.0)
    for Mx, Nx in ((3, 0), (0, 3)):
        Av = dev_view(np.ones(max(Mx * 4, 1)), (Mx, 4), (1, max(Mx, 1)))
        Bv = dev_view(np.ones(max(4 * Nx, 1)), (4, Nx), (1, 4))
        Cv = dev_view(np.full(max(Mx * Nx, 1), np.nan), (Mx, Nx), (1, max(Mx, 1)))
        plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3))
        assert plan.path == "empty"
        qs.execute(plan, 1.0, 2.0)


def test_out_of_bounds_rejected_before_any_write():
    case("bounds: each operand one element short")
    M, K, N = 8, 6, 5
    for sa, sb, sc in ((M * K - 1, K * N, M * N), (M * K, K * N - 1, M * N), (M * K, K * N, M * N - 1)):
        Cv = dev_view(np.zeros(sc), (M, N), (1, M))
        Av = dev_view(np.ones(sa), (M, K), (1, M))
        Bv = dev_view(np.ones(sb), (K, N), (1, K))
        plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3))
        with pytest.raises(qs.BoundsError):
            qs.execute(plan, 1.0, 0.0)
        torch.cuda.synchronize()
        assert np.all(storage_np(Cv.parent) == 0)


def test_aliasing_and_conjugated_output_rejected():
    case("aliasing / conjugated C rejected")
    M = torch.randn(32, dtype=torch.float64, device=DEV)
    Cv = SV(M, (4, 4), (1, 8), 0)
    Av = SV(M, (4, 4), (1, 8), 4)
    Bv = dev_view(np.ones(16), (4, 4), (1, 4))
    with pytest.raises(qs.ArgumentError):
        qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3))
    Cc = dev_view(np.zeros(16, dtype=np.complex128), (4, 4), (1, 4), 0, "conj")
    Ac = dev_view(np.ones(16, dtype=np.complex128), (4, 4), (1, 4))
    with pytest.raises(qs.ArgumentError):
        qs.plan_contract(Cc, Ac, (1, 2), Ac, (2, 3), (1, 3))


@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.float32])
def test_split_k(oracle, dtype):
    case(f"split-K M=N=64 K=8192 {np.dtype(dtype).name}")
    rng = np.random.default_rng(8)
    M, K, N = 64, 8192, 64
    Av = dev_view(randn(rng, dtype, M * K), (M, K), (1, M))
    Bv = dev_view(randn(rng, dtype, K * N), (K, N), (1, K))
    Cv = dev_view(randn(rng, dtype, M * N), (M, N), (1, M))
    plan = run_and_check(oracle, Cv, Av, (1, 2), Bv, (2, 3), (1, 3), 1.0, 0.5)
    assert plan.info["k_splits"] > 1, plan


@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4, 5])
def test_forced_split_k_ragged(oracle, tile):
    """Split-K chunks are whole stage pairs; the last split ends at a ragged K
    (the K tables' pad entries read as zero)."""
    case(f"forced split-K 3, ragged K=1000, tile {tile}")
    rng = np.random.default_rng(40 + tile)
    M, K, N = 100, 1000, 70
    for dtype in (np.float64, np.complex128, np.float32):
        Av = dev_view(randn(rng, dtype, M * K), (M, K), (1, M))
        Bv = dev_view(randn(rng, dtype, K * N), (K, N), (N, 1) if tile % 2 else (1, K))
        Cv = dev_view(randn(rng, dtype, M * N), (M, N), (1, M))
        plan = run_and_check(oracle, Cv, Av, (1, 2), Bv, (2, 3), (1, 3), 1.0, 0.5,
                             kernel=qs.MFMAKernel(tile=tile, k_splits=3), path_modes=qs.PathModes("never", "never"))
        assert plan.info["k_splits"] == 3, plan


def test_edge_rows_do_not_leak():
    """Rows past M (columns past N) in a partial tile repeat the last one; a
    NaN and an Inf there reach only that row (column) of C."""
    case("NaN / Inf in the last row of A and the last column of B, partial tiles")
    rng = np.random.default_rng(44)
    M, K, N = 70, 96, 50
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    A[M - 1, 5] = np.nan
    B[7, N - 1] = np.inf
    Av = dev_view(A.reshape(-1, order="F"), (M, K), (1, M))
    Bv = dev_view(B.reshape(-1, order="F"), (K, N), (1, K))
    Cv = dev_view(np.zeros(M * N), (M, N), (1, M))
    plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), path_modes=qs.PathModes("never", "never"))
    qs.execute(plan, 1.0, 0.0)
    torch.cuda.synchronize()
    got = values(Cv)
    want = A[: M - 1] @ B[:, : N - 1]
    assert rel_fro(got[: M - 1, : N - 1], want) <= 1e-13
    assert np.all(np.isnan(got[M - 1]))
    assert not np.any(np.isfinite(got[: M - 1, N - 1]))


def test_span_past_2gib_takes_generic_gather():
    """An operand whose accessed span exceeds the buffer-load range (2^31
    bytes) runs the generic 64-bit gather, with the same results."""
    case("A span of 2.6 GB (stride 2^23 along K), F64")
    rng = np.random.default_rng(45)
    M, K, N, ks = 64, 40, 32, 1 << 23
    st = torch.zeros((K - 1) * ks + M, dtype=torch.float64, device=DEV)
    A = rng.standard_normal((M, K))
    idx = (np.arange(M)[:, None] + ks * np.arange(K)[None, :]).reshape(-1)
    st[torch.from_numpy(idx).to(DEV)] = torch.from_numpy(A.reshape(-1)).to(DEV)
    Av = SV(st, (M, K), (1, ks), 0)
    B = rng.standard_normal((K, N))
    Bv = dev_view(B.reshape(-1, order="F"), (K, N), (1, K))
    Cv = dev_view(np.zeros(M * N), (M, N), (1, M))
    plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), path_modes=qs.PathModes("never", "never"))
    qs.execute(plan, 1.0, 0.0)
    torch.cuda.synchronize()
    assert rel_fro(values(Cv), A @ B) <= 1e-13
    del st


@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4, 5])
def test_every_tile_shape(oracle, tile):
    case(f"tile override {tile}")
    rng = np.random.default_rng(20 + tile)
    M, K, N = 200, 77, 150
    for dtype in (np.float64, np.complex64):
        Av = dev_view(randn(rng, dtype, M * K), (M, K), (1, M))
        Bv = dev_view(randn(rng, dtype, K * N), (K, N), (N, 1) if tile % 2 else (1, K))
        Cv = dev_view(randn(rng, dtype, M * N), (M, N), (1, M))
        run_and_check(oracle, Cv, Av, (1, 2), Bv, (2, 3), (1, 3), 1.0, 1.0, kernel=qs.MFMAKernel(tile=tile),
                      path_modes=qs.PathModes("never", "never"))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("akf", [False, True])
def test_paired_loads_match_single_loads(oracle, dtype, akf):
    """Tiles on the two-stage pipeline load two adjacent elements per lane when
    every pair the gathers read is adjacent.  Every tile sums K in the same
    order, so their results match the 128 x 128 tile's (one element per lane)
    bit for bit, partial tiles included; an odd M (for an M-fast A) or K
    strided by 2 turns the pairing off."""
    rng = np.random.default_rng(90 + int(akf))
    M, K, N = 250, 98, 190
    A = randn(rng, dtype, 2 * M * K)
    B = randn(rng, dtype, 2 * K * N)
    C0 = randn(rng, dtype, M * N)

    def views(m, kstride):
        a_st = (K * kstride, kstride) if akf else (1, m)   # A K-fast (rows along K) or M-fast
        Av = dev_view(A, (m, K), a_st)
        Bv = dev_view(B, (K, N), (kstride, K * kstride))     # B K-fast
        Cv = dev_view(C0, (m, N), (1, m))
        return Av, Bv, Cv

    # An odd M breaks the row pairs of an M-fast A only; a K-fast A pairs along K.
    for m, kstride, paired in ((M, 1, True), (M - 1, 1, akf), (M, 2, False)):
        case(f"paired loads {np.dtype(dtype).name}, A {'K' if akf else 'M'}-fast, M={m}, K stride {kstride}")
        results = []
        for tile in range(6):
            Av, Bv, Cv = views(m, kstride)
            plan = run_and_check(oracle, Cv, Av, (1, 2), Bv, (2, 3), (1, 3), 1.5, 0.5,
                                 kernel=qs.MFMAKernel(tile=tile, k_splits=1), path_modes=qs.PathModes("never", "never"))
            assert plan.info["k_splits"] == 1, plan
            on = plan.info["paired_loads"] == 1
            assert on == (paired and plan.info["tile_m"] * plan.info["tile_n"] <= 64 * 128), (plan, plan.info)
            results.append(values(Cv).copy())
        for r in results[1:]:
            assert np.array_equal(r, results[0])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("akf", [False, True])
def test_one_tile_row_streams_b(oracle, dtype, akf):
    """M <= 32 on 32-row tiles: one tile row, so each B element is read by one
    workgroup and B's loads are non-temporal.  The results match the oracle
    and, bit for bit, the 128 x 128 tile's; ragged N and K, and an odd M (no
    paired loads for an M-fast A) included."""
    rng = np.random.default_rng(140 + int(akf))
    K, N = 301, 517
    for m in (16, 32, 31):
        case(f"streamed B {np.dtype(dtype).name}, A {'K' if akf else 'M'}-fast, M={m}, K={K}, N={N}")
        A = randn(rng, dtype, m * K)
        B = randn(rng, dtype, K * N)
        C0 = randn(rng, dtype, m * N)
        results = []
        for tile in (4, 0):
            Av = dev_view(A, (m, K), (K, 1) if akf else (1, m))
            Bv = dev_view(B, (K, N), (1, K))
            Cv = dev_view(C0, (m, N), (1, m))
            plan = run_and_check(oracle, Cv, Av, (1, 2), Bv, (2, 3), (1, 3), 1.5, 0.5,
                                 kernel=qs.MFMAKernel(tile=tile, k_splits=1), path_modes=qs.PathModes("never", "never"))
            assert plan.info["k_splits"] == 1 and plan.info["tile_m"] == (32 if tile == 4 else 128), plan.info
            assert not plan.swapped and plan.info["m"] == m, plan
            results.append(values(Cv).copy())
        assert np.array_equal(results[0], results[1])


def test_host_memory_operands(oracle):
    case("host-memory operands staged through the library")
    rng = np.random.default_rng(12)
    M, K, N = 70, 90, 50
    A = torch.from_numpy(rng.standard_normal(M * K))
    B = torch.from_numpy(rng.standard_normal(K * N + 3))
    Cst = torch.from_numpy(rng.standard_normal(2 * M * N))
    Av, Bv, Cv = SV(A, (M, K), (1, M)), SV(B, (K, N), (1, K), 3), SV(Cst, (M, N), (2, 2 * M))
    c0 = Cst.numpy().copy()
    plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3))
    qs.execute(plan, 0.75, -0.5)
    want = oracle_of_plan(oracle, plan, 0.75, -0.5, A.numpy(), B.numpy(), c0)
    assert rel_fro(Cst.numpy(), want) <= 1e-13
    assert np.array_equal(Cst.numpy()[1::2], c0[1::2])


# ------------------------------------------- TensorOperations backend path

@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.complex64, np.complex128])
def test_tensorcontract_backend(oracle, dtype):
    # test/integrations/test_tensoroperations.jl:21-42 cases
    rng = np.random.default_rng(1234567)
    backend = qs.QuasiStridedBackend()
    cases = (
        ((3, 20, 5, 3, 4), (4, 6, 20, 3), ((3, 1, 4), (2, 5)), ((3, 1), (4, 2)), ((3, 1, 4), (5, 2)), (3, 5, 3, 6, 3)),
        ((4, 5, 6), (3, 5, 6), ((1,), (2, 3)), ((2, 3), (1,)), ((2, 1), ()), (3, 4)),
    )
    cplx = np.dtype(dtype).kind == "c"
    for szA, szB, pA, pB, pAB, szC in cases:
        for conjA, conjB in (((False, False), (True, True)) if cplx else ((False, False),)):
            for alpha, beta in ((1.0, 0.0), (0.5, 0.0), (0.5, 0.25)):
                case(f"tensorcontract {szA} {szB} {pA} {pB} {pAB} conj={conjA},{conjB} {np.dtype(dtype).name}")
                A = dev_view(randn(rng, dtype, int(np.prod(szA))), szA, colmajor_strides(szA))
                B = dev_view(randn(rng, dtype, int(np.prod(szB))), szB, colmajor_strides(szB))
                c0 = randn(rng, dtype, int(np.prod(szC)))
                if beta == 0:
                    c0[:] = np.nan
                C = dev_view(c0, szC, colmajor_strides(szC))
                cstart = values(C)
                backend.tensorcontract(C, A, pA, conjA, B, pB, conjB, pAB, alpha, beta)
                got = values(C)
                assert np.all(np.isfinite(got))
                indA, indB, indC = qs.contraction_labels(pA, pB, pAB)
                ref = brute_ref(A, indA, B, indB, cstart, indC, alpha, beta, conjA, conjB)
                tol = 1e-12 if np.dtype(dtype).itemsize in (8, 16) and dtype != np.complex64 else 2e-5
                assert rel_fro(got, ref) <= tol


# --------------------------------------------- full-size property checks

def test_config4_rank4_f64_checksum():
    case("config4 C[i,j,k,l] := A[i,m,j,n]*B[m,k,n,l], all 64 (M=N=K=4096), F64: checksum property")
    rng = np.random.default_rng(4)
    d = 64
    ua = rng.integers(-20, 21, size=(d * d,))
    ub = rng.integers(-20, 21, size=(d * d,))
    A = rng.standard_normal((d, d, d, d))                      # A[i,m,j,n], numpy C order
    B = rng.standard_normal((d, d, d, d))                      # B[m,k,n,l]
    A *= np.exp2(ua.reshape(d, d))[:, None, :, None]           # per (i, j)
    B *= np.exp2(ub.reshape(d, d))[None, :, None, :]           # per (k, l)
    # column-major storage: flatten in Fortran order
    Av = dev_view(A.reshape(-1, order="F"), (d,) * 4, colmajor_strides((d,) * 4))
    Bv = dev_view(B.reshape(-1, order="F"), (d,) * 4, colmajor_strides((d,) * 4))
    Cv = dev_view(np.full(d ** 4, np.nan), (d,) * 4, colmajor_strides((d,) * 4))
    plan = qs.plan_contract(Cv, Av, (1, 5, 2, 6), Bv, (5, 3, 6, 4), (1, 2, 3, 4))
    assert plan.path == "gemm"
    qs.execute(plan, 1.0, 0.0)
    torch.cuda.synchronize()
    Cg = storage_np(Cv.parent).reshape((d,) * 4, order="F")     # C[i,j,k,l]
    assert np.all(np.isfinite(Cg))
    Am = A.transpose(0, 2, 1, 3).reshape(d * d, d * d, order="F")   # [(i,j),(m,n)]
    Bm = B.transpose(0, 2, 1, 3).reshape(d * d, d * d, order="F")   # [(m,n),(k,l)]
    Cm = Cg.reshape(d * d, d * d, order="F")
    for s in range(3):
        r = np.random.default_rng(100 + s)
        u = r.standard_normal(d * d)
        v = r.standard_normal(d * d)
        lhs = u @ Cm @ v
        rhs = (u @ Am) @ (Bm @ v)
        scale = np.abs(u) @ np.abs(Am) @ np.abs(Bm) @ np.abs(v)
        assert abs(lhs - rhs) <= 1e-12 * scale
    # rows: C e_q = A (B e_q) for a few columns, elementwise relative to |A||B|
    for q in (0, 777, 4095):
        col = Am @ Bm[:, q]
        scale = np.abs(Am) @ np.abs(Bm[:, q])
        assert np.all(np.abs(Cm[:, q] - col) <= 1e-12 * scale + 1e-300)


def test_autotune_keeps_results_and_c(oracle):
    case("autotune on the permuted all-32 F64 contraction")
    rng = np.random.default_rng(2)
    d = 32
    P = randn(rng, np.float64, d * d * 2 * d * d)
    pst = colmajor_strides((d, d, 2 * d, d))
    A = dev_view(P, (d,) * 4, (pst[0], pst[1], 2 * pst[2], pst[3]))
    B = dev_view(randn(rng, np.float64, d ** 4), (d,) * 4, colmajor_strides((d,) * 4))
    C = dev_view(randn(rng, np.float64, d ** 4), (d,) * 4, colmajor_strides((d,) * 4))
    qs.autotune_clear()
    plan = qs.plan_contract(C, A, (1, 2, 3, 4), B, (4, 5, 2, 6), (1, 3, 5, 6))
    c0 = storage_np(C.parent)
    res = qs.autotune(plan, reps=3)
    torch.cuda.synchronize()
    assert np.array_equal(storage_np(C.parent), c0)          # tuning writes a scratch C
    assert res["candidates"] > 1 and res["best_ms"] <= res["default_ms"]
    print("[tune]", res, plan, flush=True)
    # the tuned plan still matches the oracle; a new plan of the class starts tuned
    run_and_check(oracle, C, A, (1, 2, 3, 4), B, (4, 5, 2, 6), (1, 3, 5, 6), 1.0, 0.5)
    p2 = qs.plan_contract(C, A, (1, 2, 3, 4), B, (4, 5, 2, 6), (1, 3, 5, 6))
    assert (p2.info["tile_m"], p2.info["tile_n"], p2.info["k_splits"]) == \
        (plan.info["tile_m"], plan.info["tile_n"], plan.info["k_splits"])
    qs.autotune_clear()


def test_backend_plan_cache_rebinds_layouts():
    """The TensorOperations backend plans every call; the layout cache keeps
    one device plan per layout and rebinds storages (no device allocation per
    call), a non-default allocator bypasses it, and a short storage is still
    rejected before any write."""
    case("backend plan cache: repeated layouts, new storages")
    pl = qs.plan
    pl.clear_plan_cache()
    rng = np.random.default_rng(97)
    be = qs.QuasiStridedBackend()
    for i in range(3):
        A = torch.from_numpy(rng.standard_normal((30, 20, 24))).to(DEV)
        B = torch.from_numpy(rng.standard_normal((20, 16))).to(DEV)
        Cd = torch.zeros(30, 24, 16, dtype=torch.float64, device=DEV)
        be.tensorcontract(Cd, A, ((1, 3), (2,)), False, B, ((1,), (2,)), False, ((1, 2, 3), ()), 1.0, 0.0)
        torch.cuda.synchronize()
        want = np.einsum("akb,kn->abn", A.cpu().numpy(), B.cpu().numpy())
        assert rel_fro(Cd.cpu().numpy(), want) <= 1e-13
        assert len(pl._plan_cache) == 1
    A2 = torch.from_numpy(rng.standard_normal((32, 40))).to(DEV)
    B2 = torch.from_numpy(rng.standard_normal((40, 8))).to(DEV)
    C2 = torch.zeros(32, 8, dtype=torch.float64, device=DEV)
    be.tensorcontract(C2, A2, ((1,), (2,)), False, B2, ((1,), (2,)), False, ((1, 2), ()), 1.0, 0.0)
    assert len(pl._plan_cache) == 2
    al = qs.TorchAllocator()
    be.tensorcontract(C2, A2, ((1,), (2,)), False, B2, ((1,), (2,)), False, ((1, 2), ()), 2.0, 0.0, al)
    torch.cuda.synchronize()
    assert len(pl._plan_cache) == 2 and al.live_bytes == 0
    assert rel_fro(C2.cpu().numpy(), 2.0 * (A2 @ B2).cpu().numpy()) <= 1e-13
    # the cached plan of the first layout, rebound to a C storage one element short
    A = torch.from_numpy(rng.standard_normal((30, 20, 24))).to(DEV)
    B = torch.from_numpy(rng.standard_normal((20, 16))).to(DEV)
    cpar = rng.standard_normal(30 * 24 * 16 - 1)
    Cv = SV(torch.from_numpy(cpar.copy()).to(DEV), (30, 24, 16), (24 * 16, 16, 1))   # the first layout
    with pytest.raises(qs.BoundsError):
        be.tensorcontract(Cv, A, ((1, 3), (2,)), False, B, ((1,), (2,)), False, ((1, 2, 3), ()), 1.0, 0.0)
    assert np.array_equal(Cv.parent.cpu().numpy(), cpar)
    assert len(pl._plan_cache) == 2
    pl.clear_plan_cache()
    assert len(pl._plan_cache) == 0


@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4, 5])
def test_mixed_domain_split_k(oracle, tile):
    """Mixed-domain kernels with forced split-K (partials of both planes, the
    reduce kernel on the buffered and the 64-bit gather variants), ragged K."""
    rng = np.random.default_rng(120 + tile)
    M, K, N = 90, 700, 60
    for TA, TB, TC in ((np.complex128, np.float64, np.complex128), (np.float64, np.complex128, np.complex128),
                       (np.complex64, np.float64, np.complex128)):
        case(f"mixed domain split-K 3, tile {tile}, {np.dtype(TA).name} x {np.dtype(TB).name}")
        Av = dev_view(randn(rng, TA, M * K), (M, K), (1, M))
        Bv = dev_view(randn(rng, TB, K * N), (K, N), (1, K))
        Cv = dev_view(randn(rng, TC, M * N), (M, N), (1, M))
        plan = run_and_check(oracle, Cv, Av, (1, 2), Bv, (2, 3), (1, 3), 0.75, 0.5 - 0.25j,
                             kernel=qs.MFMAKernel(tile=tile, k_splits=3), path_modes=qs.PathModes("never", "never"))
        assert plan.info["k_splits"] == 3 and plan.info["mixed_domain"] in (1, 2), (plan, plan.info)


@pytest.mark.parametrize("tile", [1, 4])
def test_mixed_domain_every_gather_direction(oracle, tile):
    """Complex x real and real x complex on each of the four gather
    directions (A K-fast or M-fast, B K-fast or N-fast): the two-stage
    32 x 64 tile and the one-stage 64 x 64 tile, against the oracle."""
    rng = np.random.default_rng(130 + tile)
    M, K, N = 70, 90, 50
    for TA, TB in ((np.complex128, np.float64), (np.float64, np.complex128)):
        for akf in (False, True):
            for bkf in (False, True):
                case(f"mixed domain {np.dtype(TA).name} x {np.dtype(TB).name}, A K-fast {akf}, "
                     f"B K-fast {bkf}, tile {tile}")
                Av = dev_view(randn(rng, TA, M * K), (M, K), (K, 1) if akf else (1, M))
                Bv = dev_view(randn(rng, TB, K * N), (K, N), (1, K) if bkf else (N, 1))
                Cv = dev_view(randn(rng, np.complex128, M * N), (M, N), (1, M))
                plan = run_and_check(oracle, Cv, Av, (1, 2), Bv, (2, 3), (1, 3), 1.25, 0.5 + 0.5j,
                                     conjA=akf, conjB=bkf,
                                     kernel=qs.MFMAKernel(tile=tile, k_splits=1),
                                     path_modes=qs.PathModes("never", "never"))
                assert plan.info["mixed_domain"] in (1, 2), (plan, plan.info)
