"""Pins of the oracle (oracle/, the CPU restatement) against the reference's
known-answer tests, closed forms, numpy.einsum and brute force.  CPU only."""
import itertools

import numpy as np
import pytest

from helpers import randn

F32, F64, C32, C64 = 0, 1, 2, 3


def brute_offsets(lengths, strides):
    """CartesianIndices oracle (test/layout/test_axis_group.jl:8-11): first dim fastest."""
    out = []
    ranges = [range(L) for L in lengths]
    for idx in itertools.product(*reversed(ranges)):
        c = list(reversed(idx))
        out.append(tuple(sum(c[d] * s[d] for d in range(len(lengths))) for s in strides))
    return out


def test_fill_offsets_worked_example(oracle):
    # test/layout/test_axis_group.jl:14-28
    M = ((3, 2), [(1, 15), (1, 12)])
    N = ((4,), [(5,), (3,)])
    K = ((5,), [(3,), (1,)])
    for (lengths, strides), e1, e2 in ((M, [0, 1, 2, 15, 16, 17], [0, 1, 2, 12, 13, 14]),
                                      (N, [0, 5, 10, 15], [0, 3, 6, 9]),
                                      (K, [0, 3, 6, 9, 12], [0, 1, 2, 3, 4])):
        Q = int(np.prod(lengths))
        r, (b1, b2) = oracle.fill_offsets(lengths, strides, 0, Q)
        assert r == 0 and b1 == e1 and b2 == e2


def test_fill_offsets_randomized_vs_brute_force(oracle):
    # test/layout/test_axis_group.jl:42-73
    rng = np.random.default_rng(0xA51CE1)
    for _ in range(200):
        D, P = int(rng.integers(0, 4)), int(rng.integers(1, 4))
        lengths = tuple(int(x) for x in rng.integers(0, 5, D))
        strides = [tuple(int(x) for x in rng.integers(-8, 9, D)) for _ in range(P)]
        full = brute_offsets(lengths, strides)
        Q = len(full)
        assert Q == int(np.prod(lengths)) if D else Q == 1
        isramp, steps = oracle.affine_ramp(lengths, strides)
        if Q > 0:
            step1 = full[1] if Q > 1 else tuple([0] * P)
            assert isramp == (full == [tuple(s * q for s in step1) for q in range(Q)])
            if isramp:
                assert steps == step1
        for _ in range(4):
            first = int(rng.integers(0, Q + 1))
            count = int(rng.integers(0, Q - first + 1))
            r, bufs = oracle.fill_offsets(lengths, strides, first, count)
            assert r == 0
            for p in range(P):
                assert bufs[p] == [full[i][p] for i in range(first, first + count)]


def test_fill_offsets_invalid_requests(oracle):
    # test/layout/test_axis_group.jl:99-112: 1 ArgumentError, 3 BoundsError
    lengths, strides = (3, 2), [(1, 10), (2, 20)]
    Q = 6
    for first, count, err in ((0, -1, 1), (Q, 1, 3), (-1, 1, 3), (1, Q, 3)):
        r, _ = oracle.fill_offsets(lengths, strides, first, count)
        assert r == err
    r, _ = oracle.fill_offsets(lengths, strides, Q, 0)
    assert r == 0


def test_validate_group_overflow_policy(oracle):
    # test/layout/test_axis_group.jl:88-97: 1 ArgumentError, 4 OverflowError
    big = (1 << 63) - 1
    assert oracle.validate_group((-1,), [(1,)]) == 1
    assert oracle.validate_group((3, -2), [(1, 1)]) == 1
    assert oracle.validate_group((3,), []) == 1
    assert oracle.validate_group((big // 2 + 2, big // 2 + 2), [(1, 1)]) == 4
    assert oracle.validate_group((big,), [(2,)]) == 4
    assert oracle.validate_group((2,), [(-(1 << 63),)]) == 4
    assert oracle.validate_group((0,), [(big,)]) == 0
    assert oracle.validate_group((2,), [(big // 2,)]) == 0
    assert oracle.validate_group((big // 2, big // 2, 0), [(1, 1, 1)]) == 0


def test_describe_block_edge_cases(oracle):
    # test/layout/test_axis_group.jl:114-127
    imin, imax = -(1 << 63), (1 << 63) - 1
    assert oracle.describe_block([], 0, 0)[1] == (0, 0, 0, True)
    assert oracle.describe_block([42], 0, 1)[1] == (42, 0, 1, True)
    assert oracle.describe_block([5, 8, 11, 14], 0, 4)[1] == (5, 3, 4, True)
    assert oracle.describe_block([0, 5, 1, 6], 0, 4)[1] == (0, 0, 4, False)
    assert oracle.describe_block([imin, imax], 0, 2)[1] == (imin, 0, 2, False)
    assert oracle.describe_block([1, 2, 42, 3], 2, 1)[1] == (42, 0, 1, True)
    assert oracle.describe_block(list(range(0, 21, 2)), 3, 4)[1] == (6, 2, 4, True)
    assert oracle.describe_block(list(range(1, 11)), 10, 0)[1] == (0, 0, 0, True)
    assert oracle.describe_block([1, 2, 3], 0, -1)[0] == 1
    assert oracle.describe_block([1, 2], 0, 3)[0] == 2
    assert oracle.describe_block([1, 2, 3], -1, 1)[0] == 1
    assert oracle.describe_block([1, 2, 3], 2, 2)[0] == 2


def test_affine_ramp_cases(oracle):
    # test/layout/test_axis_group.jl:197-213
    ar = oracle.affine_ramp
    assert ar((), [(), ()]) == (True, (0, 0))
    assert ar((7,), [(3,), (-2,)]) == (True, (3, -2))
    assert ar((1, 7, 1), [(99, 3, -4), (5, -2, 8)]) == (True, (3, -2))
    assert ar((4, 5), [(1, 4), (2, 8)]) == (True, (1, 2))
    assert ar((4, 5), [(1, 4), (2, 9)])[0] is False
    assert ar((16, 16, 16), [(1, 256, 4096), (1, 256, 4096)])[0] is False
    assert ar((2, 3, 4), [(1, 2, 6), (5, 10, 30)]) == (True, (1, 5))
    assert ar((2, 3, 4), [(1, 2, 4), (5, 10, 30)])[0] is False
    assert ar((0, 3), [(1, 4), (1, 4)])[0] is True


def _worked():
    """_worked_fixture (test/helpers.jl:49-57): C[a,n,b] = sum_k A[a,k,b] B[k,n]."""
    A = np.arange(1.0, 31.0).reshape(3, 5, 2, order="F")
    B = np.arange(1.0, 21.0).reshape(5, 4, order="F")
    Cref = np.zeros((3, 4, 2))
    for b in range(2):
        for n in range(4):
            for a in range(3):
                Cref[a, n, b] = sum(A[a, k, b] * B[k, n] for k in range(5))
    return A, B, Cref


WORKED_GROUPS = (((3, 2), (1, 15), (1, 12)),   # M (A, C)
                 ((4,), (5,), (3,)),           # N (B, C)
                 ((5,), (3,), (1,)))           # K (A, B)


@pytest.mark.parametrize("blis", [False, True])
def test_worked_fixture_exact(oracle, blis):
    A, B, Cref = _worked()
    a = A.reshape(-1, order="F").copy()
    b = B.reshape(-1, order="F").copy()
    c = np.full(24, np.nan)
    st = oracle.contract(a, 0, b, 0, c, 0, *WORKED_GROUPS, F64, 1.0, 0.0, blis=blis)
    assert st == 0
    assert np.array_equal(c.reshape(3, 4, 2, order="F"), Cref)   # integers: exact


def test_worked_fixture_closed_form():
    # The closed form sum_k (a + 3k + 15b + 1)(k + 5n + 1) over k = 0..4 of
    # the 1:30 / 1:20 fixtures, independent of any loop above.
    A, B, Cref = _worked()
    for a, n, b in itertools.product(range(3), range(4), range(2)):
        x0 = a + 15 * b + 1
        y0 = 5 * n + 1
        val = sum((x0 + 3 * k) * (y0 + k) for k in range(5))
        assert Cref[a, n, b] == val


def _strided_case(rng, dtype, conj_a, conj_b):
    """Permuted A with a zero stride, negative-stride B, offset C (scattered_fixture, test/helpers.jl:75-82)."""
    an, kn, bn, nn = 6, 7, 3, 5
    A2 = randn(rng, dtype, an * kn)                       # A[k, b, a] = A2[a, k]
    Bst = randn(rng, dtype, kn * nn)                      # B[k, n] at (kn-1) - k + kn n
    Cst = randn(rng, dtype, (an + 2) * (nn + 3) * (bn + 1))
    # Labels: A (k=2, b=3, a=1) -> indA = (2, 3, 1); B (k, n) = (2, 4); C (a, n, b) = (1, 4, 3)
    # M = (a, b): A strides (1, 0); C strides (1, (an+2)(nn+3))
    # N = (n,): B stride kn; C stride an+2
    # K = (k,): A stride an; B stride -1
    ldc = an + 2
    cplane = (an + 2) * (nn + 3)
    mg = ((an, bn), (1, 0), (1, cplane))
    ng = ((nn,), (kn,), (ldc,))
    kg = ((kn,), (an,), (-1,))
    c_base = 1 + ldc
    b_base = kn - 1
    # brute force
    ref = np.zeros((an, nn, bn), dtype=np.complex128)
    for a_, n_, b_ in itertools.product(range(an), range(nn), range(bn)):
        s = 0
        for k in range(kn):
            x = complex(A2[a_ + an * k])
            y = complex(Bst[b_base - k + kn * n_])
            if conj_a:
                x = x.conjugate()
            if conj_b:
                y = y.conjugate()
            s += x * y
        ref[a_, n_, b_] = s
    return A2, Bst, Cst, mg, ng, kg, c_base, b_base, ref, (an, nn, bn), ldc, cplane


@pytest.mark.parametrize("dtype,code", [(np.float64, F64), (np.float32, F32), (np.complex128, C64),
                                        (np.complex64, C32)])
@pytest.mark.parametrize("beta", [0.0, 0.5, 1.0])
def test_scattered_against_brute_force(oracle, dtype, code, beta):
    rng = np.random.default_rng(11)
    cplx = np.dtype(dtype).kind == "c"
    conj_a, conj_b = (True, False) if cplx else (False, False)
    A2, Bst, Cst, mg, ng, kg, c_base, b_base, ref, (an, nn, bn), ldc, cplane = _strided_case(rng, dtype, conj_a, conj_b)
    alpha = (1.3 - 0.4j) if cplx else 1.3
    beta_v = (beta + 0.2j) if (cplx and beta not in (0.0, 1.0)) else beta
    c0 = Cst.copy()
    for blis in (False, True):
        if blis and conj_a is None:
            continue
        c = c0.copy()
        st = oracle.contract(A2, 0, Bst, b_base, c, c_base, mg, ng, kg, code, alpha, beta_v, conj_a=conj_a,
                             conj_b=conj_b, blis=blis)
        assert st == 0
        got = np.zeros((an, nn, bn), dtype=np.complex128)
        want = np.zeros_like(got)
        touched = np.zeros(c.size, bool)
        for a_, n_, b_ in itertools.product(range(an), range(nn), range(bn)):
            idx = c_base + a_ + ldc * n_ + cplane * b_
            touched[idx] = True
            got[a_, n_, b_] = c[idx]
            want[a_, n_, b_] = alpha * ref[a_, n_, b_] + (beta_v * complex(c0[idx]) if beta != 0 else 0)
        tol = 1e-13 if np.dtype(dtype).itemsize >= 8 and dtype != np.complex64 else 1e-5
        assert np.linalg.norm(got - want) <= tol * np.linalg.norm(want)
        assert np.array_equal(c[~touched], c0[~touched])   # nothing outside C's view is written


def test_beta_zero_never_reads_c_and_alpha_zero_never_reads_ab(oracle):
    # test/execution/test_execute.jl:25-36
    M, K, N = 5, 6, 4
    mg, ng, kg = ((M,), (1,), (1,)), ((N,), (K,), (M,)), ((K,), (M,), (1,))
    rng = np.random.default_rng(2)
    a, b = rng.standard_normal(M * K), rng.standard_normal(K * N)
    c = np.full(M * N, np.nan)
    assert oracle.contract(a, 0, b, 0, c, 0, mg, ng, kg, F64, 2.5, 0.0) == 0
    assert np.all(np.isfinite(c))
    an, bn = np.full(M * K, np.nan), np.full(K * N, np.nan)
    c0 = rng.standard_normal(M * N)
    for beta in (0.5, 0.0):
        c = c0.copy()
        assert oracle.contract(an, 0, bn, 0, c, 0, mg, ng, kg, F64, 0.0, beta) == 0
        assert np.array_equal(c, beta * c0)
    # K = 0: C *= beta
    c = c0.copy()
    assert oracle.contract(an, 0, bn, 0, c, 0, mg, ng, ((0,), (M,), (1,)), F64, 1.0, 0.5) == 0
    assert np.array_equal(c, 0.5 * c0)


def test_out_of_bounds_rejected_before_any_write(oracle):
    # test/execution/test_execute.jl:57-70: each operand one element short
    M, K, N = 8, 6, 5
    mg, ng, kg = ((M,), (1,), (1,)), ((N,), (K,), (M,)), ((K,), (M,), (1,))
    for sa, sb, sc in ((M * K - 1, K * N, M * N), (M * K, K * N - 1, M * N), (M * K, K * N, M * N - 1)):
        a, b, c = np.ones(sa), np.ones(sb), np.zeros(sc)
        assert oracle.contract(a, 0, b, 0, c, 0, mg, ng, kg, F64, 1.0, 0.0) == 3
        assert np.all(c == 0)


@pytest.mark.parametrize("TA,TC,code,beta", [(np.float64, np.float32, F64, 0.5), (np.float32, np.float32, F64, 1.0),
                                             (np.float64, np.float32, F64, 0.0)])
def test_one_rounding_across_k(oracle, TA, TC, code, beta):
    # test/execution/test_c_panel.jl:5-41: 1, 2^-30, -1 at K positions 0, K/2, K-1;
    # the sum survives only when accumulated in the compute type and rounded once.
    M, N, K = 13, 9, 20
    A = np.zeros((M, K), dtype=TA)
    A[:, 0] = 1
    A[:, K // 2] = 2.0 ** -30
    A[:, K - 1] = -1
    a = A.reshape(-1, order="F").copy()
    b = np.ones(K * N, dtype=TA)
    c = np.full(M * N, np.nan if beta == 0 else 2.0 ** -28, dtype=TC)
    mg, ng, kg = ((M,), (1,), (1,)), ((N,), (K,), (M,)), ((K,), (M,), (1,))
    assert oracle.contract(a, 0, b, 0, c, 0, mg, ng, kg, code, 1.0, beta) == 0
    want = TC(2.0 ** -30 + beta * 2.0 ** -28)
    assert np.all(c == want)


@pytest.mark.parametrize("dtype,code", [(np.float64, F64), (np.complex128, C64), (np.float32, F32),
                                        (np.complex64, C32)])
def test_dense_against_einsum(oracle, dtype, code):
    # Random dense GEMMs against numpy.einsum in float64/complex128.
    rng = np.random.default_rng(5)
    for M, K, N in ((1, 70, 29), (37, 1, 21), (23, 37, 19), (64, 64, 64), (100, 7, 3)):
        A = randn(rng, dtype, M * K)
        B = randn(rng, dtype, K * N)
        c0 = randn(rng, dtype, M * N)
        mg, ng, kg = ((M,), (1,), (1,)), ((N,), (K,), (M,)), ((K,), (M,), (1,))
        cplx = np.dtype(dtype).kind == "c"
        alpha, beta = (0.5 + 1.5j, -1.0 + 0.25j) if cplx else (1.5, -0.75)
        ref = alpha * (A.reshape(M, K, order="F").astype(np.complex128) @ B.reshape(K, N, order="F")) + \
            beta * c0.reshape(M, N, order="F")
        for blis in (False, True):
            c = c0.copy()
            assert oracle.contract(A, 0, B, 0, c, 0, mg, ng, kg, code, alpha, beta, blis=blis) == 0
            tol = 1e-13 if dtype in (np.float64, np.complex128) else 2e-5
            assert np.linalg.norm(c.reshape(M, N, order="F") - ref) <= tol * np.linalg.norm(ref)


def test_mixed_storage_types_convert_on_load(oracle):
    # F32 A and F64 B into F64 C, compute F64: products of the exactly promoted inputs.
    rng = np.random.default_rng(9)
    M, K, N = 7, 11, 5
    A = rng.standard_normal(M * K).astype(np.float32)
    B = rng.standard_normal(K * N)
    c = np.zeros(M * N)
    mg, ng, kg = ((M,), (1,), (1,)), ((N,), (K,), (M,)), ((K,), (M,), (1,))
    assert oracle.contract(A, 0, B, 0, c, 0, mg, ng, kg, F64, 1.0, 0.0) == 0
    ref = A.astype(np.float64).reshape(M, K, order="F") @ B.reshape(K, N, order="F")
    assert np.linalg.norm(c.reshape(M, N, order="F") - ref) <= 1e-14 * np.linalg.norm(ref)


def test_blis_kernel_name(oracle):
    name = oracle.blis_kernel(F64)
    assert ("avx512" in name or "avx2" in name) and "x6" in name
