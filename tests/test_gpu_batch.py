"""Batch labels on the GPU: every batch coordinate's contraction in one
grouped launch, against numpy.einsum over the views' values (einsum takes a
label in all three operands as a batch label).  Tolerances as in the parity
tests: 1e-13 for FP64 / ComplexF64, 1e-5 for FP32."""
import numpy as np
import pytest
import torch

from helpers import RTOL, SV, brute_ref, qs, randn, rel_fro, values

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


def dev_view(st_np, size, strides, offset=0, op="identity"):
    return SV(torch.from_numpy(np.ascontiguousarray(st_np)).to(DEV), size, strides, offset, op)


def run_batched(Cv, Av, indA, Bv, indB, indC, alpha, beta, conjA=False, conjB=False, **kw):
    cstart = values(Cv)
    c_before = Cv.parent.cpu().numpy().copy()
    bp = qs.plan_batched(Cv, Av, indA, Bv, indB, indC, conjA=conjA, conjB=conjB, **kw)
    bp.execute(alpha, beta)
    torch.cuda.synchronize()
    got = values(Cv)
    ref = brute_ref(Av, indA, Bv, indB, cstart, indC, alpha, beta, conjA, conjB)
    err = rel_fro(got, ref)
    assert err <= RTOL[bp.plan.T], (err, bp.plan)
    # nothing outside C's view changed
    mask = np.ones(c_before.size, bool)
    idx = np.full(Cv.size, Cv.offset, dtype=np.int64)
    for d, (n, s) in enumerate(zip(Cv.size, Cv.strides)):
        shp = [1] * len(Cv.size)
        shp[d] = n
        idx = idx + (np.arange(n) * s).reshape(shp)
    mask[idx.reshape(-1)] = False
    after = Cv.parent.cpu().numpy()
    assert np.array_equal(after[mask], c_before[mask], equal_nan=True)
    return bp


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.complex128])
def test_one_batch_label(dtype):
    case(f"C[i,j,b] := A[i,k,b] * B[k,j,b], b=7, {np.dtype(dtype).name}")
    rng = np.random.default_rng(60)
    i, j, k, b = 33, 40, 50, 7
    Av = dev_view(randn(rng, dtype, i * k * b), (i, k, b), colmajor((i, k, b)))
    Bv = dev_view(randn(rng, dtype, k * j * b), (k, j, b), colmajor((k, j, b)))
    Cv = dev_view(randn(rng, dtype, i * j * b), (i, j, b), colmajor((i, j, b)))
    cplx = np.dtype(dtype).kind == "c"
    bp = run_batched(Cv, Av, (1, 2, 3), Bv, (2, 4, 3), (1, 4, 3), 1.5, (0.5 + 0.25j) if cplx else 0.5)
    assert bp.n == b and bp.batch == (3,)
    assert bp.flops == (8 if cplx else 2) * i * j * k * b


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_paired_loads_on_every_tile(dtype):
    case(f"batched paired loads on every tile: 48 x 40 x 66, b=5, {np.dtype(dtype).name}")
    rng = np.random.default_rng(66)
    i, j, k, b = 48, 40, 66, 5
    Av = dev_view(randn(rng, dtype, i * k * b), (i, k, b), colmajor((i, k, b)))
    Bv = dev_view(randn(rng, dtype, k * j * b), (k, j, b), colmajor((k, j, b)))
    c0 = randn(rng, dtype, i * j * b)
    for tile in range(6):
        Cv = dev_view(c0, (i, j, b), colmajor((i, j, b)))
        bp = run_batched(Cv, Av, (1, 2, 3), Bv, (2, 4, 3), (1, 4, 3), 1.25, 0.5, tile=tile)
        assert bp.plan.info["paired_loads"] == 1, tile
        bp.release()


@pytest.mark.parametrize("k", [3, 16, 17, 130])
@pytest.mark.parametrize("akf,bkf", [(False, False), (False, True), (True, False), (True, True)])
def test_f64_64x64_shallow_stages_every_direction(k, akf, bkf):
    # Float64 64 x 64 batched tiles run K stages of 16 at four workgroups per
    # CU: every gather direction (A or B with K as its unit-stride axis), odd
    # extents and an odd storage offset, K below, at and past one stage,
    # ragged M and N.
    case(f"batched F64 64x64 shallow stages: K={k}, A K-fast {akf}, B K-fast {bkf}")
    rng = np.random.default_rng(1000 + k)
    i, j, b = 67, 45, 4
    if akf:   # A[k,i,b] storage, K unit stride
        Av = dev_view(randn(rng, np.float64, 1 + k * i * b), (i, k, b), (k, 1, k * i), offset=1)
    else:
        Av = dev_view(randn(rng, np.float64, 1 + i * k * b), (i, k, b), colmajor((i, k, b)), offset=1)
    if bkf:
        Bv = dev_view(randn(rng, np.float64, k * j * b), (k, j, b), colmajor((k, j, b)))
    else:     # B[j,k,b] storage, J unit stride
        Bv = dev_view(randn(rng, np.float64, j * k * b), (k, j, b), (j, 1, j * k))
    Cv = dev_view(randn(rng, np.float64, i * j * b), (i, j, b), colmajor((i, j, b)))
    bp = run_batched(Cv, Av, (1, 2, 3), Bv, (2, 4, 3), (1, 4, 3), 0.75, 0.5, tile=1)
    print("[plan]", bp.plan.info.get("paired_loads"), bp.plan, flush=True)
    bp.release()


def test_two_batch_labels_permuted_conj():
    case("C[b2,j,i,b1] := conj(A[b1,i,k,b2]) * B[k,b2,j,b1], ComplexF64, strided parents")
    rng = np.random.default_rng(61)
    b1, b2, i, j, k = 3, 4, 20, 24, 36
    asz = (b1, i, k, b2)
    A_par = randn(rng, np.complex128, 2 * int(np.prod(asz)))
    Av = dev_view(A_par, asz, tuple(2 * s for s in colmajor(asz)), 1)
    bsz = (k, b2, j, b1)
    Bv = dev_view(randn(rng, np.complex128, int(np.prod(bsz))), bsz, colmajor(bsz))
    csz = (b2, j, i, b1)
    big = (b2 + 1, j, i, b1)
    Cv = dev_view(randn(rng, np.complex128, int(np.prod(big))), csz, colmajor(big), 0)
    bp = run_batched(Cv, Av, (1, 2, 3, 4), Bv, (3, 4, 5, 1), (4, 5, 2, 1), 1.0 - 0.5j, 0.3, conjA=True)
    assert bp.n == b1 * b2


def test_swapped_inner_plan():
    case("C[j,i,b] := A[i,k,b] * B[k,j,b] with C N-fast (the planner swaps A and B), F64")
    rng = np.random.default_rng(62)
    i, j, k, b = 48, 64, 40, 5
    Av = dev_view(randn(rng, np.float64, i * k * b), (i, k, b), colmajor((i, k, b)))
    Bv = dev_view(randn(rng, np.float64, k * j * b), (k, j, b), colmajor((k, j, b)))
    Cv = dev_view(randn(rng, np.float64, j * i * b), (j, i, b), colmajor((j, i, b)))
    bp = run_batched(Cv, Av, (1, 2, 3), Bv, (2, 4, 3), (4, 1, 3), 1.0, 1.0)
    assert bp.n == b


def test_broadcast_operand_along_batch():
    case("A shared by every batch coordinate (zero batch stride), F64")
    rng = np.random.default_rng(63)
    i, j, k, b = 40, 30, 24, 6
    Av = dev_view(randn(rng, np.float64, i * k), (i, k, b), (1, i, 0))
    Bv = dev_view(randn(rng, np.float64, k * j * b), (k, j, b), colmajor((k, j, b)))
    Cv = dev_view(np.zeros(i * j * b), (i, j, b), colmajor((i, j, b)))
    run_batched(Cv, Av, (1, 2, 3), Bv, (2, 4, 3), (1, 4, 3), 1.0, 0.0)


def test_many_small_batches_and_graph_replay():
    case("1000 batches of 16 x 16 x 16, F64, captured graph replayed 3 times")
    rng = np.random.default_rng(64)
    i = j = k = 16
    b = 1000
    Av = dev_view(randn(rng, np.float64, i * k * b), (i, k, b), colmajor((i, k, b)))
    Bv = dev_view(randn(rng, np.float64, k * j * b), (k, j, b), colmajor((k, j, b)))
    Cv = dev_view(np.zeros(i * j * b), (i, j, b), colmajor((i, j, b)))
    bp = run_batched(Cv, Av, (1, 2, 3), Bv, (2, 4, 3), (1, 4, 3), 1.0, 0.0)
    want = values(Cv).copy()
    # The graph reads the coefficients the group last uploaded, so the eager
    # executes between the replays use the same ones.
    g = bp.capture(repeats=1, alpha=2.0, beta=0.0)
    for _ in range(3):
        Cv.parent.zero_()
        g.launch()
        torch.cuda.synchronize()
        assert rel_fro(values(Cv), 2.0 * want) <= 1e-13
        Cv.parent.zero_()
        bp.execute(2.0, 0.0)   # eager work between the replays
        torch.cuda.synchronize()
        assert rel_fro(values(Cv), 2.0 * want) <= 1e-13
    g.release()
    bp.release()


def test_batch_spans_checked_before_any_write():
    case("a batch coordinate past B's storage is rejected before anything is written")
    rng = np.random.default_rng(65)
    i, j, k, b = 8, 8, 8, 4
    Av = dev_view(randn(rng, np.float64, i * k * b), (i, k, b), colmajor((i, k, b)))
    Bst = randn(rng, np.float64, k * j * b)
    Bv = dev_view(Bst, (k, j, b), (1, k, k * j + 1))   # the last batch slice runs past the storage
    Cv = dev_view(np.zeros(i * j * b), (i, j, b), colmajor((i, j, b)))
    # every coordinate's spans are checked when the batched group is built
    with pytest.raises(qs.BoundsError):
        qs.plan_batched(Cv, Av, (1, 2, 3), Bv, (2, 4, 3), (1, 4, 3))
    torch.cuda.synchronize()
    assert np.all(values(Cv) == 0)


def test_autotune_and_per_coordinate_coefficients():
    case("C[i,j,b] := alpha_b A[i,k,b] B[k,j,b] + beta_b C[i,j,b], b=9, F64: every tile, then per-coordinate coefficients")
    rng = np.random.default_rng(66)
    i, j, k, b = 70, 90, 45, 9
    A = randn(rng, np.float64, i * k * b).reshape((i, k, b), order="F")
    B = randn(rng, np.float64, k * j * b).reshape((k, j, b), order="F")
    C0 = randn(rng, np.float64, i * j * b).reshape((i, j, b), order="F")
    Av = dev_view(A.ravel(order="F"), (i, k, b), colmajor((i, k, b)))
    Bv = dev_view(B.ravel(order="F"), (k, j, b), colmajor((k, j, b)))
    Cv = dev_view(C0.ravel(order="F"), (i, j, b), colmajor((i, j, b)))
    prod = np.einsum("ikb,kjb->ijb", A, B)
    bp = qs.plan_batched(Cv, Av, (1, 2, 3), Bv, (2, 4, 3), (1, 4, 3))
    times = bp.autotune(reps=2)
    assert set(times) == {-1, 0, 1, 2, 3, 4, 5} and bp.tile in times
    assert np.array_equal(values(Cv), C0)   # autotune restores C
    bp.execute(1.0, 0.5)                    # one alpha and beta: the batched launch
    torch.cuda.synchronize()
    assert rel_fro(values(Cv), prod + 0.5 * C0) <= 1e-13
    Cv.parent.copy_(torch.from_numpy(C0.ravel(order="F")))
    alphas = [1.0 + 0.25 * t for t in range(b)]
    betas = [0.1 * t for t in range(b)]
    bp.execute(alphas, betas)               # per-coordinate coefficients: the grouped launch
    torch.cuda.synchronize()
    want = prod * np.array(alphas) + C0 * np.array(betas)
    assert rel_fro(values(Cv), want) <= 1e-13
    bp.release()
