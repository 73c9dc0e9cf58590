"""The host planner (labels, groups, conjugation, swap, compute type) against
the reference's known answers (test/planning/test_labels.jl,
test/planning/test_plan.jl, test/layout/test_axis_group.jl,
src/integrations/tensoroperations.jl:37-38).  CPU only."""
import random

import numpy as np
import pytest
import torch

from helpers import SV, qs


def lo_view(size, strides):
    """A StridedView of the given layout over a big enough storage (_lo_view)."""
    n = 1 + sum((s - 1) * abs(st) for s, st in zip(size, strides))
    off = sum((s - 1) * -st for s, st in zip(size, strides) if st < 0)
    return SV(torch.zeros(max(n, 1), dtype=torch.float64), size, strides, off)


def col_major(size, dtype=torch.float64):
    strides, s = [], 1
    for n in size:
        strides.append(s)
        s *= n
    return SV(torch.zeros(max(s, 1), dtype=dtype), size, tuple(strides), 0)


def test_classify_labels():
    cases = (
        ((1, 2), (2, 3), (1, 3), ((1,), (3,), (2,))),
        ((1, 2), (2, 3), (3, 1), ((1,), (3,), (2,))),
        ((1, 2, 3, 4), (3, 5, 6, 7), (4, 6, 7, 1, 2, 5), ((1, 2, 4), (5, 6, 7), (3,))),
        ((1, 2), (3, 1, 4, 5), (3, 2, 4, 5), ((2,), (3, 4, 5), (1,))),
        ((1, 2, 3, 4), (3, 4, 5, 6), (1, 2, 5, 6), ((1, 2), (5, 6), (3, 4))),
        ((1,), (1, 2), (2,), ((), (2,), (1,))),
        ((1, 2), (3, 4), (1, 2, 3, 4), ((1, 2), (3, 4), ())),
    )
    for indA, indB, indC, want in cases:
        got = qs.classify_labels(indA, indB, indC)
        assert got == want
        assert tuple(map(len, got)) == qs.group_ranks(len(indA), len(indB), len(indC))
    for bad in (((1, 1), (1, 2), (1, 2)), ((1, 2), (2, 3), (1, 2, 3)), ((1, 2), (3, 4), (1, 3)),
                ((1, 2), (3, 4), (1, 2, 3, 9))):
        with pytest.raises(qs.ArgumentError):
            qs.classify_labels(*bad)


def test_order_free_labels():
    Cv = lo_view((5, 3, 2, 4), (12, 1, 60, 3))
    indC = (10, 20, 30, 40)
    assert qs.order_free_labels((10, 20, 30, 40), indC, Cv) == (20, 40, 10, 30)
    assert qs.order_free_labels((40, 10), indC, Cv) == (40, 10)
    assert qs.order_free_labels((30, 20), indC, Cv) == (20, 30)
    Ct = lo_view((2, 3, 4), (1, 1, 1))
    assert qs.order_free_labels((7, 8, 9), (7, 8, 9), Ct) == (7, 8, 9)
    assert qs.order_free_labels((9, 7, 8), (7, 8, 9), Ct) == (9, 7, 8)
    assert qs.order_free_labels((20, 10), (10, 20), lo_view((3, 4), (-1, 4))) == (10, 20)
    assert qs.order_free_labels((10, 20), (10, 20), lo_view((3, 4), (4, -1))) == (20, 10)
    assert qs.order_free_labels((), indC, Cv) == ()
    assert qs.order_free_labels((30,), indC, Cv) == (30,)
    rnd = random.Random(20260921)
    for _ in range(100):
        nd = rnd.randint(1, 5)
        C = col_major(tuple(rnd.randint(1, 4) for _ in range(nd)))
        labels = rnd.sample(list(range(1, nd + 1)), rnd.randint(1, nd))
        got = qs.order_free_labels(tuple(labels), tuple(range(1, nd + 1)), C)
        want = [l for _, l in sorted(((abs(C.strides[l - 1]), i), l) for i, l in enumerate(labels))]
        assert list(got) == want


def test_leading_unit_run_and_prefer_swap():
    d = 4
    C6 = col_major((d,) * 6)
    indC = (1, 2, 3, 4, 5, 6)
    a, b, c, i, j, k = indC
    run = qs.leading_unit_run
    assert run((a, i, j), indC, C6) == d
    assert run((a, b, k), indC, C6) == d ** 2
    assert run((a, b, c, i, j, k), indC, C6) == d ** 6
    assert run((b, i, j), indC, C6) == 1
    assert run((a, c, b), indC, C6) == d
    assert run((), indC, C6) == 1
    assert run((10, 20), (10, 20), lo_view((3, 4), (-1, 3))) == 1
    assert run((10, 20), (10, 20), lo_view((1, 6), (5, 1))) == 6
    assert run((10, 20), (10, 20), lo_view((6, 1), (1, 17))) == 6
    assert run((10, 20), (10, 20), lo_view((0, 6), (1, 1))) == 0

    def swap(m, n, mt, mt2=None):
        return qs.prefer_swap(run(m, indC, C6), run(n, indC, C6), mt, mt2)

    assert swap((b, i, j), (a, c, k), 4)
    assert not swap((b, i, j), (a, c, k), 8)
    assert swap((c, i, j), (a, b, k), 8)
    assert swap((c, i, j), (a, b, k), 16)
    assert not swap((c, i, j), (a, b, k), 32)
    assert not swap((a, i, j), (b, c, k), 4)
    assert not swap((a, i, j), (b, c, k), 8)
    assert swap((b, i, j), (a, c, k), 8, 4)
    assert not swap((b, i, j), (a, c, k), 8, 8)
    assert not swap((a, i, j), (b, c, k), 4, 4)
    assert not qs.prefer_swap(run((b, i, j), indC, C6), run((), indC, C6), 1)


def test_leading_unit_run_is_single_map_ramp():
    # test/planning/test_labels.jl:97-124
    rng = random.Random(0x01E3A3E1)
    tested = 0
    while tested < 300:
        D = rng.randint(1, 5)
        lens = tuple(rng.choice((1, 2, 3, 5)) for _ in range(D))
        strides = tuple(rng.choice((-3, -1, 1, 2, 3, 7)) for _ in range(D))
        Cr = lo_view(lens, strides)
        order = tuple(rng.sample(list(range(1, D + 1)), rng.randint(1, D)))
        m_length = int(np.prod([lens[l - 1] for l in order]))
        if m_length == 1:
            continue
        tested += 1
        r = qs.leading_unit_run(order, tuple(range(1, D + 1)), Cr)
        g = qs.AxisGroup(tuple(lens[l - 1] for l in order), (tuple(strides[l - 1] for l in order),))
        isramp, steps = qs.affine_ramp(g)
        assert (r == m_length) == (isramp and steps[0] == 1)


def _ko_order(Av, indA, Bv, indB, Cv, indC, l2bytes):
    m, n, k = qs.classify_labels(indA, indB, indC)
    mo = qs.order_free_labels(m, indC, Cv)
    no = qs.order_free_labels(n, indC, Cv)
    ext = lambda ls: int(np.prod([Cv.size[indC.index(l)] for l in ls])) if ls else 1
    return qs.order_contract_labels(k, indA, Av, mo, indB, Bv, no, ext(m), ext(n), l2bytes=l2bytes)


def test_k_order_cost_model_known_answers():
    # test/planning/test_labels.jl:314-347 (host-independent at explicit l2bytes)
    ext = {1: 9, 2: 3, 3: 4, 4: 5, 5: 6}
    arr = lambda ind: col_major(tuple(ext[l] for l in ind))
    Av = arr((1, 2, 3, 4))
    Cv = col_major((ext[1], ext[5]))
    Bs = arr((4, 3, 2, 5))
    assert _ko_order(Av, (1, 2, 3, 4), Bs, (4, 3, 2, 5), Cv, (1, 5), 0) == (4, 3, 2)
    assert _ko_order(Av, (1, 2, 3, 4), Bs, (4, 3, 2, 5), Cv, (1, 5), 1 << 20) == (2, 3, 4)
    for indB in ((2, 3, 4, 5), (2, 5, 3, 4)):
        assert _ko_order(Av, (1, 2, 3, 4), arr(indB), indB, Cv, (1, 5), 0) == (2, 3, 4)
    At = col_major((8,) * 5)
    Bt = col_major((8,) * 3)
    Ct = col_major((8,) * 4)
    for l2 in (0, 1 << 20):
        assert _ko_order(At, (1, 2, 3, 4, 5), Bt, (5, 6, 2), Ct, (1, 3, 4, 6), l2) == (2, 5)
    A1 = col_major((9, 1, 4, 5))
    B1 = col_major((1, 5, 4, 6))
    o = _ko_order(A1, (1, 2, 3, 4), B1, (2, 4, 3, 5), Cv, (1, 5), 0)
    assert sorted(o) == [2, 3, 4]
    assert tuple(x for x in o if x != 2) == (4, 3)


def test_axis_group_validation_and_label_ctor():
    # test/layout/test_axis_group.jl:88-97, 146-153
    big = (1 << 63) - 1
    with pytest.raises(qs.ArgumentError):
        qs.AxisGroup((-1,), ((1,),))
    with pytest.raises(qs.ArgumentError):
        qs.AxisGroup((3,), ())
    with pytest.raises(OverflowError):
        qs.AxisGroup((big // 2 + 2, big // 2 + 2), ((1, 1),))
    with pytest.raises(OverflowError):
        qs.AxisGroup((2,), ((-(1 << 63),),))
    assert qs.AxisGroup((0,), ((big,),)).axis_length() == 0
    assert qs.AxisGroup((2,), ((big // 2,),)).axis_length() == 2
    from quasistrided_amd.labels import group_from_labels
    A, B, C = col_major((3, 5, 2)), col_major((5, 4)), col_major((3, 4, 2))
    M = group_from_labels((1, 3), ((1, 2, 3), A), ((1, 4, 3), C))
    assert (M.lengths, M.strides) == ((3, 2), ((1, 15), (1, 12)))
    K = group_from_labels((2,), ((1, 2, 3), A), ((2, 4), B))
    assert (K.lengths, K.strides) == ((5,), ((3,), (1,)))
    with pytest.raises(qs.DimensionMismatch):
        group_from_labels((2,), ((1, 2, 3), A), ((2, 4), col_major((6, 4))))


def test_affine_ramp_python():
    ar = qs.affine_ramp
    G = qs.AxisGroup
    assert ar(G((), ((), ()))) == (True, (0, 0))
    assert ar(G((7,), ((3,), (-2,)))) == (True, (3, -2))
    assert ar(G((1, 7, 1), ((99, 3, -4), (5, -2, 8)))) == (True, (3, -2))
    assert ar(G((4, 5), ((1, 4), (2, 8)))) == (True, (1, 2))
    assert ar(G((4, 5), ((1, 4), (2, 9))))[0] is False
    assert ar(G((2, 3, 4), ((1, 2, 6), (5, 10, 30)))) == (True, (1, 5))
    assert ar(G((0, 3), ((1, 4), (1, 4))))[0] is True


def test_contraction_labels_example():
    # tensoroperations.jl:37-38
    pA, pB, pAB = ((3, 1, 4), (2, 5)), ((3, 1), (2, 4)), ((4, 2), (5, 1, 3))
    assert qs.contraction_labels(pA, pB, pAB) == ((2, -1, 1, 3, -2), (-2, 4, -1, 5), (4, 2, 5, 1, 3))


def test_compute_type_rules():
    f32, f64, c32, c64 = torch.float32, torch.float64, torch.complex64, torch.complex128
    ct = qs.compute_type
    assert ct(f32, f64, f64) == f64
    assert ct(f32, f32, f32) == f32
    assert ct(f64, c64, c64) == c64
    assert ct(c32, f64, c64) == c64
    assert ct(f32, f32, f32, torch.float64) == f64
    assert ct(c64, f32, c64, torch.float32) == c32
    assert ct(f64, f64, c64) == c64
    with pytest.raises(qs.ArgumentError):
        ct(c64, f64, f64)
    with pytest.raises(qs.ArgumentError):
        ct(torch.int32, f64, f64)


def test_op_conjugates_and_isconj_table():
    # test/planning/test_plan.jl:55-73
    assert qs.op_conjugates("identity") is False
    assert qs.op_conjugates("conj") is True
    assert qs.op_conjugates("transpose") is False
    assert qs.op_conjugates("adjoint") is True
    with pytest.raises(qs.ArgumentError):
        qs.op_conjugates("sin")
    for op in ("identity", "conj", "transpose", "adjoint"):
        for flag in (False, True):
            v = SV(torch.zeros(16, dtype=torch.float64), (4, 4), (1, 4), 0, op)
            assert qs.isconj(v, flag) is False
            vc = SV(torch.zeros(16, dtype=torch.complex128), (4, 4), (1, 4), 0, op)
            assert qs.isconj(vc, flag) == (flag ^ (op in ("conj", "adjoint")))


def test_contraction_labels_roundtrip_random():
    rng = random.Random(7)
    for _ in range(50):
        NoA, Nk, NoB = rng.randint(0, 3), rng.randint(0, 3), rng.randint(0, 3)
        pa = list(range(1, NoA + Nk + 1))
        rng.shuffle(pa)
        pb = list(range(1, Nk + NoB + 1))
        rng.shuffle(pb)
        pc = list(range(1, NoA + NoB + 1))
        rng.shuffle(pc)
        pA, pB, pAB = (tuple(pa[:NoA]), tuple(pa[NoA:])), (tuple(pb[:Nk]), tuple(pb[Nk:])), (tuple(pc), ())
        indA, indB, indC = qs.contraction_labels(pA, pB, pAB)
        m, n, k = qs.classify_labels(indA, indB, indC)
        assert (len(m), len(n), len(k)) == (NoA, NoB, Nk)
        # contracted axes pair up: A's pA[1][i] carries the label of B's pB[0][i]
        for i in range(Nk):
            assert indA[pA[1][i] - 1] == indB[pB[0][i] - 1] < 0
