"""Batch labels (a label in A, B and C): the host-side checks that run before
any device call, and the rejection `plan_contract` keeps
(src/planning/labels.jl:48-54)."""
import numpy as np
import pytest

from helpers import dense, qs

disjoint = qs.batch._disjoint


def test_plan_contract_still_rejects_batch_labels():
    A = dense(np.ones((3, 4, 5)))
    B = dense(np.ones((4, 6, 5)))
    C = dense(np.zeros((3, 6, 5)))
    with pytest.raises(qs.ArgumentError, match="appears in indA, indB, and indC"):
        qs.plan_contract(C, A, (1, 2, 3), B, (2, 4, 3), (1, 4, 3))


def test_batch_labels_in_c_order():
    assert qs.batch_labels((1, 2, 3, 7), (2, 7, 3, 4), (7, 1, 4, 3)) == (7, 3)
    assert qs.batch_labels((1, 2), (2, 3), (1, 3)) == ()


def test_plan_batched_argument_checks():
    A = dense(np.ones((3, 4, 5)))
    B = dense(np.ones((4, 6, 5)))
    C = dense(np.zeros((3, 6, 5)))
    with pytest.raises(qs.ArgumentError, match="no label appears"):
        qs.plan_batched(dense(np.zeros((3, 6))), dense(np.ones((3, 4))), (1, 2), dense(np.ones((4, 6))), (2, 4), (1, 4))
    with pytest.raises(qs.DimensionMismatch, match="batch label 3"):
        qs.plan_batched(C, A, (1, 2, 3), dense(np.ones((4, 6, 4))), (2, 4, 3), (1, 4, 3))
    with pytest.raises(qs.ArgumentError, match="repeated label"):
        qs.plan_batched(C, A, (1, 1, 3), B, (2, 4, 3), (1, 4, 3))
    with pytest.raises(qs.ArgumentError, match="plain precision"):
        qs.plan_batched(C, A, (1, 2, 3), B, (2, 4, 3), (1, 4, 3), precision="split")
    # C slices that overlap: a zero stride along the batch axis of C
    Cz = qs.StridedView(C.parent, (3, 6, 5), (1, 3, 0), 0)
    with pytest.raises(qs.ArgumentError, match="may overlap"):
        qs.plan_batched(Cz, A, (1, 2, 3), B, (2, 4, 3), (1, 4, 3))


@pytest.mark.parametrize("sizes,strides,ok", [
    ((3, 6, 5), (1, 3, 18), True),      # column-major
    ((3, 6, 5), (30, 5, 1), True),      # row-major
    ((3, 6, 5), (1, 3, 17), False),     # last slice overlaps the previous one
    ((3, 6, 5), (1, 3, 0), False),      # zero stride
    ((3, 1, 5), (1, 0, 3), True),       # extent-1 axes do not count
    ((4, 4), (1, -4), True),            # negative strides
])
def test_disjoint_sufficient_test(sizes, strides, ok):
    assert disjoint(sizes, strides) == ok
