"""tensoradd! / tensortrace! argument checks, which run on the host before
anything reaches the device (TensorOperations' argcheck / dimcheck rules)."""
import numpy as np
import pytest
import torch

from helpers import qs


def t(*shape, dtype=torch.float64):
    return torch.zeros(*shape, dtype=dtype)


def test_tensoradd_invalid_permutation():
    with pytest.raises(qs.ArgumentError):
        qs.tensoradd(t(3, 4), t(4, 3), ((1, 1), ()), False, 1.0, 0.0)
    with pytest.raises(qs.ArgumentError):
        qs.tensoradd(t(3, 4, 1), t(4, 3), ((2, 1), ()), False, 1.0, 0.0)


def test_tensoradd_dimension_mismatch():
    with pytest.raises(qs.DimensionMismatch):
        qs.tensoradd(t(3, 4), t(3, 4), ((2, 1), ()), False, 1.0, 0.0)


def test_tensortrace_checks():
    with pytest.raises(qs.DimensionMismatch):   # traced axes differ in length
        qs.tensortrace(t(4), t(3, 4, 5), ((2,), ()), ((1,), (3,)), False, 1.0, 0.0)
    with pytest.raises(qs.ArgumentError):       # pairs of unequal count
        qs.tensortrace(t(4), t(3, 4, 3), ((2,), ()), ((1, 3), (3,)), False, 1.0, 0.0)
    with pytest.raises(qs.ArgumentError):       # an axis both kept and traced
        qs.tensortrace(t(3), t(3, 4, 3), ((1,), ()), ((1,), (3,)), False, 1.0, 0.0)


def test_tensoradd_domain_alias_and_conj_output():
    with pytest.raises(qs.ArgumentError):       # complex A into a real C
        qs.tensoradd(t(3, 4), t(4, 3, dtype=torch.complex128), ((2, 1), ()), False, 1.0, 0.0)
    x = t(4, 4)
    with pytest.raises(qs.ArgumentError):       # C aliases A
        qs.tensoradd(x, x.t(), ((2, 1), ()), False, 1.0, 0.0)
    with pytest.raises(qs.ArgumentError):       # conjugated output view
        qs.tensoradd(t(3, dtype=torch.complex128).conj(), t(3, dtype=torch.complex128), ((1,), ()), False, 1.0, 0.0)


def test_tensoradd_needs_device_operands():
    with pytest.raises(qs.UnsupportedError):
        qs.tensoradd(t(3, 4), t(4, 3), ((2, 1), ()), False, 1.0, 0.0)
