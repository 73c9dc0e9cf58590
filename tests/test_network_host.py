"""`ncon`'s host planning (`ncon_steps`): label checks, traces, the pairwise
contraction sequence and outer products, against hand-worked networks."""
import pytest

from helpers import qs

steps_of = qs.ncon_steps


def test_reference_network():
    # test/integrations/test_tensoroperations.jl:137: [[-1,1,-3,2],[3,-4,1],[3,2,-2]]
    traces, steps, out = steps_of([[-1, 1, -3, 2], [3, -4, 1], [3, 2, -2]])
    assert traces == {} and out == (-1, -2, -3, -4)
    # label 1 joins tensors 0 and 1 (sharing only 1); then 2 joins node 3 and tensor 2, sharing 2 and 3
    assert steps == [(0, 1, (1,), (-1, -3, 2, 3, -4)), (2, 3, (3, 2), (-2, -1, -3, -4))]


def test_order_and_output_override():
    traces, steps, out = steps_of([[1, -1], [1, 2], [2, -2]], order=[2, 1], output=[-2, -1])
    assert out == (-2, -1)
    assert steps == [(1, 2, (2,), (1, -2)), (0, 3, (1,), (-1, -2))]


def test_partial_trace_and_outer_product():
    traces, steps, out = steps_of([[1, -1, 1], [-2]])
    assert traces == {0: ((-1,), (1,))}
    assert steps == [(0, 1, (), (-1, -2))]
    assert out == (-1, -2)


def test_full_contraction_to_a_scalar():
    traces, steps, out = steps_of([[1, 2], [2, 1]])
    assert out == () and steps == [(0, 1, (1, 2), ())]


@pytest.mark.parametrize("network,msg", [
    ([[1, -1], [-1, 1]], "open label -1 appears 2"),
    ([[1, 1, 1]], "contracted label 1 appears 3"),
    ([[1, -1], [2, -2]], "contracted label 1 appears 1"),
    ([[0, -1]], "label 0"),
])
def test_invalid_networks(network, msg):
    with pytest.raises(qs.ArgumentError, match=msg):
        steps_of(network)


def test_bad_order_and_output():
    with pytest.raises(qs.ArgumentError, match="order"):
        steps_of([[1, -1], [1, -2]], order=[2])
    with pytest.raises(qs.ArgumentError, match="output"):
        steps_of([[1, -1], [1, -2]], output=[-1, -3])
