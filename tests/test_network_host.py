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


# ---- contraction-order search (`optimal_order`, `contraction_cost`)

def test_matrix_chain_known_answer():
    # A[-1,2] 10x100, B[2,1] 100x5, C[1,-2] 5x50: increasing labels contract
    # B C first (100*5*50 + 10*100*50 = 75000 multiplications); (A B) C costs
    # 10*100*5 + 10*5*50 = 7500, the textbook matrix-chain optimum.
    net = [[-1, 2], [2, 1], [1, -2]]
    sizes = {-1: 10, 2: 100, 1: 5, -2: 50}
    assert qs.contraction_cost(net, sizes) == 75000
    assert qs.optimal_order(net, sizes) == [2, 1]
    assert qs.contraction_cost(net, sizes, [2, 1]) == 7500


def _brute_min(net, sizes):
    import itertools
    labels = sorted({l for ix in net for l in ix if l > 0})
    return min(qs.contraction_cost(net, sizes, list(p)) for p in itertools.permutations(labels))


def _random_network(rng, n_tensors, n_bonds, n_open):
    import numpy as np
    net = [[] for _ in range(n_tensors)]
    for b in range(1, n_bonds + 1):   # a spanning chain first, so the network is connected
        if b < n_tensors:
            i, j = b - 1, b
        else:
            i, j = rng.choice(n_tensors, size=2, replace=False)
        net[int(i)].append(b)
        net[int(j)].append(b)
    for o in range(1, n_open + 1):
        net[int(rng.integers(n_tensors))].append(-o)
    for ix in net:
        rng.shuffle(ix)
    sizes = {l: int(rng.choice([2, 3, 5, 8, 16, 40])) for ix in net for l in ix}
    return [list(map(int, ix)) for ix in net], sizes


@pytest.mark.parametrize("seed", range(12))
def test_optimal_order_matches_brute_force(seed):
    # every order of the contracted labels (each a different pairwise tree or
    # the same tree in another sequence) against the dynamic program
    import numpy as np
    rng = np.random.default_rng(seed)
    n = 3 + seed % 3
    net, sizes = _random_network(rng, n, n + 1 + seed % 2, 3)
    order = qs.optimal_order(net, sizes)
    assert sorted(order) == sorted({l for ix in net for l in ix if l > 0})
    assert qs.contraction_cost(net, sizes, order) == _brute_min(net, sizes)


def test_optimal_order_traces_and_components():
    # a partial trace on tensor 0 and two components joined by an outer product
    net = [[1, 1, -1, 2], [2, 3, -2], [3, -3], [4, -4], [4, 5], [5, -5]]
    sizes = {1: 3, -1: 4, 2: 30, 3: 2, -2: 50, -3: 60, 4: 7, -4: 9, 5: 100, -5: 2}
    order = qs.optimal_order(net, sizes)
    assert sorted(order) == [1, 2, 3, 4, 5]
    assert qs.contraction_cost(net, sizes, order) == _brute_min(net, sizes)


def test_optimal_order_rejects_bad_networks():
    with pytest.raises(qs.ArgumentError, match="contracted label 1 appears 1"):
        qs.optimal_order([[1, -1], [2, -2]], {1: 2, 2: 2, -1: 2, -2: 2})


# ---- greedy order search (`greedy_order`), for networks past 16 tensors

def test_greedy_order_matrix_chain_known_answer():
    # (A B) leaves a 10 x 5 result from operands of 1000 and 500 entries,
    # (B C) a 100 x 50 one: greedy takes A B first, the optimum (7500).
    net = [[-1, 2], [2, 1], [1, -2]]
    sizes = {-1: 10, 2: 100, 1: 5, -2: 50}
    assert qs.greedy_order(net, sizes) == [2, 1]
    assert qs.contraction_cost(net, sizes, qs.greedy_order(net, sizes)) == 7500


@pytest.mark.parametrize("seed", range(8))
def test_greedy_order_is_a_valid_order_no_cheaper_than_optimal(seed):
    import numpy as np
    rng = np.random.default_rng(100 + seed)
    n = 3 + seed % 4
    net, sizes = _random_network(rng, n, n + 1 + seed % 3, 3)
    order = qs.greedy_order(net, sizes)
    assert sorted(order) == sorted({l for ix in net for l in ix if l > 0})
    assert qs.contraction_cost(net, sizes, order) >= qs.contraction_cost(net, sizes, qs.optimal_order(net, sizes))


def test_greedy_order_traces_components_and_large_networks():
    net = [[1, 1, -1, 2], [2, 3, -2], [3, -3], [4, -4], [4, 5], [5, -5]]
    sizes = {1: 3, -1: 4, 2: 30, 3: 2, -2: 50, -3: 60, 4: 7, -4: 9, 5: 100, -5: 2}
    assert sorted(qs.greedy_order(net, sizes)) == [1, 2, 3, 4, 5]
    # an MPS overlap <psi|psi> of 12 sites: 24 tensors, past the exhaustive
    # search; greedy runs the zipper (site by site), bond^3 * d work per step
    L, D, d = 12, 32, 2
    net, sizes = [], {}
    for i in range(L):
        left, right = 3 * i, 3 * (i + 1)         # ket bonds
        bl, br, p = 3 * i + 1, 3 * (i + 1) + 1, 3 * i + 2   # bra bonds, physical
        ket = ([right] if i == 0 else [left, right]) if i < L - 1 else [left]
        bra = ([br] if i == 0 else [bl, br]) if i < L - 1 else [bl]
        net.append(ket + [p])
        net.append(bra + [p])
    net = [[l + 1 for l in ix] for ix in net]
    for ix in net:
        for l in ix:
            sizes[l] = d if (l - 1) % 3 == 2 else D
    with pytest.raises(qs.ArgumentError, match="beyond the exhaustive search"):
        qs.optimal_order(net, sizes)
    order = qs.greedy_order(net, sizes)
    assert sorted(order) == sorted({l for ix in net for l in ix})
    # the zipper's cost: per site about two steps of D^3 d, far below the
    # D^(L) of a bad order
    assert qs.contraction_cost(net, sizes, order) <= 4 * L * D ** 3 * d
