"""`ncon` on the device against numpy.einsum: the reference's integration
network (test/integrations/test_tensoroperations.jl:132-140), partial traces,
conjugation, explicit order and output, outer products and a full
contraction to a scalar.  Tolerance 1e-13 relative Frobenius (F64, C64)."""
import numpy as np
import pytest
import torch

from helpers import qs, rel_fro

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def case(msg):
    print(f"[case] {msg}", flush=True)


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to(DEV)


def test_reference_integration_network():
    case("ncon [[-1,1,-3,2],[3,-4,1],[3,2,-2]], ComplexF64, all extents 5")
    rng = np.random.default_rng(112233)
    A, B, C = crandn(rng, (5, 5, 5, 5)), crandn(rng, (5, 5, 5)), crandn(rng, (5, 5, 5))
    D = qs.ncon([dev(A), dev(B), dev(C)], [[-1, 1, -3, 2], [3, -4, 1], [3, 2, -2]])
    want = np.einsum("aecf,gde,gfb->abcd", A, B, C)
    assert tuple(D.shape) == (5, 5, 5, 5)
    assert rel_fro(D.cpu().numpy(), want) <= 1e-13
    qs.plan.clear_plan_cache()


def test_mps_environment_update_f64():
    case("ncon MPS transfer: L[a,w,a'] psi[a',s,b] W[w,v,s,t] conj(psi)[a,t,c], F64, bonds 24/40")
    rng = np.random.default_rng(5)
    D, d, w = 24, 3, 5
    L = rng.standard_normal((D, w, D))
    psi = rng.standard_normal((D, d, 40))
    W = rng.standard_normal((w, w, d, d))
    # new L'[b, v, c] = L[a, w, a'] psi[a', s, b] W[w, v, s, t] psi[a, t, c]
    out = qs.ncon([dev(L), dev(psi), dev(W), dev(psi)], [[1, 2, 3], [3, 4, -1], [2, -2, 4, 5], [1, 5, -3]],
                  order=[3, 2, 4, 1, 5])
    want = np.einsum("awx,xsb,wvst,atc->bvc", L, psi, W, psi)
    assert rel_fro(out.cpu().numpy(), want) <= 1e-13


def test_trace_conj_output_and_outer_product():
    case("ncon: partial trace, conj, explicit output order, an outer product, C64")
    rng = np.random.default_rng(7)
    T = crandn(rng, (6, 4, 6, 3))
    v = crandn(rng, (5,))
    # R[x, y, z] over T[i, x, i, y] (traced) times conj(v)[z], output (-3, -1, -2)
    R = qs.ncon([dev(T), dev(v)], [[1, -1, 1, -2], [-3]], conjlist=[False, True], output=[-3, -1, -2])
    want = np.einsum("ixiy,z->zxy", T, np.conj(v))
    assert tuple(R.shape) == (5, 4, 3)
    assert rel_fro(R.cpu().numpy(), want) <= 1e-13


def test_full_contraction_to_scalar():
    case("ncon [[1,2],[2,1]] -> a scalar, F64")
    rng = np.random.default_rng(9)
    A, B = rng.standard_normal((30, 20)), rng.standard_normal((20, 30))
    s = qs.ncon([dev(A), dev(B)], [[1, 2], [2, 1]])
    assert s.dim() == 0
    assert abs(s.item() - np.trace(A @ B)) <= 1e-12 * abs(np.trace(A @ B))


def test_label_length_mismatch():
    case("ncon: a label with two lengths raises DimensionMismatch")
    with pytest.raises(qs.DimensionMismatch):
        qs.ncon([dev(np.zeros((3, 4))), dev(np.zeros((5, 2)))], [[1, -1], [1, -2]])


def test_optimal_order_mps_transfer():
    case("ncon order='optimal' on the MPS transfer network, F64, bonds 24/40, against einsum and the given order")
    rng = np.random.default_rng(11)
    D, d, w = 24, 3, 5
    L = rng.standard_normal((D, w, D))
    psi = rng.standard_normal((D, d, 40))
    W = rng.standard_normal((w, w, d, d))
    net = [[1, 2, 3], [3, 4, -1], [2, -2, 4, 5], [1, 5, -3]]
    sizes = {1: D, 2: w, 3: D, 4: d, -1: 40, -2: w, 5: d, -3: 40}
    order = qs.optimal_order(net, sizes)
    assert qs.contraction_cost(net, sizes, order) <= qs.contraction_cost(net, sizes, [3, 2, 4, 1, 5])
    out = qs.ncon([dev(L), dev(psi), dev(W), dev(psi)], net, order="optimal")
    want = np.einsum("awx,xsb,wvst,atc->bvc", L, psi, W, psi)
    assert rel_fro(out.cpu().numpy(), want) <= 1e-13
    # a chain where the search departs from increasing labels: (A B) C
    A, B, C = rng.standard_normal((10, 100)), rng.standard_normal((100, 5)), rng.standard_normal((5, 50))
    chain = [[-1, 2], [2, 1], [1, -2]]
    assert qs.optimal_order(chain, {-1: 10, 2: 100, 1: 5, -2: 50}) == [2, 1]
    got = qs.ncon([dev(A), dev(B), dev(C)], chain, order="optimal")
    assert rel_fro(got.cpu().numpy(), A @ B @ C) <= 1e-13
    got = qs.ncon([dev(L), dev(psi), dev(W), dev(psi)], net, order="greedy")
    assert rel_fro(got.cpu().numpy(), want) <= 1e-13
    with pytest.raises(qs.ArgumentError, match="optimal"):
        qs.ncon([dev(L), dev(psi), dev(W), dev(psi)], net, order="fastest")
    qs.plan.clear_plan_cache()


def test_greedy_order_mps_overlap_past_the_exhaustive_search():
    case("ncon order='greedy': <psi|psi> of a 12-site MPS (24 tensors), ComplexF64, D=16, d=2, against einsum")
    rng = np.random.default_rng(12)
    Lsites, D, d = 12, 16, 2
    ket = [crandn(rng, (d, D)) / np.sqrt(D)] + [crandn(rng, (D, d, D)) / np.sqrt(D) for _ in range(Lsites - 2)] + \
          [crandn(rng, (D, d)) / np.sqrt(D)]
    # labels: ket bonds k_i, bra bonds b_i, physical p_i (all contracted)
    k = lambda i: 3 * i + 1
    b = lambda i: 3 * i + 2
    p = lambda i: 3 * i + 3
    net, tensors, subs = [], [], []
    letters = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ"
    sym = {}

    def s_of(l):   # one einsum letter per label
        if l not in sym:
            sym[l] = letters[len(sym)]
        return sym[l]
    for i, T in enumerate(ket):
        if i == 0:
            kl, bl = [p(i), k(i)], [p(i), b(i)]
        elif i == Lsites - 1:
            kl, bl = [k(i - 1), p(i)], [b(i - 1), p(i)]
        else:
            kl, bl = [k(i - 1), p(i), k(i)], [b(i - 1), p(i), b(i)]
        net += [kl, bl]
        tensors += [T, np.conj(T)]
        subs += ["".join(s_of(l) for l in kl), "".join(s_of(l) for l in bl)]
    want = np.einsum(",".join(subs) + "->", *tensors, optimize="greedy")
    got = qs.ncon([dev(t) for t in tensors], net, order="greedy")
    assert abs(complex(got.cpu().numpy().reshape(())) - want) <= 1e-12 * abs(want)
    qs.plan.clear_plan_cache()
