"""tensoradd! / tensortrace! on the device against numpy's transpose, einsum
and trace over the views' values (the closed forms TensorOperations'
StridedNative computes).  Tolerances: relative Frobenius <= 1e-13 (F64 /
C64), <= 1e-5 (F32 / C32)."""
import numpy as np
import pytest
import torch

from helpers import SV, dense, qs, rel_fro, values

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def case(msg):
    print(f"[case] {msg}", flush=True)


def test_tensoradd_transpose_f64():
    case("tensoradd F64, 3-D transposing permutation, alpha 1.5, beta 0.5")
    rng = np.random.default_rng(1)
    a = rng.standard_normal((64, 80, 48))
    c0 = rng.standard_normal((48, 64, 80))
    Cv = dense(c0, DEV)
    qs.tensoradd(Cv, dense(a, DEV), ((3, 1, 2), ()), False, 1.5, 0.5)
    assert rel_fro(values(Cv), 0.5 * c0 + 1.5 * np.transpose(a, (2, 0, 1))) <= 1e-13


def test_tensoradd_large_transpose_and_identity():
    case("tensoradd F64 2048 x 2048 transpose (beta 0) and identity permutation (beta 1)")
    rng = np.random.default_rng(2)
    a = rng.standard_normal((2048, 2048))
    Cv = dense(np.full((2048, 2048), np.nan), DEV)
    qs.tensoradd(Cv, dense(a, DEV), ((2, 1), ()), False, 1.0, 0.0)
    assert np.array_equal(values(Cv), a.T)
    c0 = rng.standard_normal((2048, 2048))
    Cv = dense(c0, DEV)
    qs.tensoradd(Cv, dense(a, DEV), ((1,), (2,)), False, -2.0, 1.0)
    assert rel_fro(values(Cv), c0 - 2.0 * a) <= 1e-13


def test_tensoradd_strided_views():
    case("tensoradd F64: A a step-2 slice with a negative stride, C padded with an offset")
    rng = np.random.default_rng(3)
    parent = rng.standard_normal(2 * 30 * 40 + 7)
    # A[i, j] = parent[6 + 2 i + 60 (39 - j)]: the j axis runs backwards.
    Av = SV(torch.from_numpy(parent).to(DEV), (30, 40), (2, -60), 6 + 60 * 39)
    a = np.array([[parent[6 + 2 * i + 60 * (39 - j)] for j in range(40)] for i in range(30)])
    cparent = rng.standard_normal(5 + 50 * 31)
    Cv = SV(torch.from_numpy(cparent.copy()).to(DEV), (40, 30), (1, 50), 5)
    qs.tensoradd(Cv, Av, ((2, 1), ()), False, 1.0, 0.25)
    got = torch.as_tensor(Cv.parent).cpu().numpy()
    want = cparent.copy()
    for i in range(40):
        for j in range(30):
            want[5 + i + 50 * j] = 0.25 * cparent[5 + i + 50 * j] + a[j, i]
    assert rel_fro(got, want) <= 1e-13
    assert np.array_equal(got[:5], cparent[:5])   # outside the view: untouched


def test_tensoradd_complex_conj():
    case("tensoradd C64: conjA xor a lazily conjugated view, complex alpha and beta")
    rng = np.random.default_rng(4)
    a = rng.standard_normal((24, 36)) + 1j * rng.standard_normal((24, 36))
    c0 = rng.standard_normal((36, 24)) + 1j * rng.standard_normal((36, 24))
    for op, conjA in (("identity", True), ("conj", False), ("conj", True)):
        Cv = dense(c0, DEV)
        qs.tensoradd(Cv, dense(a, DEV, op=op), ((2, 1), ()), conjA, 0.5 - 1j, 2 + 0.5j)
        eff = np.conj(a) if (op == "conj") != conjA else a
        assert rel_fro(values(Cv), (2 + 0.5j) * c0 + (0.5 - 1j) * eff.T) <= 1e-13, (op, conjA)


def test_tensoradd_mixed_f32_into_f64():
    case("tensoradd: Float32 A into a Float64 C (compute Float64)")
    rng = np.random.default_rng(5)
    a = rng.standard_normal((33, 65, 17)).astype(np.float32)
    c0 = rng.standard_normal((17, 33, 65))
    Cv = dense(c0, DEV)
    qs.tensoradd(Cv, dense(a, DEV), ((3,), (1, 2)), False, 1.0, 1.0)
    assert rel_fro(values(Cv), c0 + np.transpose(a.astype(np.float64), (2, 0, 1))) <= 1e-13


def test_tensortrace_one_and_two_pairs():
    case("tensortrace F64: one pair into a vector, two pairs into a permuted matrix")
    rng = np.random.default_rng(6)
    a = rng.standard_normal((32, 40, 32))
    c0 = rng.standard_normal(40)
    Cv = dense(c0, DEV)
    qs.tensortrace(Cv, dense(a, DEV), ((2,), ()), ((1,), (3,)), False, 2.0, 0.5)
    assert rel_fro(values(Cv), 0.5 * c0 + 2.0 * np.einsum("iji->j", a)) <= 1e-13
    b = rng.standard_normal((8, 30, 12, 8, 40, 12))
    c1 = rng.standard_normal((40, 30))
    Cv = dense(c1, DEV)
    qs.tensortrace(Cv, dense(b, DEV), ((5, 2), ()), ((1, 3), (4, 6)), False, 1.0, -1.0)
    assert rel_fro(values(Cv), -c1 + np.einsum("ajbaib->ij", b)) <= 1e-13


def test_tensortrace_full_to_scalar_c64():
    case("tensortrace C64: full trace of a 50 x 50 matrix into a scalar, conjA")
    rng = np.random.default_rng(7)
    a = rng.standard_normal((50, 50)) + 1j * rng.standard_normal((50, 50))
    Cv = SV(torch.zeros(1, dtype=torch.complex128, device=DEV), (), (), 0)
    qs.tensortrace(Cv, dense(a, DEV), ((), ()), ((1,), (2,)), True, 1.0, 0.0)
    got = Cv.parent.cpu().numpy()[0]
    want = np.conj(np.trace(a))
    assert abs(got - want) <= 1e-13 * abs(want)


def test_tensoradd_short_circuits():
    case("tensoradd: beta == 0 never reads C (NaN), alpha == 0 never reads A (NaN)")
    rng = np.random.default_rng(8)
    a = rng.standard_normal((70, 90))
    Cv = dense(np.full((90, 70), np.nan), DEV)
    qs.tensoradd(Cv, dense(a, DEV), ((2, 1), ()), False, 1.0, 0.0)
    assert np.array_equal(values(Cv), a.T)
    c0 = rng.standard_normal((90, 70))
    Cv = dense(c0, DEV)
    qs.tensoradd(Cv, dense(np.full((70, 90), np.nan), DEV), ((2, 1), ()), False, 0.0, 0.5)
    assert np.array_equal(values(Cv), 0.5 * c0)


def test_tensoradd_bounds_checked_before_write():
    case("tensoradd: a C view past its storage raises BoundsError, C untouched")
    rng = np.random.default_rng(9)
    a = rng.standard_normal((16, 16))
    cparent = rng.standard_normal(16 * 16 - 1)
    Cv = SV(torch.from_numpy(cparent.copy()).to(DEV), (16, 16), (1, 16), 0)
    with pytest.raises(qs.BoundsError):
        qs.tensoradd(Cv, dense(a, DEV), ((2, 1), ()), False, 1.0, 0.0)
    assert np.array_equal(Cv.parent.cpu().numpy(), cparent)


def test_backend_tensoradd_and_tensortrace_torch():
    case("QuasiStridedBackend.tensoradd / tensortrace on torch tensors (one-based TO index tuples)")
    be = qs.QuasiStridedBackend()
    g = torch.Generator(device=DEV)
    g.manual_seed(10)
    A = torch.randn(20, 30, 40, dtype=torch.float64, device=DEV, generator=g)
    C = torch.zeros(40, 20, 30, dtype=torch.float64, device=DEV)
    be.tensoradd(C, A, ((3, 1, 2), ()), False, 1.0, 0.0)
    assert torch.equal(C, A.permute(2, 0, 1))
    T = torch.randn(12, 7, 12, dtype=torch.float64, device=DEV, generator=g)
    v = torch.zeros(7, dtype=torch.float64, device=DEV)
    be.tensortrace(v, T, ((2,), ()), ((1,), (3,)), False, 1.0, 0.0)
    want = torch.einsum("iji->j", T.cpu())
    assert rel_fro(v.cpu().numpy(), want.numpy()) <= 1e-13


def test_plan_cache_reuses_layouts_and_rebinds():
    case("tensoradd plan cache: one plan per layout, storages rebound; a short storage still raises")
    add = qs.add
    add.clear_plan_cache()
    rng = np.random.default_rng(12)
    for i in range(3):   # same layout, new storages every time
        a = rng.standard_normal((40, 24, 16))
        c0 = rng.standard_normal((16, 40, 24))
        Cv = dense(c0, DEV)
        qs.tensoradd(Cv, dense(a, DEV), ((3, 1, 2), ()), False, 1.0, -0.5)
        assert rel_fro(values(Cv), -0.5 * c0 + np.transpose(a, (2, 0, 1))) <= 1e-13
        assert len(add._cache) == 1
    t = rng.standard_normal((6, 9, 6))
    v = dense(np.zeros(9), DEV)
    qs.tensortrace(v, dense(t, DEV), ((2,), ()), ((1,), (3,)), False, 1.0, 0.0)
    assert rel_fro(values(v), np.einsum("iji->j", t)) <= 1e-13
    assert len(add._cache) == 2
    # a cached plan rebound to a storage one element short raises before writing
    a = rng.standard_normal((40, 24, 16))
    cparent = rng.standard_normal(16 * 40 * 24 - 1)
    Cv = SV(torch.from_numpy(cparent.copy()).to(DEV), (16, 40, 24), (1, 16, 16 * 40), 0)
    with pytest.raises(qs.BoundsError):
        qs.tensoradd(Cv, dense(a, DEV), ((3, 1, 2), ()), False, 1.0, 0.0)
    assert np.array_equal(Cv.parent.cpu().numpy(), cparent)
    add.clear_plan_cache()
    assert len(add._cache) == 0
