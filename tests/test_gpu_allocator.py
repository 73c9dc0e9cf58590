"""The workspace allocator on the GPU (plan_contract's `allocator`,
src/planning/plan.jl:125-131; tensoralloc / release!,
src/execution/workspace.jl:83-123): a plan's split-K partials and
split-precision buffer come from the caller's allocator, results are those of
the library's own allocation, and the workspace goes back in reverse order."""
import numpy as np
import pytest
import torch

from helpers import SV, qs, randn, rel_fro, values

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def case(msg):
    print(f"[case] {msg}", flush=True)


def dense_views(rng, M, K, N, dtype=np.float64):
    A = torch.from_numpy(randn(rng, dtype, M * K)).to(DEV)
    B = torch.from_numpy(randn(rng, dtype, K * N)).to(DEV)
    Cst = torch.from_numpy(randn(rng, dtype, M * N)).to(DEV)
    return SV(A, (M, K), (1, M)), SV(B, (K, N), (1, K)), SV(Cst, (M, N), (1, M))


def reference(Av, Bv, c0, alpha, beta):
    return alpha * (values(Av) @ values(Bv)) + beta * c0


class Counting(qs.TorchAllocator):
    pass


def test_split_k_partials_from_torch_pool():
    case("forced split-K 3: partials from torch's caching allocator, F64 200 x 1000 x 150")
    rng = np.random.default_rng(70)
    Av, Bv, Cv = dense_views(rng, 200, 1000, 150)
    c0 = values(Cv).copy()
    al = Counting()
    plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), allocator=al,
                            kernel=qs.MFMAKernel(tile=1, k_splits=3), path_modes=qs.PathModes("never", "never"))
    assert plan.info["k_splits"] == 3
    assert al.allocs == 1 and al.log[0][2] == 3 * 200 * 150 * 8
    assert plan.info["workspace_bytes"] == 3 * 200 * 150 * 8
    qs.execute(plan, 1.5, 0.5)
    torch.cuda.synchronize()
    assert rel_fro(values(Cv), reference(Av, Bv, c0, 1.5, 0.5)) <= 1e-13
    ptr = al.log[0][1]
    plan.release()
    assert al.frees == 1 and al.log[-1] == ("free", ptr, 3 * 200 * 150 * 8) and al.live_bytes == 0


def test_split_precision_workspace_from_torch_pool():
    case("split precision: residues and scales from the allocator, F64 512^3, released in reverse order")
    rng = np.random.default_rng(71)
    Av, Bv, Cv = dense_views(rng, 512, 512, 512)
    C2 = SV(torch.zeros(512 * 512, dtype=torch.float64, device=DEV), (512, 512), (1, 512))
    al = Counting()
    split = qs.plan_contract(C2, Av, (1, 2), Bv, (2, 3), (1, 3), precision="split", allocator=al)
    plain = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3))
    assert al.allocs >= 1 and al.live_bytes > 0
    assert split.info["workspace_bytes"] == al.live_bytes
    qs.execute(split, 1.0, 0.0)
    qs.execute(plain, 1.0, 0.0)
    torch.cuda.synchronize()
    assert rel_fro(values(C2), values(Cv)) <= 1e-13
    acquired = [(p, n) for op, p, n in al.log if op == "alloc"]
    split.release_workspace()
    assert split.info["workspace_bytes"] == 0
    freed = [(p, n) for op, p, n in al.log if op == "free"]
    assert freed == acquired[::-1] and al.live_bytes == 0
    with pytest.raises(qs.ArgumentError):
        qs.execute(split, 1.0, 0.0)
    split.release()
    assert al.frees == len(acquired)   # destroy hands nothing back twice
    plain.release()


def test_released_plan_is_refused_everywhere():
    case("a released plan: execute, autotune and groups refuse it")
    rng = np.random.default_rng(72)
    Av, Bv, Cv = dense_views(rng, 64, 64, 64)
    al = Counting()
    plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), allocator=al)
    plan.release_workspace()
    with pytest.raises(qs.ArgumentError):
        qs.execute(plan, 1.0, 0.0)
    with pytest.raises(qs.ArgumentError):
        qs.autotune(plan)
    with pytest.raises(qs.ArgumentError):
        qs.GroupPlan([plan])
    plan.release()


def test_allocator_without_memory():
    case("an allocator with no memory: split-K runs unsplit; split precision raises")
    rng = np.random.default_rng(73)

    class Empty(qs.TorchAllocator):
        def alloc(self, nbytes):
            self.allocs += 1
            return 0

    Av, Bv, Cv = dense_views(rng, 100, 900, 80)
    c0 = values(Cv).copy()
    al = Empty()
    plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), allocator=al,
                            kernel=qs.MFMAKernel(tile=1, k_splits=4), path_modes=qs.PathModes("never", "never"))
    assert al.allocs == 1 and plan.info["k_splits"] == 1
    qs.execute(plan, 1.0, 1.0)
    torch.cuda.synchronize()
    assert rel_fro(values(Cv), reference(Av, Bv, c0, 1.0, 1.0)) <= 1e-13
    plan.release()
    with pytest.raises(MemoryError):
        qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), precision="split", allocator=Empty())


def test_misaligned_block_is_handed_back():
    case("an allocator returning a block not aligned to 256 bytes: handed back, treated as no memory")
    rng = np.random.default_rng(74)

    class Skewed(qs.TorchAllocator):
        def alloc(self, nbytes):
            return super().alloc(nbytes + 512) + 8

        def free(self, ptr, nbytes):
            super().free(ptr - 8, nbytes + 512)

    Av, Bv, Cv = dense_views(rng, 100, 900, 80)
    al = Skewed()
    plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), allocator=al,
                            kernel=qs.MFMAKernel(tile=1, k_splits=4), path_modes=qs.PathModes("never", "never"))
    assert plan.info["k_splits"] == 1 and al.allocs == 1 and al.frees == 1 and al.live_bytes == 0
    plan.release()


def test_backend_checkpoint_and_reset():
    case("TensorOperations backend: allocator_checkpoint! / allocator_reset! around the contraction")
    rng = np.random.default_rng(75)

    class Arena(qs.TorchAllocator):
        def __init__(self):
            super().__init__()
            self.marks = []

        def checkpoint(self):
            self.marks.append("checkpoint")
            return len(self.log)

        def reset(self, cp):
            self.marks.append(("reset", cp))

    A = torch.from_numpy(randn(rng, np.float64, 40 * 30)).to(DEV).reshape(30, 40).t()
    B = torch.from_numpy(randn(rng, np.float64, 30 * 20)).to(DEV).reshape(20, 30).t()
    Cd = torch.zeros(40, 20, dtype=torch.float64, device=DEV)
    al = Arena()
    qs.QuasiStridedBackend().tensorcontract(Cd, A, ((1,), (2,)), False, B, ((1,), (2,)), False, ((1, 2), ()),
                                            1.0, 0.0, al)
    torch.cuda.synchronize()
    assert al.marks == ["checkpoint", ("reset", 0)]
    assert rel_fro(Cd.cpu().numpy(), (A @ B).cpu().numpy()) <= 1e-13


def test_pool_allocator_workspaces_and_reuse():
    case("PoolAllocator: split-K partials and split-precision buffers from the library's hipMemPool, F64")
    rng = np.random.default_rng(76)
    pool = qs.PoolAllocator()
    assert pool.stats()["live_blocks"] == 0 and pool.stats()["used_bytes"] == 0
    Av, Bv, Cv = dense_views(rng, 200, 1000, 150)
    c0 = values(Cv).copy()
    C_lib = SV(Cv.parent.clone(), Cv.size, Cv.strides)
    kern = qs.MFMAKernel(tile=1, k_splits=3)
    modes = qs.PathModes("never", "never")
    plan = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), allocator=pool, kernel=kern, path_modes=modes)
    own = qs.plan_contract(C_lib, Av, (1, 2), Bv, (2, 3), (1, 3), kernel=kern, path_modes=modes)
    assert plan.info["k_splits"] == 3 and plan.info["workspace_bytes"] == 3 * 200 * 150 * 8
    st = pool.stats()
    assert st["live_blocks"] == 1 and st["used_bytes"] >= 3 * 200 * 150 * 8
    qs.execute(plan, 1.5, 0.5)
    qs.execute(own, 1.5, 0.5)
    torch.cuda.synchronize()
    assert torch.equal(Cv.parent, C_lib.parent)   # same kernel, same split: the same bits
    assert rel_fro(values(Cv), reference(Av, Bv, c0, 1.5, 0.5)) <= 1e-13
    own.release()

    case("PoolAllocator: split precision 512^3 against the plain FP64 result")
    Ap, Bp, Cp = dense_views(rng, 512, 512, 512)
    C2 = SV(torch.zeros(512 * 512, dtype=torch.float64, device=DEV), (512, 512), (1, 512))
    split = qs.plan_contract(C2, Ap, (1, 2), Bp, (2, 3), (1, 3), precision="split", allocator=pool)
    plain = qs.plan_contract(Cp, Ap, (1, 2), Bp, (2, 3), (1, 3))
    assert pool.stats()["live_blocks"] >= 2
    qs.execute(split, 1.0, 0.0)
    qs.execute(plain, 1.0, 0.0)
    torch.cuda.synchronize()
    assert rel_fro(values(C2), values(Cp)) <= 1e-13
    plain.release()

    with pytest.raises(qs.ArgumentError):   # plans still hold blocks
        pool.close()
    split.release()
    plan.release()
    torch.cuda.synchronize()
    st = pool.stats()
    assert st["live_blocks"] == 0 and st["used_bytes"] == 0 and st["reserved_bytes"] > 0

    case("PoolAllocator: a second plan reuses the reserved memory")
    reserved = st["reserved_bytes"]
    again = qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), allocator=pool, kernel=kern, path_modes=modes)
    assert again.info["k_splits"] == 3 and pool.stats()["reserved_bytes"] == reserved
    qs.execute(again, 1.5, 0.0)
    torch.cuda.synchronize()
    assert rel_fro(values(Cv), reference(Av, Bv, c0, 1.5, 0.0)) <= 1e-13
    again.release()
    pool.close()
    with pytest.raises(ValueError):
        pool.stats()
    with pytest.raises(ValueError):
        qs.plan_contract(Cv, Av, (1, 2), Bv, (2, 3), (1, 3), allocator=pool, kernel=kern, path_modes=modes)
