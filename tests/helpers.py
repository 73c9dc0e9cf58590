"""Fixtures and checks shared by the tests (restating test/helpers.jl and
test/execution/helpers.jl of the reference)."""
from __future__ import annotations

import numpy as np
import torch

from conftest import load_package

qs = load_package()
SV = qs.StridedView

TORCH_OF = {np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64,
            np.dtype(np.complex64): torch.complex64, np.dtype(np.complex128): torch.complex128}
NP_OF = {v: k for k, v in TORCH_OF.items()}
CODE_OF = {torch.float32: 0, torch.float64: 1, torch.complex64: 2, torch.complex128: 3}

# Relative Frobenius tolerances of the GPU path against the oracle.
RTOL = {torch.float64: 1e-13, torch.complex128: 1e-13, torch.float32: 1e-5, torch.complex64: 1e-5}


def randn(rng, dtype, n):
    """n iid N(0,1) entries (complex: re and im each N(0,1))."""
    dt = np.dtype(dtype)
    if dt.kind == "c":
        re = rng.standard_normal(n)
        im = rng.standard_normal(n)
        return (re + 1j * im).astype(dt)
    return rng.standard_normal(n).astype(dt)


def view(parent_np, size, strides, offset=0, op="identity", device="cpu"):
    """A StridedView over a 1-D numpy storage copied to `device`."""
    t = torch.from_numpy(np.ascontiguousarray(parent_np)).to(device)
    return SV(t, size, strides, offset, op)


def dense(arr_np, device="cpu", op="identity"):
    """A column-major dense view of a numpy array's values (Julia layout)."""
    a = np.asarray(arr_np)
    flat = a.reshape(-1, order="F").copy()
    strides, s = [], 1
    for n in a.shape:
        strides.append(s)
        s *= n
    return view(flat, a.shape, tuple(strides), 0, op, device)


def values(v: SV) -> np.ndarray:
    """The view's logical values (op applied) as a numpy array."""
    t = v.to_tensor()
    return t.cpu().numpy()


def brute_ref(Av, indA, Bv, indB, Cstart, indC, alpha, beta, conjA=False, conjB=False, dtype=np.complex128):
    """_brute_ref / _lo_reference (test/execution/helpers.jl:6-32): einsum in
    float64 / complex128 over the views' values (ops applied)."""
    a = values(Av).astype(dtype)
    b = values(Bv).astype(dtype)
    if conjA:
        a = np.conj(a)
    if conjB:
        b = np.conj(b)
    letters = {}
    for l in list(indA) + list(indB) + list(indC):
        if l not in letters:
            letters[l] = chr(ord("a") + len(letters))
    spec = "".join(letters[l] for l in indA) + "," + "".join(letters[l] for l in indB) + "->" + \
        "".join(letters[l] for l in indC)
    ab = np.einsum(spec, a, b, optimize=True)
    out = alpha * ab
    if beta != 0:
        out = out + beta * Cstart.astype(dtype)
    return out


def rel_fro(x, ref) -> float:
    x = np.asarray(x, dtype=np.complex128)
    ref = np.asarray(ref, dtype=np.complex128)
    den = np.linalg.norm(ref)
    num = np.linalg.norm(x - ref)
    return float(num / den) if den > 0 else float(num)


def oracle_of_plan(oracle, plan, alpha, beta, a_np, b_np, c_np, k_block=0):
    """Run the oracle on the plan's own groups and storages (host copies);
    returns the oracle's C storage."""
    c = c_np.copy()
    T = plan.T
    g = lambda G: (G.lengths, G.strides[0], G.strides[1])
    st = oracle.contract(a_np, plan.Abase, b_np, plan.Bbase, c, plan.Cbase, g(plan.mgroup), g(plan.ngroup),
                         g(plan.kgroup), CODE_OF[T], alpha, beta, conj_a=plan.atransform == "conj",
                         conj_b=plan.btransform == "conj", k_block=k_block)
    assert st == 0, st
    return c


def storage_np(t: torch.Tensor) -> np.ndarray:
    return t.detach().cpu().numpy().copy()
