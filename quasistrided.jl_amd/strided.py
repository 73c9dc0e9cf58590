"""StridedView: the operand type of the engine, as StridedViews.jl's
`StridedView(parent, size, strides, offset, op)`.

`parent` is a one-dimensional torch tensor holding the storage (device memory
for `QS_MEM_DEVICE`, host memory for `QS_MEM_HOST`); `size` and `strides` are
in elements and may be zero or negative; `offset` is the zero-based element
index of the view's first element; `op` is one of identity / conj / transpose
/ adjoint (src/planning/plan.jl:176-186 folds it into the conjugation flags).
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from .errors import ArgumentError

OPS = ("identity", "conj", "transpose", "adjoint")

DTYPE_CODE = {torch.float32: 0, torch.float64: 1, torch.complex64: 2, torch.complex128: 3}
CODE_DTYPE = {v: k for k, v in DTYPE_CODE.items()}


@dataclass
class StridedView:
    parent: torch.Tensor
    size: tuple
    strides: tuple
    offset: int = 0
    op: str = "identity"

    def __post_init__(self):
        if self.parent.dim() != 1:
            raise ArgumentError("StridedView: parent must be a one-dimensional storage tensor")
        self.size = tuple(int(s) for s in self.size)
        self.strides = tuple(int(s) for s in self.strides)
        if len(self.size) != len(self.strides):
            raise ArgumentError("StridedView: size and strides differ in length")
        if any(s < 0 for s in self.size):
            raise ArgumentError("StridedView: negative size")
        if self.op not in OPS:
            raise ArgumentError(f"unsupported StridedView.op {self.op}: QuasiStrided recognizes only "
                                "identity/conj/transpose/adjoint")
        self.offset = int(self.offset)

    # ------------------------------------------------------------ builders
    @classmethod
    def of(cls, t) -> "StridedView":
        """The view a torch tensor (any strides, possibly a lazy conj) describes."""
        if isinstance(t, StridedView):
            return t
        if not isinstance(t, torch.Tensor):
            raise ArgumentError(f"QuasiStridedBackend requires strided arrays, got {type(t).__name__}")
        op = "conj" if t.is_conj() else "identity"
        storage = t.untyped_storage()
        n = storage.nbytes() // t.element_size()
        parent = torch.empty(0, dtype=t.dtype, device=t.device)
        parent.set_(storage, 0, (n,), (1,))
        return cls(parent, tuple(t.shape), tuple(t.stride()), t.storage_offset(), op)

    # ------------------------------------------------------------ queries
    @property
    def dtype(self):
        return self.parent.dtype

    @property
    def ndim(self) -> int:
        return len(self.size)

    @property
    def is_complex(self) -> bool:
        return self.parent.dtype.is_complex

    def numel(self) -> int:
        n = 1
        for s in self.size:
            n *= s
        return n

    def data_range(self):
        """[start, end) byte range of the parent storage (for the alias check)."""
        p = self.parent
        return p.data_ptr(), p.data_ptr() + p.numel() * p.element_size()

    def permute(self, perm) -> "StridedView":
        return StridedView(self.parent, tuple(self.size[p] for p in perm), tuple(self.strides[p] for p in perm),
                           self.offset, self.op)

    def conj(self) -> "StridedView":
        flip = {"identity": "conj", "conj": "identity", "transpose": "adjoint", "adjoint": "transpose"}
        return StridedView(self.parent, self.size, self.strides, self.offset, flip[self.op])

    def to_tensor(self) -> torch.Tensor:
        """The view's values as a dense tensor (op applied), for checks."""
        if self.numel() == 0:
            return torch.zeros(self.size, dtype=self.dtype, device=self.parent.device)
        idx = torch.full(self.size, self.offset, dtype=torch.int64)
        for d, (n, s) in enumerate(zip(self.size, self.strides)):
            shape = [1] * len(self.size)
            shape[d] = n
            idx = idx + (torch.arange(n, dtype=torch.int64) * s).reshape(shape)
        vals = self.parent[idx.reshape(-1).to(self.parent.device)].reshape(self.size)
        if self.op in ("conj", "adjoint") and self.is_complex:
            vals = vals.conj().resolve_conj()
        return vals.clone()
