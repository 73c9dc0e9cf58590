"""MI355X (gfx950) block-scatter tensor contraction: a drop-in engine for
QuasiStrided.jl's `plan_contract` / `execute!` / `contract!` /
`QuasiStridedBackend` path.

Host planning (labels -> M/N/K axis groups) is done here as the reference
does it; the contraction runs in libqs_amd.so, hand-written HIP kernels for
gfx950 behind the C ABI of include/qs_amd.h.
"""
from .errors import ArgumentError, BoundsError, DeviceError, DimensionMismatch, UnsupportedError
from .strided import StridedView
from .labels import (AxisGroup, affine_ramp, axis_length, classify_labels, group_ranks, leading_unit_run,
                     map_ramp_step, order_contract_labels, order_free_labels, prefer_swap)
from .plan import (ContractPlan, MFMAKernel, PathModes, autotune, autotune_clear, autotune_load, autotune_save,
                   compute_type, contract, contract_, error_report, execute, execute_, isconj, op_conjugates, plan_contract, planned,
                   release, set_error_report)
from .backend import QuasiStridedBackend, contraction_labels, tensorcontract
from .add import tensoradd, tensortrace
from .target import measure_peaks, target_profile
from .group import GroupPlan, Graph
from .batch import BatchedPlan, batch_labels, contract_batched, plan_batched
from .server import Server
from .allocators import DefaultAllocator, PoolAllocator, TorchAllocator
from .network import contraction_cost, greedy_order, ncon, ncon_steps, optimal_order
from . import _lib

__all__ = [
    "ArgumentError", "BoundsError", "DeviceError", "DimensionMismatch", "UnsupportedError", "StridedView",
    "AxisGroup", "affine_ramp", "axis_length", "classify_labels", "group_ranks", "leading_unit_run",
    "map_ramp_step", "order_contract_labels", "order_free_labels", "prefer_swap", "ContractPlan", "MFMAKernel",
    "PathModes", "autotune", "autotune_clear", "autotune_load", "autotune_save", "compute_type", "contract", "contract_", "error_report", "set_error_report", "execute", "execute_", "isconj", "op_conjugates",
    "plan_contract", "planned", "release", "QuasiStridedBackend", "contraction_labels", "tensorcontract",
    "tensoradd", "tensortrace",
    "target_profile", "measure_peaks", "GroupPlan", "Graph", "Server", "BatchedPlan", "batch_labels", "contract_batched",
    "plan_batched", "DefaultAllocator", "PoolAllocator", "TorchAllocator", "ncon", "ncon_steps", "optimal_order", "greedy_order", "contraction_cost",
]
