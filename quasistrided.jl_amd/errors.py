"""The reference's exception classes, and the map from the C ABI's status codes."""
from __future__ import annotations


class ArgumentError(ValueError):
    """Julia's ArgumentError."""


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch."""


class BoundsError(IndexError):
    """Julia's BoundsError."""


class DeviceError(RuntimeError):
    """A HIP runtime failure reported by the library."""


class UnsupportedError(NotImplementedError):
    """Valid input that this build does not implement."""


def raise_status(status: int, message: str):
    """Raise the exception class the reference throws for `status`."""
    cls = {
        1: ArgumentError,
        2: DimensionMismatch,
        3: BoundsError,
        4: OverflowError,
        5: DeviceError,
        6: MemoryError,
        7: UnsupportedError,
    }.get(status, RuntimeError)
    raise cls(message)
