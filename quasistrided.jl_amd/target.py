"""The gfx950 TargetProfile (the device counterpart of src/hardware/target.jl)."""
from __future__ import annotations

import ctypes as C

from . import _lib
from .errors import raise_status

# The MI355X datasheet figures the roofline is priced against (dense rates).
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_TFLOPS = 2500.0
PEAK_INT8_TOPS = 5000.0
PEAK_HBM_GBS = 8000.0


def target_profile() -> dict:
    """hipGetDeviceProperties of the current device plus the MFMA table."""
    p = _lib.QSTargetProfile()
    st = _lib.lib().qs_target_profile_get(C.byref(p))
    if st != 0:
        raise_status(st, _lib.last_error())
    out = {f: getattr(p, f) for f, _ in _lib.QSTargetProfile._fields_}
    out["name"] = p.name.decode()
    return out


def measure_peaks(stream=None) -> dict:
    """Measured FP64 MFMA rate (TFLOP/s) and device copy rate (GB/s) of the
    current device, beside the datasheet figures (qs_target_measure)."""
    import torch
    s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
    f, b = C.c_double(), C.c_double()
    st = _lib.lib().qs_target_measure(C.byref(f), C.byref(b), C.c_void_p(s))
    if st != 0:
        raise_status(st, _lib.last_error())
    return {"fp64_mfma_tflops": f.value, "copy_gbs": b.value,
            "datasheet_fp64_tflops": PEAK_FP64_TFLOPS, "datasheet_hbm_gbs": PEAK_HBM_GBS}
