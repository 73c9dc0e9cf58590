"""The gfx950 TargetProfile (counterpart of src/hardware/target.jl:26-47,
181-191) read on the device: the facts the planner takes its tile and
split-K choices from."""
import ctypes as C

import pytest

from helpers import qs

pytestmark = pytest.mark.gpu


def test_target_profile_is_gfx950():
    t = qs._lib.QSTargetProfile()
    assert qs._lib.lib().qs_target_profile_get(C.byref(t)) == 0, qs._lib.last_error()
    print(f"[target] {t.name.decode()} CUs {t.compute_units} wave {t.wavefront_size} "
          f"LDS/CU {t.lds_bytes_per_cu} LDS/block {t.lds_bytes_per_block} XCDs {t.xcds} L2 {t.l2_bytes} "
          f"HBM {t.hbm_bytes / 2**30:.0f} GiB clock {t.clock_khz} kHz", flush=True)
    assert t.name.decode().startswith("gfx950")
    assert t.compute_units == 256
    assert t.wavefront_size == 64
    assert t.lds_bytes_per_cu == 160 * 1024
    assert t.xcds == 8
    assert (t.mfma_f64_m, t.mfma_f64_n, t.mfma_f64_k) == (16, 16, 4)
    assert (t.mfma_i8_m, t.mfma_i8_n, t.mfma_i8_k) == (16, 16, 64)
    assert t.hbm_bytes > 250 * 2**30
