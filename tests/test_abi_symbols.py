"""libsrs_hip.so loads without a GPU and exports every entry point that
include/srs_hip.h declares; the no-device path fails loudly."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "srs_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(srs_hip_\w+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("srs_hip_device_count", "srs_hip_spectrum", "srs_hip_flux", "srs_hip_power_density",
                 "srs_hip_flux_device", "srs_hip_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from srsoriginal_amd import abi
    lib = ctypes.CDLL(abi.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-GPU path")
def test_no_device_is_an_error_not_a_fallback():
    from srsoriginal_amd import abi
    assert abi.device_count() == 0
    t = np.zeros((9, 4))
    with pytest.raises(abi.SrsError, match="none were found"):
        abi.flux(t, 1e-12, [[0, 0, 30]], [100.0], -1.6e-19, 0.5)
