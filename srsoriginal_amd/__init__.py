"""srsoriginal_amd: MI355X-native radiation hot path of OSCARS SR.

The Lienard-Wiechert spectrum / flux / power-density integrals run as
hand-written gfx950 HIP kernels in lib/libsrs_hip.so behind the C ABI of
include/srs_hip.h (Python view: srsoriginal_amd.abi).  The OSCARSSR-compatible
engine and its `sr` Python type live in lib/_sr*.so (srsoriginal_amd.sr).

One HIP runtime per process: the ROCm PyTorch wheel ships its own
libamdhip64.so under the same soname as the system one that libsrs_hip.so
links.  Whichever is loaded first serves both, but if libsrs_hip.so loads
first the wheel's copy is mapped as well, and a second runtime in the process
sees no device (and could not use the first one's device pointers anyway).
So torch, where it is installed, is imported before any native library of
this package is loaded.
"""
__version__ = "0.1.0"

try:  # load torch's HIP runtime first, so libsrs_hip.so binds to it
    import torch as _torch  # noqa: F401
except ImportError:  # no torch: libsrs_hip.so brings the system runtime
    pass
