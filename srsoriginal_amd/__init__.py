"""srsoriginal_amd: MI355X-native radiation hot path of OSCARS SR.

The Lienard-Wiechert spectrum / flux / power-density integrals run as
hand-written gfx950 HIP kernels in lib/libsrs_hip.so behind the C ABI of
include/srs_hip.h (Python view: srsoriginal_amd.abi).  The OSCARSSR-compatible
engine and its `sr` Python type live in lib/_sr*.so (srsoriginal_amd.sr).
"""
__version__ = "0.1.0"
