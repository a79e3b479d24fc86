"""Shared helpers of the GPU parity tests."""
import sys

import numpy as np


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def normwise(got, ref):
    """max|got - ref| / max|ref| (the uniform parity metric); 0 or inf for an all-zero reference."""
    d = float(np.max(np.abs(got - ref)))
    m = float(np.max(np.abs(ref)))
    if m == 0.0:
        return 0.0 if d == 0.0 else float("inf")
    return d / m


def rms_rel(got, ref):
    return float(np.sqrt(np.mean((got - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))


def pointwise(got, ref, floor=1e-3):
    """max |got - ref| / |ref| over points with |ref| >= floor * max|ref|."""
    m = np.abs(ref) >= floor * np.max(np.abs(ref))
    return float(np.max(np.abs(got[m] - ref[m]) / np.abs(ref[m])))
