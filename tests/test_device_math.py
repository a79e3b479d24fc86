"""CPU unit test of the kernels' scalar FP64 math (srs_math.h, compiled here for the
host from the same source): large-argument sin/cos against glibc (the reference's
libm), correctly rounded quotients, 2*pi reduction.  Runs without a GPU."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
P = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def dm(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("dm") / "dm.so")
    # FMA must be real (hardware or libm-exact), as on the GPU
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-mfma", "-shared", "-fPIC", "-o", so,
                    os.path.join(HERE, "device_math_host.cc")], check=True)
    return C.CDLL(so)


def _call(f, *arrays, nout=1):
    n = len(arrays[0])
    outs = [np.zeros(n) for _ in range(nout)]
    f(*[np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(P) for a in arrays], C.c_long(n),
      *[o.ctypes.data_as(P) for o in outs])
    return outs


def test_sincos_matches_libm_over_all_ranges(dm):
    rng = np.random.default_rng(1)
    for lo, hi in ((-1.0, 1.0), (-200.0, 200.0), (-1e6, 1e6), (1e9, 5e11), (-5e11, -1e9), (1e13, 1e15)):
        x = rng.uniform(lo, hi, 200000)
        s, c = _call(dm.dm_sincos, x, nout=2)
        # within 2 ulp of the result magnitude scale (errors of the reduced argument
        # are ~1 ulp of r <= pi/4)
        assert np.max(np.abs(s - np.sin(x))) <= 4.5e-16, (lo, hi)
        assert np.max(np.abs(c - np.cos(x))) <= 4.5e-16, (lo, hi)
    # quadrant boundaries and signs, negative arguments included
    x = np.array([0.0, -0.0, np.pi / 2, -np.pi / 2, np.pi, -np.pi, 3 * np.pi / 2, -3 * np.pi / 2, -2.0, -1e-300])
    s, c = _call(dm.dm_sincos, x, nout=2)
    assert np.allclose(s, np.sin(x), atol=1e-15) and np.allclose(c, np.cos(x), atol=1e-15)


def test_div_rn_is_correctly_rounded(dm):
    rng = np.random.default_rng(2)
    b = np.concatenate([rng.uniform(0.5, 60, 300000), np.full(300000, 2.99792458e8)])
    a = np.concatenate([rng.uniform(-0.1, 0.1, 300000), rng.uniform(0.1, 100, 300000)])
    (q,) = _call(dm.dm_div, a, b)
    assert np.array_equal(q, a / b)


def test_reduce_two_pi(dm):
    x = np.random.default_rng(3).uniform(1e8, 5e11, 100000)
    (r,) = _call(dm.dm_reduce, x)
    assert np.all(np.abs(r) <= np.pi + 1e-9)
    assert np.max(np.abs(np.sin(r) - np.sin(x))) <= 1e-15
