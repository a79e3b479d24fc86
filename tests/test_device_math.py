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


def test_sqrt_rn_is_correctly_rounded_from_a_rough_estimate(dm):
    """sqrt_rn(x, y0) equals the correctly rounded square root (numpy's, IEEE) for any
    estimate y0 of 1/sqrt(x) within 2^-16 relative -- the hardware v_rsq_f64 is far
    closer -- over the squared distances the flux terms meet and near perfect squares,
    where a misrounding would show first."""
    rng = np.random.default_rng(4)
    x = np.concatenate([
        10.0 ** rng.uniform(-30, 30, 400000),              # 1e-15 .. 1e15 m distances, squared
        rng.uniform(899.0, 1001.0, 200000),               # ~30 m
        np.ldexp(rng.uniform(0.5, 1.0, 100000), rng.integers(-900, 1000, 100000)),
    ])
    k = rng.integers(1, 2 ** 26, 100000).astype(np.float64)
    sq = k * k                                            # exact squares and their neighbours
    x = np.concatenate([x, sq, np.nextafter(sq, 0), np.nextafter(sq, np.inf)])
    exact = np.sqrt(x)
    for rel in (0.0, 2.0 ** -40, -(2.0 ** -30), 2.0 ** -20, -(2.0 ** -16)):
        y0 = (1.0 / exact) * (1.0 + rel * rng.uniform(0.0, 1.0, len(x)))
        (g,) = _call(dm.dm_sqrt, x, y0)
        bad = np.flatnonzero(g != exact)
        assert bad.size == 0, (rel, x[bad[:5]], g[bad[:5]], exact[bad[:5]])
