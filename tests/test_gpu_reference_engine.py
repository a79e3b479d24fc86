"""The drop-in at the C++ boundary: the reference's own engine (its src/*.cc compiled with
its GPU switch, -DCUDA, by oracle/Makefile) with srsoriginal_amd/adapter/OSCARSSR_Cuda_hip.cc
in place of OSCARSSR_Cuda.cu, run through the reference's gpu=1 routes
(OSCARSSR.cc:894-900, :1876-1882, :2582-2588).  Its results must match the same engine's
CPU routes (the golden fixtures) at 1e-10: the reference computes the trajectories and
drives the per-particle loops and containers, libsrs_hip.so computes the integrals.

The harness binary is built where the reference sources exist; where it was not built,
these tests are skipped."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import cases
from conftest import GOLDEN, ROOT, golden
from gpu_util import normwise, say

pytestmark = pytest.mark.gpu

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness_hip")
needs_harness = pytest.mark.skipif(not os.path.exists(HARNESS), reason="reference engine with the adapter not built")


def _run(mode, name, extra=()):
    args = ["%s=%s" % kv for kv in cases.args(name).items()] + ["gpu=1"] + list(extra)
    with tempfile.NamedTemporaryFile(suffix=".bin") as f:
        say("reference engine, %s route gpu=1: %s" % (mode, name))
        r = subprocess.run([HARNESS, mode] + args + ["out=" + f.name], cwd=GOLDEN, capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr[-2000:]
        return np.fromfile(f.name, dtype="<f8")


def _container(a):
    n = int(a[0])
    return a[1:1 + 4 * n].reshape(n, 4)


@needs_harness
@pytest.mark.parametrize("name", ["spectrum_c0_small", "spectrum_points_offaxis", "multispectrum_c3_small",
                                  "spectrum_beam_x"])
def test_reference_engine_spectrum_gpu_route(name):
    got = _run("spectrum", name)[1:].reshape(-1, 2)
    ref = golden(name)["spectrum"]
    assert np.array_equal(got[:, 0], ref[:, 0])
    err = normwise(got[:, 1], ref[:, 1])
    say("  normwise vs the CPU route %.3e" % err)
    assert err <= 1e-10


@needs_harness
@pytest.mark.parametrize("name", ["flux_c1_small", "flux_points_list", "flux_fields_mixed", "multiflux_c3_small",
                                  "multiflux_two_beams", "flux_beam_x", "flux_taper_center", "flux_x0x1x2",
                                  "flux_efield_only", "flux_efield_bfield"])
def test_reference_engine_flux_gpu_route(name):
    got = _container(_run("flux", name))
    ref = golden(name)["container"]
    # the CPU fixture's first N entries hold the values; the gpu=1 route adds N more 3-D
    # points per particle after them (OSCARSSR.cc:2570-2572)
    a = cases.args(name)
    if "points" in a:
        n = len(cases.points(a)[0])
    elif "x0x1x2" in a or "nx" in a:
        nx = [int(float(v)) for v in a["nx"].split(",")]
        n = nx[0] * nx[1]
    assert np.array_equal(got[:n, :3], ref[:n, :3])
    err = normwise(got[:n, 3], ref[:n, 3])
    say("  normwise vs the CPU route %.3e" % err)
    assert err <= 1e-10


@needs_harness
@pytest.mark.parametrize("name", ["pd_c2_small", "pd_directional_tilted", "pd_directional_inside", "pd_beam_x",
                                  "pd_points_rotated_flip", "pd_x0x1x2_normal", "pd_efield_bfield"])
def test_reference_engine_power_density_gpu_route(name):
    got = _container(_run("pd", name))
    ref = golden(name)["container"]
    a = cases.args(name)
    if "points" in a:
        n = len(cases.points(a)[0])
    else:
        nx = [int(float(v)) for v in a["nx"].split(",")]
        n = nx[0] * nx[1]
    err = normwise(got[:n, 3], ref[:n, 3])
    say("  normwise vs the CPU route %.3e" % err)
    assert err <= 1e-10


@needs_harness
def test_reference_engine_reports_the_library_device_count():
    # OSCARSSR::CheckGPU (OSCARSSR.cc:499-505) answers through the adapter: a flux on a 1x1
    # grid succeeds only if it saw at least one device
    got = _container(_run("flux", "flux_c1_small", ["nx=2,2"]))
    assert got.shape[0] >= 4 and np.all(np.isfinite(got[:, 3]))
