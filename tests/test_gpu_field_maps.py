"""GPU tests of the field maps (TField3D_Grid, kind SRS_FIELD_GRID of the C ABI): the
device RK4 through two maps (an OSCARS 3-D map and a shifted, scaled OSCARS1D map)
against the reference engine's trajectory through the same files, and the flux map
of that trajectory on both routes (host RK4 + srs_hip_flux, device RK4 + batch)
against the reference's flux."""
import numpy as np
import pytest

import cases
from conftest import golden
from gpu_util import normwise, say

pytestmark = pytest.mark.gpu


def _engine(name):
    from test_host_engine import _engine as e
    return e(name)


def test_device_trajectory_through_field_maps_matches_reference():
    say("device RK4 through OSCARS + OSCARS1D field maps vs the reference trajectory (6001 points)")
    osr = _engine("traj_fieldmaps")
    osr.calculate_trajectory(gpu=1)
    td, dt = osr.trajectory_array()
    g = golden("traj_fieldmaps")
    ref = g["traj"]
    assert td.shape == ref.shape and dt == float(g["dt"])
    for k, scale in ((0, 1e-5), (1, 1e-7), (2, 1.0), (3, 1e-3), (4, 1e-6), (5, 1.0)):
        err = np.max(np.abs(td[k] - ref[k])) / scale
        say("  component %d max deviation / scale %.3e" % (k, err))
        assert err < 1e-12


def test_flux_through_field_maps_routes_match_reference():
    from srsoriginal_amd import abi
    from srsoriginal_amd.lib import _sr
    say("flux 9x9 at 400 eV through field maps: host-RK4 route and device-RK4 batch vs the reference")
    osr = _engine("flux_fieldmaps")
    a = cases.args("flux_fieldmaps")
    ref = golden("flux_fieldmaps")["container"][:81, 3]
    out = osr.calculate_flux_rectangle(energy_eV=400.0, npoints=[9, 9], plane="XY", width=[0.006, 0.006],
                                       translation=[0, 0, 30], gpu=1)
    v = np.array([x for _, x in out])[:81]
    err = normwise(v, ref)
    say("  per-particle route normwise %.3e" % err)
    assert err <= 1e-10
    st = osr._e.particle_state()
    c0, c1, n = cases.ct(a)
    row = dict(x0=st["x0"], beta0=st["beta0"], qmg=st["qmg"], t0=st["t0"], ct_start=c0, ct_stop=c1, npoints=n)
    q, cur = osr._e.particle_charge_current()
    xyz = _sr.Surface.rectangle("XY", 9, 9, 0.006, 0.006, [0, 0, 0], [0, 0, 30], 0).xyz
    for prec, tol in ((abi.FP64, 1e-10), (abi.MIXED, 1e-5)):
        got = abi.batch_flux(osr._e.undulators(), abi.particles([row]), q, cur, xyz, [400.0], 1.0, prec, 0)[0]
        err = normwise(got, ref)
        say("  device RK4 batch (%s) normwise %.3e" % ("fp64" if prec == abi.FP64 else "mixed", err))
        assert err <= tol
