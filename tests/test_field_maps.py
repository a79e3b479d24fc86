"""Field maps (TField3D_Grid): the four file formats and the parameter-interpolated map,
read by the host engine, against the fields and the trajectory the reference engine
computed from the same files (tests/golden/fields/, fixtures bfield_maps,
traj_fieldmaps).  Both are bitwise: the readers parse with the same iostream
extraction and the interpolation keeps the reference's operation order."""
import os
import sys

import numpy as np
import pytest

import cases
from conftest import golden

sys.path.insert(0, cases.GOLDEN)
from make_golden import MAP_CASES  # noqa: E402  (the per-map harness arguments)


def _kv(args):
    return dict(a.split("=", 1) for a in args)


def _osr_with(map_args):
    from srsoriginal_amd.sr import sr
    osr = sr()
    cases.add_field_maps(osr, _kv(map_args))
    return osr


@pytest.mark.parametrize("name", list(MAP_CASES))
def test_field_map_matches_reference_bitwise(name):
    g = golden("bfield_maps")
    osr = _osr_with(MAP_CASES[name])
    got = np.array([osr.get_bfield(list(p)) for p in g["points"]])
    ref = g[name]
    assert np.count_nonzero(ref) > 0
    bad = np.flatnonzero(np.any(got != ref, axis=1))
    assert bad.size == 0, "%s: %d points differ, first %s: %s vs %s" % (name, bad.size, g["points"][bad[0]],
                                                                          got[bad[0]], ref[bad[0]])


def test_trajectory_through_field_maps_matches_reference_bitwise():
    from test_host_engine import _engine
    osr = _engine("traj_fieldmaps")
    osr.calculate_trajectory()
    t, dt = osr.trajectory_array()
    g = golden("traj_fieldmaps")
    assert dt == float(g["dt"]) and t.shape[1] == int(g["n"])
    ref = g["traj"]
    assert np.array_equal(t[:6], ref[:6])
    assert np.array_equal(t[6:, 1:], ref[6:, 1:])  # A/c of point 0 is unset in the reference


def test_field_map_errors():
    from srsoriginal_amd.sr import sr
    osr = sr()
    with pytest.raises(ValueError, match="'ifile' or 'iformat' is blank"):
        osr.add_bfield_file(ifile="", iformat="OSCARS")
    with pytest.raises(ValueError, match="Could not import magnetic field"):
        osr.add_bfield_file(ifile=os.path.join(cases.GOLDEN, "fields", "map_oscars.dat"), iformat="NOSUCH")
    with pytest.raises(ValueError, match="Could not import magnetic field"):
        osr.add_bfield_file(ifile=os.path.join(cases.GOLDEN, "fields", "no_such_file.dat"), iformat="OSCARS")
    with pytest.raises(ValueError, match="Could not import magnetic field"):
        osr.add_bfield_file(ifile=os.path.join(cases.GOLDEN, "fields", "map_oscars.dat"), iformat="OSCARS1D Q By")
    maps = [[10, os.path.join(cases.GOLDEN, "fields", "map_oscars_gap10.dat")],
            [20, os.path.join(cases.GOLDEN, "fields", "map_oscars_gap20.dat")]]
    with pytest.raises(ValueError, match="Could not import magnetic field"):
        osr.add_bfield_interpolated(mapping=maps, iformat="OSCARS", parameter=25.0)  # outside the mapping
    with pytest.raises(ValueError, match="Could not import magnetic field"):
        osr.add_bfield_interpolated(mapping=maps, iformat="SRW", parameter=15.0)
    with pytest.raises(ValueError, match="Incorrect format in 'mapping'"):
        osr.add_bfield_interpolated(mapping=[[10]], iformat="OSCARS", parameter=15.0)
    assert osr._e.n_bfields() == 0


def test_interpolated_map_at_the_mapping_ends():
    """At a mapping value the interpolated map is that file's map: exactly at the lower
    end (the reference's unset-index case: the first map is taken), to rounding of
    F0 + ((F1 - F0)/dp) dp at the upper end."""
    f10 = os.path.join(cases.GOLDEN, "fields", "map_oscars_gap10.dat")
    f20 = os.path.join(cases.GOLDEN, "fields", "map_oscars_gap20.dat")
    from srsoriginal_amd.sr import sr
    for p, f in ((10.0, f10), (20.0, f20)):
        a, b = sr(), sr()
        a.add_bfield_interpolated(mapping=[[20, f20], [10, f10]], iformat="OSCARS", parameter=p)
        b.add_bfield_file(ifile=f, iformat="OSCARS")
        for x in ([0, 0, 0.013], [0.0011, -0.0004, -0.2], [-0.0015, 0.0009, 0.17]):
            ga, gb = np.array(a.get_bfield(x)), np.array(b.get_bfield(x))
            if p == 10.0:
                assert np.array_equal(ga, gb)
            else:
                assert np.all(np.abs(ga - gb) <= 4e-16 * np.abs(gb).max())
