"""Host side of the drop-in (srsoriginal_amd.sr over lib/_sr): trajectory,
fields, beams, surfaces, validation and error behaviour, on the CPU.  The
product's host arithmetic must match the reference bit for bit, because the GPU
phase integrals consume it."""
import hashlib
import os
import subprocess
import sys

import numpy as np
import pytest

import cases
from conftest import ROOT, golden

HAS_GPU = os.path.exists("/dev/kfd")


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def _engine(name):
    return cases.make_sr(name)


@pytest.mark.parametrize("name", ["traj_c0_20001", "traj_epu_40001"])
def test_host_trajectory_matches_reference_digest(name):
    osr = _engine(name)
    lst = osr.calculate_trajectory()
    t, dt = osr.trajectory_array()
    g = golden(name)
    assert len(lst) == int(g["n"]) and dt == float(g["dt"])
    assert _digest(t[:6]) == str(g["digest_xb"])
    assert _digest(t[6:, 1:]) == str(g["digest_a_tail"])
    assert lst[5][0] == list(t[0:3, 5]) and lst[5][1] == list(t[3:6, 5])


def test_host_fields_match_oracle(oracle):
    osr = _engine("traj_epu_40001")
    a = cases.args("traj_epu_40001")
    F = oracle.FieldSet(cases.undulators(a))
    for z in np.linspace(-1.2, 1.2, 997):
        x = [1e-5, -2e-6, float(z)]
        assert osr.get_bfield(x) == list(oracle.bfield(F, x))


def test_pole_scaled_field_matches_oracle(oracle):
    from srsoriginal_amd.sr import sr
    d = np.random.default_rng(4).normal(0, 0.005, 82)
    osr = sr()
    osr.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.05], nperiods=41, pole_scales=list(d))
    F = oracle.FieldSet([], dict(field=[0, 1, 0], period=[0, 0, 0.05], nperiods=41), d)
    for z in np.linspace(-1.1, 1.1, 1001):
        assert osr.get_bfield([0, 0, float(z)]) == list(oracle.bfield(F, [0, 0, float(z)]))
    with pytest.raises(ValueError, match="2\\*nperiods"):
        osr.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.05], nperiods=41, pole_scales=[0.1])


def test_surfaces_match_oracle(oracle):
    from srsoriginal_amd.lib import _sr
    for plane, rot, trans, normal in (("XY", [0, 0, 0], [0, 0, 30], 0), ("XZ", [0.1, 0, 0], [0, 0.001, 30], 1),
                                      ("zy", [0.02, -0.3, 0.7], [1, 2, 3], -1), ("YX", [0, 0, 0], [0, 0, 30], 0)):
        S = _sr.Surface.rectangle(plane, 5, 7, 0.004, 0.006, rot, trans, normal)
        xyz, nrm, x12 = oracle.rect_surface(plane, 5, 7, 0.004, 0.006, rot, trans, normal)
        assert np.array_equal(S.xyz, xyz) and np.array_equal(S.normals, nrm)


def test_random_particles_match_reference(tmp_path):
    # fresh interpreter: the global normal distribution caches its second value
    code = r"""
import sys, numpy as np
sys.path.insert(0, %r); sys.path.insert(0, %r)
import cases
from srsoriginal_amd.sr import sr
a = cases.args("beam_c3_seed1234")
osr = sr()
osr.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.05], nperiods=41)
osr.set_particle_beam(type="electron", name="b", energy_GeV=3, x0=[0, 0, -1.5], d0=[0, 0, 1], current=0.5,
                      sigma_energy_GeV=float(a["sigE"]), beta=cases.vec(a, "beta"), emittance=cases.vec(a, "emit"),
                      horizontal_direction=cases.vec(a, "hdir"), lattice_reference=cases.vec(a, "latref"))
osr.set_seed(1234)
rows = []
for _ in range(32):
    osr.set_new_particle()
    rows.append(osr.get_particle_x0() + osr.get_particle_beta0())
np.save(sys.argv[1], np.array(rows))
""" % (ROOT, os.path.join(ROOT, "tests"))
    out = str(tmp_path / "beam_rows.npy")
    subprocess.run([sys.executable, "-c", code, out], check=True)
    rows = np.load(out)
    ref = golden("beam_c3_seed1234")["particles"]
    assert np.array_equal(rows, ref[:, :6])


def test_two_weighted_beams_match_reference(tmp_path):
    """sr.add_particle_beam twice (weights 1 and 3), set_seed, set_new_particle(): the beam is
    drawn by weight (TParticleBeamContainer.cc:109-142) and then the particle, as the
    reference does after set_seed (fresh interpreter, as above); rand() continues the same
    uniform stream (TRandomA.cc:61-64)."""
    code = r"""
import sys, numpy as np
sys.path.insert(0, %r); sys.path.insert(0, %r)
import cases
from srsoriginal_amd.sr import sr
a = cases.args("beams_two_weighted")
(b1, w1), (b2, w2) = cases.beams(a)
osr = sr()
osr.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.05], nperiods=41)
for name, b, w in (("b1", b1, w1), ("b2", b2, w2)):
    osr.add_particle_beam(type="electron", name=name, energy_GeV=b["energy_gev"], x0=b["x0"], d0=b["d0"],
                          current=b["current"], weight=w, sigma_energy_GeV=b["sigma_e_gev"], beta=b["beta"],
                          emittance=b["emit"], horizontal_direction=b["hdir"], lattice_reference=b["lattice_ref"])
osr.set_seed(int(a["seed"]))
rows = []
for _ in range(48):
    osr.set_new_particle()
    rows.append(osr.get_particle_x0() + osr.get_particle_beta0() + [osr.get_particle_e0()])
osr.set_seed(int(cases.args("uniforms_seed4321")["seed"]))
u = [osr.rand() for _ in range(64)]
np.save(sys.argv[1], np.array(rows))
np.save(sys.argv[2], np.array(u))
""" % (ROOT, os.path.join(ROOT, "tests"))
    out, uout = str(tmp_path / "rows.npy"), str(tmp_path / "u.npy")
    subprocess.run([sys.executable, "-c", code, out, uout], check=True)
    rows = np.load(out)
    ref = golden("beams_two_weighted")["particles"]
    assert np.array_equal(rows, ref[:, :7])
    e0 = ref[:, 6]  # both beams drawn: energies near 2.5 and near 3 GeV
    assert np.any(np.abs(e0 - 2.5) < 0.05) and np.any(np.abs(e0 - 3.0) < 0.05)
    # the uniform stream after another set_seed, without the normals' cache mattering
    assert np.array_equal(np.load(uout), golden("uniforms_seed4321")["uniforms"])


def test_total_power_matches_reference_and_oracle(oracle):
    osr = _engine("totalpower_c2")
    assert osr.calculate_total_power() == float(golden("totalpower_c2")["total_power"])
    osr = _engine("totalpower_c0")
    F, p, t, dt = cases.oracle_setup(oracle, "totalpower_c0")
    assert osr.calculate_total_power() == oracle.total_power(t, dt, p.q, p.current, p.gamma)


def test_validation_messages_follow_reference():
    from srsoriginal_amd.sr import sr
    osr = sr()
    with pytest.raises(ValueError, match="No particle beam defined"):
        osr.calculate_flux_rectangle(energy_eV=100, npoints=[3, 3], plane="XY", width=[0.01, 0.01])
    osr = _engine("traj_c0_2001")
    kw = dict(energy_eV=100, npoints=[3, 3], plane="XY", width=[0.01, 0.01], translation=[0, 0, 30])
    with pytest.raises(ValueError, match="'gpu' must be 0 or 1"):
        osr.calculate_flux_rectangle(gpu=2, **kw)
    with pytest.raises(ValueError, match="Both are not currently allowed"):
        osr.calculate_flux_rectangle(gpu=1, nthreads=2, **kw)
    with pytest.raises(ValueError, match="'nthreads' must be > 0"):
        osr.calculate_flux_rectangle(nthreads=-1, **kw)
    with pytest.raises(ValueError, match="'dim' must be 2 or 3"):
        osr.calculate_flux_rectangle(dim=4, **kw)
    with pytest.raises(ValueError, match="an entry in 'npoints' is <= 0"):
        osr.calculate_flux_rectangle(energy_eV=100, npoints=[0, 3], plane="XY", width=[0.01, 0.01])
    with pytest.raises(ValueError, match="not a valid surface string"):
        osr.calculate_flux_rectangle(energy_eV=100, npoints=[3, 3], plane="QQ", width=[0.01, 0.01])
    with pytest.raises(ValueError, match="normal must be -1, 0, or 1"):
        osr.calculate_power_density_rectangle(npoints=[3, 3], plane="XY", width=[0.01, 0.01], normal=2)
    with pytest.raises(ValueError, match="'energy_range_eV' must be a list of length 2"):
        osr.calculate_spectrum(obs=[0, 0, 30], energy_range_eV=[1, 2, 3], npoints=4)
    with pytest.raises(ValueError, match="'npoints' must be > 0"):
        osr.calculate_spectrum(obs=[0, 0, 30], energy_range_eV=[1, 2])
    with pytest.raises(ValueError, match="Incorrect format in 'obs'"):
        osr.calculate_spectrum(obs=[0, 30], energy_range_eV=[1, 2], npoints=3)
    with pytest.raises(ValueError, match="'precision' must be"):
        osr.calculate_spectrum(obs=[0, 0, 30], energy_range_eV=[1, 2], npoints=3, precision="half")
    with pytest.raises(ValueError, match="'particle' must be 'ideal' or 'random'"):
        osr.set_new_particle(particle="other")
    with pytest.raises(ValueError, match="possibly 'name' already exists"):
        osr.add_particle_beam(type="electron", name="beam", energy_GeV=3, d0=[0, 0, 1], x0=[0, 0, -1.5])


def test_trajectory_errors():
    from srsoriginal_amd.sr import sr
    osr = sr()
    osr.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.05], nperiods=41)
    osr.set_particle_beam(type="electron", name="b", energy_GeV=3, x0=[0, 0, -1.5], d0=[0, 0, 1], current=0.5)
    with pytest.raises(ValueError, match="start time is greater than stop time"):
        osr.calculate_trajectory()
    osr.set_ctstartstop(1, 3)
    with pytest.raises(ValueError, match="start time is greater than T0"):
        osr.calculate_trajectory()
    # set_ctstartstop resets the count to 10000 points per metre (OSCARSSR.cc:448-458)
    osr.set_ctstartstop(0, 3)
    assert osr.get_npoints_trajectory() == 30000


@pytest.mark.skipif(HAS_GPU, reason="checks the no-GPU path")
def test_integrals_without_gpu_raise():
    osr = _engine("traj_c0_2001")
    with pytest.raises(ValueError, match="none were found"):
        osr.calculate_flux_rectangle(energy_eV=100, npoints=[3, 3], plane="XY", width=[0.01, 0.01],
                                     translation=[0, 0, 30], gpu=1)


def test_host_trajectory_gaussian_and_box_fields_bitwise():
    """Undulator + rotated Gaussian kicker + rotated uniform box, summed in insertion
    order, both RK4 legs: the host engine's trajectory equals the reference's bit for bit
    (A/c except the uninitialised point 0)."""
    osr = _engine("traj_fields_mixed")
    osr.calculate_trajectory()
    t, dt = osr.trajectory_array()
    g = golden("traj_fields_mixed")
    assert dt == float(g["dt"])
    ref = g["traj"]
    assert t.shape == ref.shape
    assert np.array_equal(t[:6], ref[:6])
    nb = int(np.argmin(np.abs(t[2] - (-1.5))))  # T0 point (A/c uninitialised there in the reference)
    a = np.delete(t[6:], nb, axis=1)
    r = np.delete(ref[6:], nb, axis=1)
    assert np.array_equal(a, r)
    # the fields are where the fixture puts them
    assert osr.get_bfield([0, 0, 1.0])[1] > 0.009 and abs(osr.get_bfield([0, 0, -1.0])[0] - 0.005) < 1e-3


def test_field_validation():
    from srsoriginal_amd.sr import sr
    osr = sr()
    with pytest.raises(ValueError, match="Incorrect format in 'sigma'"):
        osr.add_bfield_gaussian(bfield=[0, 1, 0], sigma=[1, 2])
    with pytest.raises(ValueError, match="Incorrect format in 'bfield'"):
        osr.add_bfield_uniform(bfield=[0, 1])


def test_tapered_offcentre_undulator_trajectory_bitwise():
    # taper, centre and phase of TField3D_IdealUndulator (.cc:50-139) through add_bfield_undulator
    osr = _engine("traj_taper_center")
    osr.calculate_trajectory()
    t, dt = osr.trajectory_array()
    g = golden("traj_taper_center")
    assert dt == float(g["dt"])
    assert np.array_equal(t[:6], g["traj"][:6]) and np.array_equal(t[6:, 1:], g["traj"][6:, 1:])


def test_three_corner_rectangle_matches_oracle(oracle):
    from srsoriginal_amd.lib import _sr
    for name in ("flux_x0x1x2", "pd_x0x1x2_normal"):
        a = cases.args(name)
        n = cases.vec(a, "nx")
        x0, x1, x2 = cases.corners(a)
        normal = int(float(a.get("normal", 0)))
        S = _sr.Surface.rectangle3(int(n[0]), int(n[1]), x0, x1, x2, normal)
        xyz, nrm, x12 = oracle.rect3_surface(int(n[0]), int(n[1]), x0, x1, x2, normal)
        assert np.array_equal(S.xyz, xyz) and np.array_equal(S.normals, nrm)


@pytest.mark.parametrize("name", ["traj_efield_only", "traj_efield_bfield"])
def test_electric_field_trajectories_bitwise(name):
    # add_efield_undulator / _gaussian / _uniform (OSCARSSR_Python.cc:994-1240) and the
    # E / E+B derivatives (OSCARSSR.cc:690-794)
    osr = _engine(name)
    osr.calculate_trajectory()
    t, dt = osr.trajectory_array()
    g = golden(name)
    assert dt == float(g["dt"])
    assert np.array_equal(t[:6], g["traj"][:6]) and np.array_equal(t[6:, 1:], g["traj"][6:, 1:])


def test_electric_field_values_and_clearing():
    g = golden("efield_values")
    for key, name in (("only", "traj_efield_only"), ("mixed", "traj_efield_bfield")):
        osr = _engine(name)
        got = np.array([osr.get_efield(list(x)) for x in g["points"]])
        assert np.array_equal(got, g[key])
    osr.clear_efields()
    assert osr.get_efield([0, 0, 0.1]) == [0.0, 0.0, 0.0]
    # add_field_gaussian: an efield is added; a bfield adds nothing (the reference's branch is commented out)
    osr.add_field_gaussian(sigma=[0, 0, 0.1], efield=[0, 1e6, 0])
    assert osr.get_efield([0, 0, 0]) == [0.0, 1e6, 0.0]
    osr.add_field_gaussian(sigma=[0, 0, 0.1], bfield=[0, 1, 0])
    assert osr.get_efield([0, 0, 0]) == [0.0, 1e6, 0.0]
    with pytest.raises(ValueError, match="both"):
        osr.add_field_gaussian(sigma=[0, 0, 0.1], bfield=[0, 1, 0], efield=[0, 1, 0])
