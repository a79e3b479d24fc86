"""Writes the golden fixtures in tests/golden/ from the reference engine.

The reference OSCARS SR C++ engine is compiled from its own sources under
/root/reference/src (``make -C oracle ref``) and driven by oracle/ref_harness.cc;
this script runs that binary on each case below and stores inputs and outputs
as ``<case>.npz`` (plain arrays, no pickles) plus ``manifest.json``.

Run only where /root/reference exists:

    make -C oracle ref && python tests/golden/make_golden.py [--only case,...]

Large outputs (full trajectories) are stored as a SHA-256 digest of the raw
little-endian float64 bytes plus strided samples; everything else is stored
whole.
"""
import argparse
import hashlib
import json
import os
import platform
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")

# Undulators used by the configs (bfield, period, nperiods[, phase]).
UND_C0 = "0,1,0,0,0,0.05,41"  # By = 1 T sin(2 pi z / 50 mm), 41 periods
UND_EPU_Y = "0,0.6,0,0,0,0.057,30"
UND_EPU_X = "0.6,0,0,0,0,0.057,30,-1.5707963267948966"  # phase -pi/2 -> +pi/2 shift
BEAM = ["x0=0,0,-1.5", "d0=0,0,1", "energy=3", "current=0.5"]
BEAM_C3 = BEAM + ["sigE=0.00267", "beta=1.5,0.8", "emit=0.9e-9,8e-12", "hdir=1,0,0", "latref=0,0,0"]


def run(mode, args):
    with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as f:
        path = f.name
    cmd = [HARNESS, mode] + list(args) + ["out=" + path]
    t0 = time.time()
    subprocess.run(cmd, check=True, cwd=HERE)  # field-map paths are relative to tests/golden
    dt = time.time() - t0
    a = np.fromfile(path, dtype="<f8")
    os.unlink(path)
    return a, dt


def digest(arr):
    return hashlib.sha256(np.ascontiguousarray(arr, dtype="<f8").tobytes()).hexdigest()


def kv(args):
    out = {}
    for a in args:
        k, v = a.split("=", 1)
        out[k] = v
    return out


def container(a):
    n = int(a[0])
    return a[1:1 + 4 * n].reshape(n, 4)


CASES = {}


def case(fn):
    CASES[fn.__name__] = fn
    return fn


@case
def traj_c0_2001():
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=2001"]
    a, dt = run("traj", args)
    n = int(a[0])
    return args, {"n": np.array(n), "dt": np.array(a[1]), "traj": a[2:].reshape(9, n)}, dt


def _traj_digest(args):
    a, dt = run("traj", args)
    n = int(a[0])
    t = a[2:].reshape(9, n)
    idx = np.arange(0, n, max(1, n // 400))
    return {"n": np.array(n), "dt": np.array(a[1]),
            "digest_xb": np.array(digest(t[:6])), "digest_a_tail": np.array(digest(t[6:, 1:])),
            "sample_idx": idx, "samples": t[:, idx], "a_point0": t[6:, 0]}, dt


@case
def traj_c0_20001():
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=20001"]
    out, dt = _traj_digest(args)
    return args, out, dt


@case
def traj_epu_40001():
    args = ["und1=" + UND_EPU_Y, "und2=" + UND_EPU_X] + BEAM + ["ct=0,3", "npoints=40001"]
    out, dt = _traj_digest(args)
    return args, out, dt


@case
def spectrum_c0_small():
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=2001", "ne=200", "erange=100,3000", "obs=0,0,30", "nthreads=1"]
    a, dt = run("spectrum", args)
    return args, {"spectrum": a[1:].reshape(-1, 2)}, dt


@case
def spectrum_c0_full():
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=20001", "ne=2000", "erange=100,3000", "obs=0,0,30", "nthreads=8"]
    a, dt = run("spectrum", args)
    return args, {"spectrum": a[1:].reshape(-1, 2)}, dt


@case
def spectrum_points_offaxis():
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=4001", "points_ev=140,143.7,287.5,431.2,1000.5", "obs=0.001,-0.0005,30", "nthreads=1"]
    a, dt = run("spectrum", args)
    return args, {"spectrum": a[1:].reshape(-1, 2)}, dt


def first_harmonic_energy():
    """E* = C0 grid energy with the largest on-axis flux in 120..170 eV (first harmonic)."""
    path = os.path.join(HERE, "spectrum_c0_full.npz")
    s = np.load(path)["spectrum"]
    m = (s[:, 0] > 120) & (s[:, 0] < 170)
    i = np.argmax(np.where(m, s[:, 1], -1))
    return float(s[i, 0])


@case
def flux_c1_small():
    e = first_harmonic_energy()
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=20001", "plane=XY", "width=0.02,0.02", "nx=21,21", "trans=0,0,30", "energy_ev=%.17g" % e, "nthreads=8"]
    a, dt = run("flux", args)
    return args, {"container": container(a), "energy_ev": np.array(e)}, dt


@case
def flux_c1_full():
    e = first_harmonic_energy()
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=20001", "plane=XY", "width=0.02,0.02", "nx=201,201", "trans=0,0,30", "energy_ev=%.17g" % e, "nthreads=8"]
    a, dt = run("flux", args)
    c = container(a)
    return args, {"value": c[:, 3], "energy_ev": np.array(e)}, dt


@case
def flux_1thread_route():
    # nthreads=1 route: CalculateFlux(P,...) adds N more container points (OSCARSSR.cc:2352-2358)
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=2001", "plane=XY", "width=0.004,0.002", "nx=5,3", "trans=0,0,30", "energy_ev=431.2", "nthreads=1"]
    a, dt = run("flux", args)
    return args, {"container": container(a)}, dt


@case
def flux_threads_route_dim3():
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=2001", "plane=XZ", "width=0.004,0.002", "nx=5,3", "trans=0,0.001,30", "rot=0.1,0,0", "energy_ev=287.5", "nthreads=2", "dim=3"]
    a, dt = run("flux", args)
    return args, {"container": container(a)}, dt


@case
def flux_points_list():
    pts = "0,0,30,0,0,1,0.001,0,30,0,0,1,0,0.002,31,0,0,1,-0.003,0.001,29.5,0,0,1"
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=4001", "points=" + pts, "energy_ev=143.7", "nthreads=2"]
    a, dt = run("flux", args)
    return args, {"container": container(a)}, dt


@case
def pd_c2_small():
    args = ["und1=" + UND_EPU_Y, "und2=" + UND_EPU_X] + BEAM + ["ct=0,3", "npoints=40001", "plane=XY", "width=0.06,0.06", "nx=31,31", "trans=0,0,30", "nthreads=8"]
    a, dt = run("pd", args)
    return args, {"container": container(a)}, dt


@case
def pd_c2_full():
    args = ["und1=" + UND_EPU_Y, "und2=" + UND_EPU_X] + BEAM + ["ct=0,3", "npoints=40001", "plane=XY", "width=0.06,0.06", "nx=301,301", "trans=0,0,30", "nthreads=8"]
    a, dt = run("pd", args)
    c = container(a)
    return args, {"value": c[:, 3]}, dt


@case
def pd_c0_1thread_route():
    # 1-thread route ignores Weight and adds no duplicates beyond the pre-added N
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=2001", "plane=XY", "width=0.02,0.01", "nx=7,5", "trans=0,0,30", "nthreads=1"]
    a, dt = run("pd", args)
    return args, {"container": container(a)}, dt


@case
def pd_c0_threads_route():
    # Same configuration as pd_c0_1thread_route through the threads route; the 1-thread
    # route reads the uninitialised A/c of trajectory point 0 (OSCARSSR.cc:606) and
    # returned NaN when the fixture was written.
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=2001", "plane=XY", "width=0.02,0.01", "nx=7,5", "trans=0,0,30", "nthreads=2"]
    a, dt = run("pd", args)
    return args, {"container": container(a)}, dt


@case
def pd_directional_tilted():
    # Directional: N1.n <= 0 terms skipped, no |Sum| (OSCARSSR.cc:1549-1551, 1569-1573)
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=2001", "plane=XZ", "width=0.02,0.4", "nx=6,9", "trans=0.005,0,0.5", "rot=0,0,0", "normal=-1", "nthreads=3"]
    a, dt = run("pd", args)
    return args, {"container": container(a)}, dt


@case
def pd_directional_inside():
    # Screen inside the undulator, tilted, normal +z: trajectory points behind it
    # (N1.n <= 0) are skipped and the sign is kept (OSCARSSR.cc:1549-1551, 1569-1573).
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=3001", "plane=XY", "width=0.01,0.008", "nx=7,5",
                                        "trans=0.001,0,0.4", "rot=0.3,0.1,0", "normal=1", "nthreads=2"]
    a, dt = run("pd", args)
    return args, {"container": container(a)}, dt


@case
def totalpower_c0():
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=20001"]
    a, dt = run("totalpower", args)
    return args, {"total_power": np.array(a[0])}, dt


@case
def totalpower_c2():
    args = ["und1=" + UND_EPU_Y, "und2=" + UND_EPU_X] + BEAM + ["ct=0,3", "npoints=40001"]
    a, dt = run("totalpower", args)
    return args, {"total_power": np.array(a[0])}, dt


@case
def normals_seed1234():
    args = ["und1=" + UND_C0] + BEAM + ["seed=1234", "count=64"]
    a, dt = run("normals", args)
    return args, {"normals": a[1:]}, dt


@case
def beam_c3_seed1234():
    args = ["und1=" + UND_C0] + BEAM_C3 + ["seed=1234", "count=32"]
    a, dt = run("beam", args)
    return args, {"particles": a[1:].reshape(-1, 9)}, dt


@case
def multiflux_c3_small():
    args = ["und1=" + UND_C0] + BEAM_C3 + ["seed=777", "ct=0,3", "npoints=3001", "plane=XY", "width=0.01,0.01", "nx=9,9", "trans=0,0,30", "energy_ev=431.2", "nparticles=4", "nthreads=1"]
    a, dt = run("flux", args)
    return args, {"container": container(a)}, dt


# Undulator + Gaussian kicker (sigma along z only, rotated) + rotated uniform box
# (bounded along z only): TField3D_Gaussian.cc / TField3D_UniformBox.cc GetF, summed
# in insertion order (TFieldContainer.cc:107-117).
FIELDS_MIXED = ["und1=" + UND_C0, "gauss1=0,0.01,0,0,0,0.05,0,0,1.0,0,0.1,0",
                "box1=0.005,0,0.002,0,0,0.3,0,0,-1.0,0.05,0,0.02"]


@case
def traj_fields_mixed():
    args = FIELDS_MIXED + BEAM + ["ct=-0.5,3", "npoints=4001"]
    a, dt = run("traj", args)
    n = int(a[0])
    return args, {"n": np.array(n), "dt": np.array(a[1]), "traj": a[2:].reshape(9, n)}, dt


@case
def flux_fields_mixed():
    args = FIELDS_MIXED + BEAM + ["ct=-0.5,3", "npoints=4001", "plane=XY", "width=0.01,0.01", "nx=9,9", "trans=0.002,0,30", "energy_ev=143.52176088044024", "nthreads=2"]
    a, dt = run("flux", args)
    return args, {"container": container(a)}, dt

@case
def multispectrum_c3_small():
    # Gaussian beam, 4 particles: every route adds each particle's spectrum x 1/N
    # (OSCARSSR.cc:963-982)
    args = ["und1=" + UND_C0] + BEAM_C3 + ["seed=777", "ct=0,3", "npoints=3001", "ne=64", "erange=100,1000", "obs=0.001,0,30", "nparticles=4", "nthreads=1"]
    a, dt = run("spectrum", args)
    return args, {"spectrum": a[1:].reshape(-1, 2)}, dt

@case
def multipd_c3_small():
    # Gaussian beam, 4 particles, threads route: the reference CPU sums the particles'
    # power density without the 1/N weight (OSCARSSR.cc:1579-1585); the 1-thread route
    # would read the uninitialised A/c of point 0 (OSCARSSR.cc:606)
    args = ["und1=" + UND_C0] + BEAM_C3 + ["seed=777", "ct=0,3", "npoints=3001", "plane=XY", "width=0.04,0.02", "nx=9,5", "trans=0,0,30", "nparticles=4", "nthreads=2"]
    a, dt = run("pd", args)
    return args, {"container": container(a)}, dt

# ---- fixtures for paths the first set left unpinned ------------------------------------
# Beam along x (undulator period along x, screen in the YZ plane at x = 30 m): the
# cancellation beta - n moves to the x component (every other case has it in z).
UND_X = "0,1,0,0.05,0,0,41"
BEAM_X = ["x0=-1.5,0,0", "d0=1,0,0", "energy=3", "current=0.5"]
# Beam at 36.87 deg in the x-z plane (d0 = (0.6, 0, 0.8)), undulator period along d0.
UND_DIAG = "0,1,0,0.03,0,0.04,41"
BEAM_DIAG = ["x0=-0.9,0,-1.2", "d0=0.6,0,0.8", "energy=3", "current=0.5"]


def _diag_points(n=7, w=0.01):
    """Observation points on a square perpendicular to d0 = (0.6, 0, 0.8) at 30 m, normal d0."""
    u = np.array([0.6, 0.0, 0.8])
    e1 = np.array([0.8, 0.0, -0.6])
    e2 = np.array([0.0, 1.0, 0.0])
    pts = []
    for a in np.linspace(-w / 2, w / 2, n):
        for b in np.linspace(-w / 2, w / 2, n):
            x = 30.0 * u + a * e1 + b * e2
            pts += list(x) + list(u)
    return ",".join("%.17g" % v for v in pts)


@case
def flux_beam_x():
    e = first_harmonic_energy()
    args = ["und1=" + UND_X] + BEAM_X + ["ct=0,3", "npoints=4001", "plane=YZ", "width=0.02,0.02", "nx=15,13",
                                         "trans=30,0,0", "energy_ev=%.17g" % e, "nthreads=2"]
    a, dt = run("flux", args)
    return args, {"container": container(a), "energy_ev": np.array(e)}, dt


@case
def pd_beam_x():
    args = ["und1=" + UND_X] + BEAM_X + ["ct=0,3", "npoints=4001", "plane=YZ", "width=0.04,0.04", "nx=13,11",
                                         "trans=30,0,0", "nthreads=2"]
    a, dt = run("pd", args)
    return args, {"container": container(a)}, dt


@case
def spectrum_beam_x():
    args = ["und1=" + UND_X] + BEAM_X + ["ct=0,3", "npoints=4001", "ne=300", "erange=100,1500", "obs=30,0.0004,-0.0002",
                                         "nthreads=1"]
    a, dt = run("spectrum", args)
    return args, {"spectrum": a[1:].reshape(-1, 2)}, dt


@case
def flux_beam_diagonal():
    args = ["und1=" + UND_DIAG] + BEAM_DIAG + ["ct=0,3", "npoints=4001", "points=" + _diag_points(),
                                               "energy_ev=143.5", "nthreads=2"]
    a, dt = run("flux", args)
    return args, {"container": container(a)}, dt


@case
def pd_beam_diagonal():
    args = ["und1=" + UND_DIAG] + BEAM_DIAG + ["ct=0,3", "npoints=4001", "points=" + _diag_points(9, 0.04),
                                               "pnormal=0", "nthreads=2"]
    a, dt = run("pd", args)
    return args, {"container": container(a)}, dt


@case
def uniforms_seed4321():
    args = ["und1=" + UND_C0] + BEAM + ["seed=4321", "count=64"]
    a, dt = run("uniforms", args)
    return args, {"uniforms": a[1:]}, dt


# Two weighted beams (weights 1 : 3), the second at another energy and offset, both with the
# C3 beam's sigma settings: SetNewParticle() draws the beam by weight, then the particle.
BEAMS_TWO = BEAM_C3 + ["beam2=0.0002,0,-1.5,0,0,1,2.5,0.4", "weight1=1", "weight2=3"]


@case
def beams_two_weighted():
    args = ["und1=" + UND_C0] + BEAMS_TWO + ["seed=99", "count=48"]
    a, dt = run("beam", args)
    return args, {"particles": a[1:].reshape(-1, 9)}, dt


@case
def multiflux_two_beams():
    args = ["und1=" + UND_C0] + BEAMS_TWO + ["seed=31", "ct=0,3", "npoints=3001", "plane=XY", "width=0.01,0.01",
                                              "nx=9,9", "trans=0,0,30", "energy_ev=143.5", "nparticles=4", "nthreads=1"]
    a, dt = run("flux", args)
    return args, {"container": container(a)}, dt


_PTS = "0,0,0,0,0,1, 0.002,-0.001,0,0,0,1, -0.003,0.002,0.5,0,0,-1, 0.001,0.003,-0.4,1,0,0, 0,-0.002,0.2,0,0.6,0.8"


@case
def flux_points_rotated():
    # calculate_flux(points=..., rotations=..., translation=...): points and normals rotated, points translated
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=4001", "points=" + _PTS.replace(" ", ""),
                                        "prot=0.02,-0.03,0.4", "ptrans=0.001,-0.0005,30", "energy_ev=143.5",
                                        "nthreads=2"]
    a, dt = run("flux", args)
    return args, {"container": container(a)}, dt


@case
def pd_points_rotated_flip():
    # calculate_power_density(points=..., normal=-1): normals rotated then reversed, directional
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=4001", "points=" + _PTS.replace(" ", ""),
                                        "prot=0.02,-0.03,0.4", "ptrans=0.001,-0.0005,30", "pnormal=-1", "nthreads=2"]
    a, dt = run("pd", args)
    return args, {"container": container(a)}, dt


@case
def pd_points_rotated_undirected():
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=4001", "points=" + _PTS.replace(" ", ""),
                                        "prot=0.02,-0.03,0.4", "ptrans=0.001,-0.0005,30", "pnormal=0", "nthreads=2"]
    a, dt = run("pd", args)
    return args, {"container": container(a)}, dt


# x0x1x2 rectangle: corners rotated then translated (TSurfacePoints_Rectangle::Init(NX1, NX2, X0, X1, X2, Normal))
_CORNERS = "-0.004,-0.003,0,0.005,-0.003,0.001,-0.004,0.004,-0.0005"


@case
def flux_x0x1x2():
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=4001", "x0x1x2=" + _CORNERS, "rot=0.05,0.02,0.3",
                                        "trans=0.0005,0,30", "nx=7,5", "energy_ev=143.5", "nthreads=2"]
    a, dt = run("flux", args)
    return args, {"container": container(a)}, dt


@case
def pd_x0x1x2_normal():
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=4001", "x0x1x2=" + _CORNERS, "rot=0.05,0.02,0.3",
                                        "trans=0.0005,0,30", "nx=7,5", "normal=-1", "nthreads=2"]
    a, dt = run("pd", args)
    return args, {"container": container(a)}, dt


# Tapered, off-centre undulator (TField3D_IdealUndulator.cc:50-139): taper 0.03 per metre,
# centre (0.0001, 0, 0.12), phase 0.4
UND_TAPER = "0,1,0,0,0,0.05,41,0.4,0.03,0.0001,0,0.12"


@case
def traj_taper_center():
    args = ["und1=" + UND_TAPER] + BEAM + ["ct=0,3", "npoints=4001"]
    a, dt = run("traj", args)
    n = int(a[0])
    return args, {"n": np.array(n), "dt": np.array(a[1]), "traj": a[2:].reshape(9, n)}, dt


@case
def flux_taper_center():
    args = ["und1=" + UND_TAPER] + BEAM + ["ct=0,3", "npoints=4001", "plane=XY", "width=0.01,0.01", "nx=11,9",
                                           "trans=0,0,30", "energy_ev=143.5", "nthreads=2"]
    a, dt = run("flux", args)
    return args, {"container": container(a)}, dt


# ---- electric fields (OSCARSSR::DerivativesE / DerivativesEB, OSCARSSR.cc:690-794) ----------
# E only: an electric "undulator" Ex = 1.5e8 V/m sin(2 pi z / 50 mm) (the force of a 0.5 T
# magnetic one at 3 GeV) and a vertical Gaussian kick; E + B: the C0 undulator, a vertical
# electric undulator of 30 mm period and phase 0.3, and a longitudinal uniform box.
EFIELDS_ONLY = ["eund1=1.5e8,0,0,0,0,0.05,41", "egauss1=0,2e6,0,0,0,0.05,0,0,1.0"]
EFIELDS_MIXED = ["und1=" + UND_C0, "eund1=0,6e7,0,0,0,0.03,50,0.3", "ebox1=0,0,5e5,0,0,0.4,0,0,-1.0"]


@case
def traj_efield_only():
    args = EFIELDS_ONLY + BEAM + ["ct=0,3", "npoints=4001"]
    a, dt = run("traj", args)
    n = int(a[0])
    return args, {"n": np.array(n), "dt": np.array(a[1]), "traj": a[2:].reshape(9, n)}, dt


@case
def traj_efield_bfield():
    args = EFIELDS_MIXED + BEAM + ["ct=0,3", "npoints=4001"]
    a, dt = run("traj", args)
    n = int(a[0])
    return args, {"n": np.array(n), "dt": np.array(a[1]), "traj": a[2:].reshape(9, n)}, dt


@case
def flux_efield_only():
    args = EFIELDS_ONLY + BEAM + ["ct=0,3", "npoints=4001", "plane=XY", "width=0.01,0.01", "nx=9,9",
                                  "trans=0,0,30", "energy_ev=300", "nthreads=2"]
    a, dt = run("flux", args)
    return args, {"container": container(a)}, dt


@case
def flux_efield_bfield():
    args = EFIELDS_MIXED + BEAM + ["ct=0,3", "npoints=4001", "plane=XY", "width=0.01,0.01", "nx=9,9",
                                   "trans=0,0,30", "energy_ev=143.5", "nthreads=2"]
    a, dt = run("flux", args)
    return args, {"container": container(a)}, dt


@case
def pd_efield_bfield():
    args = EFIELDS_MIXED + BEAM + ["ct=0,3", "npoints=4001", "plane=XY", "width=0.03,0.03", "nx=9,7",
                                   "trans=0,0,30", "nthreads=2"]
    a, dt = run("pd", args)
    return args, {"container": container(a)}, dt


@case
def efield_values():
    """Summed electric fields of both sets at sample points (OSCARSSR::GetE)."""
    rng = np.random.default_rng(11)
    p = np.stack([rng.uniform(-0.002, 0.002, 64), rng.uniform(-0.002, 0.002, 64), rng.uniform(-1.4, 1.4, 64)], 1)
    path = os.path.join(HERE, "fields", "epoints.bin")
    p.astype("<f8").tofile(path)
    out, t = {"points": p}, 0.0
    for name, fl in (("only", EFIELDS_ONLY), ("mixed", EFIELDS_MIXED)):
        a, dt = run("efield_at", fl + ["pts=fields/epoints.bin"])
        out[name] = a[1:].reshape(-1, 3)
        t += dt
    os.unlink(path)
    return ["sets=only,mixed"], out, t


def _bytes(path):
    with open(path, "rb") as f:
        return np.frombuffer(f.read(), dtype=np.uint8).copy()


@case
def efield_time_c0():
    # the time-domain electric field of the C0 trajectory at an off-axis point
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=2001", "obs=0.002,-0.001,30"]
    a, dt = run("efield", args)
    return args, {"container": container(a)}, dt

def _flux1(args):
    """OSCARSSR::CalculateFlux1 through the harness: the container, then the trajectory
    it used (its first A/c reads 0: the harness zeroes the stack under the call)."""
    a, dt = run("flux1", args)
    c = container(a)
    r = a[1 + 4 * c.shape[0]:]
    n = int(r[0])
    t = r[2:].reshape(9, n)
    out = {"container": c, "dt": np.array(r[1]), "a_first": t[6:, 0].copy(), "digest_traj": np.array(digest(t))}
    return out, dt


@case
def flux1_c1_small():
    # the near- plus far-field form on the C1 geometry at the first harmonic
    e = first_harmonic_energy()
    args = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=20001", "plane=XY", "width=0.02,0.02", "nx=21,21",
                                        "trans=0,0,30", "energy_ev=%.17g" % e]
    out, dt = _flux1(args)
    out["energy_ev"] = np.array(e)
    return args, out, dt


@case
def flux1_epu_offresonance():
    # EPU (two crossed undulators), off resonance where the trajectory ends weigh more,
    # on a tilted XZ screen
    args = ["und1=" + UND_EPU_Y, "und2=" + UND_EPU_X] + BEAM + ["ct=0,3", "npoints=40001", "plane=XZ",
                                                                "width=0.01,0.02", "nx=9,7", "rot=0.2,0,0",
                                                                "trans=0,0.002,30", "energy_ev=300"]
    out, dt = _flux1(args)
    out["energy_ev"] = np.array(300.0)
    return args, out, dt


@case
def files_and_averages():
    """The reference's text / binary container files and its averaging of them:
    calculate_flux_rectangle / calculate_power_density_rectangle / calculate_spectrum
    with ofile (T3DScalarContainer::WriteToFileText, TSpectrumContainer::WriteToFileText),
    average_flux over two text files and over two binary files (AverageFromFilesText /
    AverageFromFilesBinary, Kahan sums of V/N) and average_spectra over two spectra.
    The fixture holds the files' bytes (inputs and outputs) and the averaged containers."""
    d = tempfile.mkdtemp()
    P = lambda n: os.path.join(d, n)
    base = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=2001", "plane=XY", "width=0.01,0.008", "nx=5,4", "trans=0,0,30", "nthreads=2"]
    t0 = time.time()
    run("flux", base + ["energy_ev=143.5", "ofile=" + P("f1.txt")])
    run("flux", base + ["energy_ev=150.0", "ofile=" + P("f2.txt")])
    run("pd", base + ["ofile=" + P("pd.txt")])
    a_txt, _ = run("avg3d", ["ifiles=%s,%s" % (P("f1.txt"), P("f2.txt")), "dim=2", "ofile=" + P("avg.txt"), "bofile=" + P("avg.bin")])
    run("avg3d", ["ifiles=" + P("f2.txt"), "dim=2", "bofile=" + P("avg2.bin")])
    a_bin, _ = run("avg3d", ["ifiles=%s,%s" % (P("avg.bin"), P("avg2.bin")), "binary=1", "dim=2", "ofile=" + P("avgb.txt")])
    sbase = ["und1=" + UND_C0] + BEAM + ["ct=0,3", "npoints=2001", "ne=30", "erange=100,400", "nthreads=1"]
    run("spectrum", sbase + ["obs=0,0,30", "ofile=" + P("s1.txt")])
    run("spectrum", sbase + ["obs=0.001,0,30", "ofile=" + P("s2.txt")])
    sa, _ = run("avgspec", ["ifiles=%s,%s" % (P("s1.txt"), P("s2.txt")), "ofile=" + P("savg.txt")])
    out = {k.replace(".", "_"): _bytes(P(k)) for k in ("f1.txt", "f2.txt", "pd.txt", "avg.txt", "avg.bin", "avg2.bin",
                                                        "avgb.txt", "s1.txt", "s2.txt", "savg.txt")}
    out["avg_text_container"] = container(a_txt)
    out["avg_binary_container"] = container(a_bin)
    out["avg_spectrum"] = sa[1:].reshape(-1, 2)
    shutil.rmtree(d)
    return base, out, time.time() - t0

def pole_deltas(seed=4, n=82, sigma=0.005):
    return np.random.default_rng(seed).normal(0.0, sigma, n)


@case
def poles_flux_and_fd_gradient():
    """Flux for a pole-scaled undulator and the reference's central-difference
    gradient of L = sum_j (F_j(delta) - F_j(0))^2 with h = 1e-6 (config-4 loss)."""
    d = pole_deltas()
    base = ["poles=0,1,0,0,0,0.05,41"] + BEAM + ["ct=0,3", "npoints=2001", "plane=XY", "width=0.005,0.005", "nx=5,5", "trans=0,0,30", "energy_ev=143.7", "nthreads=1"]

    def flux(delta):
        a, _ = run("flux", base + ["scales=" + ",".join("%.17g" % x for x in delta)])
        return container(a)[:, 3]

    t0 = time.time()
    f0 = flux(np.zeros_like(d))
    fd = flux(d)
    h = 1e-6
    grad = np.zeros_like(d)
    for p in range(len(d)):
        dp = d.copy(); dp[p] += h
        dm = d.copy(); dm[p] -= h
        grad[p] = (np.sum((flux(dp) - f0) ** 2) - np.sum((flux(dm) - f0) ** 2)) / (2 * h)
    return base, {"deltas": d, "flux_ideal": f0, "flux_scaled": fd, "loss": np.array(np.sum((fd - f0) ** 2)), "fd_grad": grad}, time.time() - t0



@case
def poles_pd_and_fd_gradient():
    """Power density for the pole-scaled undulator and the reference's central-difference
    gradient of L = sum_j (P_j(delta) - P_j(0))^2 with h = 1e-6 (the power-density loss)."""
    d = pole_deltas(seed=5)
    # threads route: the 1-thread route reads the uninitialised A/c of point 0 (OSCARSSR.cc:606)
    base = ["poles=0,1,0,0,0,0.05,41"] + BEAM + ["ct=0,3", "npoints=2001", "plane=XY", "width=0.02,0.02", "nx=5,5", "trans=0,0,30", "nthreads=2"]

    def pd(delta):
        a, _ = run("pd", base + ["scales=" + ",".join("%.17g" % x for x in delta)])
        return container(a)[:25, 3]  # the route appends 25 zero-valued duplicates

    t0 = time.time()
    p0 = pd(np.zeros_like(d))
    pdd = pd(d)
    h = 1e-6
    grad = np.zeros_like(d)
    for p in range(len(d)):
        dp = d.copy(); dp[p] += h
        dm = d.copy(); dm[p] -= h
        grad[p] = (np.sum((pd(dp) - p0) ** 2) - np.sum((pd(dm) - p0) ** 2)) / (2 * h)
    return base, {"deltas": d, "pd_ideal": p0, "pd_scaled": pdd, "loss": np.array(np.sum((pdd - p0) ** 2)), "fd_grad": grad}, time.time() - t0

# ---- field maps (TField3D_Grid readers, TField3D_Grid.cc:318-1146) ---------------------
# Small synthetic maps in each of the reference's formats, written here with fixed content
# (tests/golden/fields/), then read by the reference engine.  An undulator-like vertical
# field By = B0 sin(2 pi z / lam) with a transverse roll-off and a small Bx, Bz.
FIELDS = os.path.join(HERE, "fields")


def _map_value(x, y, z, b0=0.8, lam=0.04):
    k = 2 * np.pi / lam
    by = b0 * np.cos(30.0 * x) * np.cosh(20.0 * y) * np.sin(k * z) * np.exp(-(z / 0.22) ** 8)
    bx = 0.02 * b0 * np.sin(30.0 * x) * np.sin(k * z + 0.3) * np.exp(-(z / 0.22) ** 8)
    bz = 0.01 * b0 * y * 20.0 * np.cos(k * z) * np.exp(-(z / 0.22) ** 8)
    return bx, by, bz


def write_field_maps():
    os.makedirs(FIELDS, exist_ok=True)
    # OSCARS: comment, then x0, dx, nx, y0, dy, ny, z0, dz, nz, then Bx By Bz x-major
    def oscars(path, b0):
        with open(path, "w") as f:
            f.write("# synthetic undulator-like map, x-major\n")
            for v, c in ((-0.002, "X start"), (0.002, "X step"), (3, "X points"), (-0.001, "Y start"),
                         (0.001, "Y step"), (3, "Y points"), (-0.3, "Z start"), (0.002, "Z step"), (301, "Z points")):
                f.write("%r  # %s\n" % (v, c))
            for ix in range(3):
                for iy in range(3):
                    for iz in range(301):
                        bx, by, bz = _map_value(-0.002 + 0.002 * ix, -0.001 + 0.001 * iy, -0.3 + 0.002 * iz, b0)
                        f.write("%.12e %.12e %.12e\n" % (bx, by, bz))
    oscars(os.path.join(FIELDS, "map_oscars.dat"), 0.8)
    oscars(os.path.join(FIELDS, "map_oscars_gap10.dat"), 1.0)
    oscars(os.path.join(FIELDS, "map_oscars_gap20.dat"), 0.5)
    # OSCARS1D "Z Bx By": irregular z, written out of order, with a comment and a blank line
    rng = np.random.default_rng(5)
    z = np.sort(rng.uniform(-0.25, 0.25, 700))
    z[0], z[-1] = -0.25, 0.25
    perm = rng.permutation(len(z))
    with open(os.path.join(FIELDS, "map_oscars1d.dat"), "w") as f:
        f.write("# z [m]  Bx [T]  By [T], irregular spacing, unsorted\n")
        for n, i in enumerate(perm):
            if n == 100:
                f.write("\n# a comment inside the data\n")
            bx, by, bz = _map_value(0.0, 0.0, z[i], 0.6, 0.05)
            f.write("%.12e %.12e %.12e\n" % (z[i], bx, by))
    # SPECTRA: comment, "dx dy dz nx ny nz" (mm), then Bx By Bz; centred on the origin
    with open(os.path.join(FIELDS, "map_spectra.dat"), "w") as f:
        f.write("# SPECTRA-style map\n")
        f.write("1 1 2 1 1 251\n")
        for iz in range(251):
            bx, by, bz = _map_value(0.0, 0.0, -0.25 + 0.002 * iz, 0.3, 0.05)
            f.write("%.12e %.12e %.12e\n" % (bx, by, bz))
    # SRW: comment, then the nine header values each after the comment character
    with open(os.path.join(FIELDS, "map_srw.dat"), "w") as f:
        f.write("#Bx [T], By [T], Bz [T] on 3D mesh: inmost loop vs Z, outmost loop vs X\n")
        for v in (-0.001, 0.001, 3, 0.0, 0.0, 1, -0.2, 0.002, 201):
            f.write("#%r\n" % v)
        for ix in range(3):
            for iz in range(201):
                bx, by, bz = _map_value(-0.001 + 0.001 * ix, 0.0, -0.2 + 0.002 * iz, 0.4, 0.05)
                f.write("%.12e %.12e %.12e\n" % (bx, by, bz))


def _field_points(n=240, seed=9):
    """Sample points inside, outside and on the edges of the maps (x, y, z), as raw float64."""
    rng = np.random.default_rng(seed)
    p = np.stack([rng.uniform(-0.0025, 0.0025, n), rng.uniform(-0.0012, 0.0012, n), rng.uniform(-0.32, 0.32, n)], 1)
    p[:6] = [[0, 0, -0.3], [0, 0, 0.3], [-0.002, 0, 0], [0.002, 0.001, 0.1], [0, 0, 0], [1e-4, -2e-4, 0.2999]]
    path = os.path.join(HERE, "fields", "points.bin")
    p.astype("<f8").tofile(path)
    return p


MAP_CASES = {
    "oscars": ["gridfile1=fields/map_oscars.dat", "gridfmt1=OSCARS"],
    "oscars_rot_scaled": ["gridfile1=fields/map_oscars.dat", "gridfmt1=oscars", "gridrot1=0.01,-0.02,0.3",
                          "gridtrans1=0.0005,-0.0002,0.01", "gridscale1=1.1,0.9,1.05,1.5,0.75,2.0"],
    "oscars1d": ["gridfile1=fields/map_oscars1d.dat", "gridfmt1=OSCARS1D Z Bx By", "gridscale1=1,1,0.8"],
    "spectra": ["gridfile1=fields/map_spectra.dat", "gridfmt1=SPECTRA", "gridrot1=0,0,0.5"],
    "srw": ["gridfile1=fields/map_srw.dat", "gridfmt1=SRW", "gridtrans1=0,0,0.02"],
    "interpolated": ["gridmap=20:fields/map_oscars_gap20.dat;10:fields/map_oscars_gap10.dat", "gridmapfmt=OSCARS",
                     "gridparam=13.5"],
}


@case
def bfield_maps():
    """Fields of each map format at the sample points (OSCARSSR::GetB)."""
    write_field_maps()
    pts = _field_points()
    out, t = {"points": pts}, 0.0
    for name, args in MAP_CASES.items():
        a, dt = run("bfield", args + ["pts=fields/points.bin"])
        out[name] = a[1:].reshape(-1, 3)
        t += dt
    return ["maps=" + ",".join(MAP_CASES)], out, t


FIELDS_MAPS = ["gridfile1=fields/map_oscars.dat", "gridfmt1=OSCARS", "gridfile2=fields/map_oscars1d.dat",
               "gridfmt2=OSCARS1D Z Bx By", "gridtrans2=0,0,0.6", "gridscale2=1,1,0.8"]


@case
def traj_fieldmaps():
    write_field_maps()
    args = FIELDS_MAPS + BEAM + ["x0=0,0,-0.5", "ct=0,1.3", "npoints=6001"]
    a, dt = run("traj", args)
    n = int(a[0])
    return args, {"n": np.array(n), "dt": np.array(a[1]), "traj": a[2:].reshape(9, n)}, dt


@case
def flux_fieldmaps():
    write_field_maps()
    args = FIELDS_MAPS + BEAM + ["x0=0,0,-0.5", "ct=0,1.3", "npoints=6001", "plane=XY", "width=0.006,0.006",
                                 "nx=9,9", "trans=0,0,30", "energy_ev=400", "nthreads=2"]
    a, dt = run("flux", args)
    return args, {"container": container(a)}, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default="")
    ns = ap.parse_args()
    if not os.path.exists(HARNESS):
        sys.exit("build the reference harness first: make -C oracle ref")
    only = [c for c in ns.only.split(",") if c]
    mpath = os.path.join(HERE, "manifest.json")
    manifest = json.load(open(mpath)) if os.path.exists(mpath) else {"cases": {}}
    gcc = subprocess.run(["g++", "--version"], capture_output=True, text=True).stdout.splitlines()[0]
    manifest["generator"] = {"compiler": gcc, "flags": "-std=c++11 -O3 -pedantic -fPIC -pthread",
                             "libc": " ".join(platform.libc_ver()), "machine": platform.machine(),
                             "reference": "dhidas/SRSOriginal (OSCARS SR 1.31.05) src/*.cc minus the Python/CUDA files"}
    for name, fn in CASES.items():
        if only and name not in only:
            continue
        args, arrays, dt = fn()
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        manifest["cases"][name] = {"args": kv(args), "seconds": round(dt, 3)}
        print("%-32s %7.2fs" % (name, dt), flush=True)
    with open(mpath, "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
