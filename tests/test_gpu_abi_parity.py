"""GPU parity of the HIP kernels, called through the C ABI (include/srs_hip.h),
against the reference's own FP64 CPU results (golden fixtures written by the
compiled reference) and the CPU oracle.

Tolerances: plain FP64 form <= 1e-10 normwise relative
(max|d|/max|ref|); mixed form <= 1e-5 normwise and RMS relative vs FP64.
"""
import numpy as np
import pytest

import cases
from conftest import golden
from gpu_util import normwise, pointwise, rms_rel, say

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-10
TOL_MIXED = 1e-5


@pytest.fixture(scope="module")
def abi():
    from srsoriginal_amd import abi as A
    assert A.device_count() >= 1, "no gfx950 device visible"
    return A


def _rect(oracle, name):
    r = cases.rect(cases.args(name))
    return oracle.rect_surface(r["plane"], r["nx1"], r["nx2"], r["w1"], r["w2"], r["rot"], r["trans"], r["normal"])


def test_spectrum_c0_full_fp64_matches_reference(abi, oracle):
    say("spectrum C0 fp64: NT=20001, NE=2000")
    F, p, t, dt = cases.oracle_setup(oracle, "spectrum_c0_full")
    s = golden("spectrum_c0_full")["spectrum"]
    got = abi.spectrum(t, dt, (0, 0, 30), s[:, 0], p.q, p.current, abi.FP64)
    err = normwise(got, s[:, 1])
    say("  normwise %.3e  pointwise(>=1e-3 max) %.3e" % (err, pointwise(got, s[:, 1])))
    assert err <= TOL_FP64
    assert pointwise(got, s[:, 1]) <= 1e-8


def test_spectrum_offaxis_and_small_fp64(abi, oracle):
    for name in ("spectrum_points_offaxis", "spectrum_c0_small"):
        say("spectrum fp64 " + name)
        F, p, t, dt = cases.oracle_setup(oracle, name)
        s = golden(name)["spectrum"]
        obs = cases.vec(cases.args(name), "obs", [0, 0, 30])
        got = abi.spectrum(t, dt, obs, s[:, 0], p.q, p.current, abi.FP64)
        assert normwise(got, s[:, 1]) <= TOL_FP64


def test_flux_c1_full_fp64_matches_reference(abi, oracle):
    say("flux C1 fp64: NT=20001, 201x201")
    F, p, t, dt = cases.oracle_setup(oracle, "flux_c1_full")
    g = golden("flux_c1_full")
    xyz, nrm, x12 = _rect(oracle, "flux_c1_full")
    got = abi.flux(t, dt, xyz, [float(g["energy_ev"])], p.q, p.current, abi.FP64)[0]
    err = normwise(got, g["value"])
    say("  normwise %.3e  pointwise %.3e" % (err, pointwise(got, g["value"])))
    assert err <= TOL_FP64


def test_flux_point_list_and_routes_fp64(abi, oracle):
    for name in ("flux_points_list", "flux_1thread_route", "flux_threads_route_dim3", "flux_c1_small"):
        say("flux fp64 " + name)
        F, p, t, dt = cases.oracle_setup(oracle, name)
        a = cases.args(name)
        if "points" in a:
            xyz, nrm = cases.points(a)
        else:
            xyz, nrm, x12 = _rect(oracle, name)
        e = float(golden(name)["energy_ev"]) if "energy_ev" in golden(name).files else float(a["energy_ev"])
        ref = golden(name)["container"][:len(xyz), 3]
        got = abi.flux(t, dt, xyz, [e], p.q, p.current, abi.FP64)[0]
        assert normwise(got, ref) <= TOL_FP64


def test_flux_multi_energy_groups_fp64(abi, oracle):
    # 6 energies -> groups of 4 + 2 inside the library; ragged point count (not a multiple of 256)
    say("flux fp64 6 energies x 77 points")
    F, p, t, dt = cases.oracle_setup(oracle, "traj_c0_2001")
    xyz, nrm, x12 = oracle.rect_surface("XY", 7, 11, 0.004, 0.006, (0, 0, 0), (0, 0, 30), 0)
    es = [143.5, 150.0, 287.1, 431.2, 700.0, 2500.0]
    got = abi.flux(t, dt, xyz, es, p.q, p.current, abi.FP64)
    for k, e in enumerate(es):
        ref = oracle.flux(t, dt, xyz, e, p.q, p.current)
        assert normwise(got[k], ref) <= TOL_FP64


def test_single_point_edges_fp64(abi, oracle):
    say("edges: 1 observation point, 1 energy, short trajectories")
    F, p, t, dt = cases.oracle_setup(oracle, "traj_c0_2001")
    for n in (1, 2, 255, 257):
        tt = np.ascontiguousarray(t[:, :n])
        obs = np.array([[0.001, -0.002, 30.0]])
        ref = oracle.flux(tt, dt, obs, 300.0, p.q, p.current)
        got = abi.flux(tt, dt, obs, [300.0], p.q, p.current, abi.FP64)[0]
        assert normwise(got, ref) <= TOL_FP64
        sref = oracle.spectrum(tt, dt, (0, 0, 30), [300.0], p.q, p.current)
        sgot = abi.spectrum(tt, dt, (0, 0, 30), [300.0], p.q, p.current, abi.FP64)
        assert normwise(sgot, sref) <= TOL_FP64


def test_power_density_c2_full_fp64_matches_reference(abi, oracle):
    say("power density C2 fp64: NT=40001, 301x301")
    F, p, t, dt = cases.oracle_setup(oracle, "pd_c2_full")
    xyz, nrm, x12 = _rect(oracle, "pd_c2_full")
    ref = golden("pd_c2_full")["value"][:len(xyz)]  # threads-route container: N values then N duplicates
    got = abi.power_density(t, dt, xyz, nrm, False, p.q, p.current, abi.FP64)
    err = normwise(got, ref)
    say("  normwise %.3e  pointwise %.3e" % (err, pointwise(got, ref)))
    assert err <= TOL_FP64


def test_power_density_directional_and_threads_route_fp64(abi, oracle):
    for name, directional in (("pd_directional_inside", True), ("pd_directional_tilted", True),
                              ("pd_c0_threads_route", False), ("pd_c2_small", False)):
        say("power density fp64 " + name)
        F, p, t, dt = cases.oracle_setup(oracle, name)
        xyz, nrm, x12 = _rect(oracle, name)
        ref = golden(name)["container"][:len(xyz), 3]
        got = abi.power_density(t, dt, xyz, nrm, directional, p.q, p.current, abi.FP64)
        assert normwise(got, ref) <= TOL_FP64


def test_mixed_flux_c1_within_1e5(abi, oracle):
    say("flux C1 mixed vs fp64")
    F, p, t, dt = cases.oracle_setup(oracle, "flux_c1_full")
    g = golden("flux_c1_full")
    xyz, nrm, x12 = _rect(oracle, "flux_c1_full")
    got = abi.flux(t, dt, xyz, [float(g["energy_ev"])], p.q, p.current, abi.MIXED)[0]
    e1, e2 = normwise(got, g["value"]), rms_rel(got, g["value"])
    say("  mixed normwise %.3e rms %.3e" % (e1, e2))
    assert e1 <= TOL_MIXED and e2 <= TOL_MIXED


def test_mixed_flux_three_energies_3rd_harmonic(abi, oracle):
    say("flux mixed 3 energies near the 3rd harmonic vs oracle")
    F, p, t, dt = cases.oracle_setup(oracle, "traj_c0_20001")
    xyz, nrm, x12 = oracle.rect_surface("XY", 21, 21, 0.01, 0.01, (0, 0, 0), (0, 0, 30), 0)
    es = [430.0, 431.2, 432.4]
    got = abi.flux(t, dt, xyz, es, p.q, p.current, abi.MIXED)
    for k, e in enumerate(es):
        ref = oracle.flux(t, dt, xyz, e, p.q, p.current, nthreads=8)
        assert normwise(got[k], ref) <= TOL_MIXED and rms_rel(got[k], ref) <= TOL_MIXED


def test_mixed_spectrum_c0_within_1e5(abi, oracle):
    say("spectrum C0 mixed vs reference fp64")
    F, p, t, dt = cases.oracle_setup(oracle, "spectrum_c0_full")
    s = golden("spectrum_c0_full")["spectrum"]
    got = abi.spectrum(t, dt, (0, 0, 30), s[:, 0], p.q, p.current, abi.MIXED)
    e1, e2 = normwise(got, s[:, 1]), rms_rel(got, s[:, 1])
    say("  mixed normwise %.3e rms %.3e" % (e1, e2))
    assert e1 <= TOL_MIXED and e2 <= TOL_MIXED


# The mixed spectrum reproduces the reference's FP64 phase RN(omega tau) for every
# energy (the recurrence over a run of 8 energies plus an exact FP64 correction), so it
# stays within ~2e-7 of the FP64 result where the FP64 result itself is 1e-5..1e-4 away
# from the integral evaluated in extended precision (high harmonics at 10 keV).
TOL_MIXED_SPECTRUM = 2e-6
SPEC_OBSERVERS = [(0.0, 0.0, 30.0), (1.2e-3, -0.4e-3, 30.0), (-2.0e-3, 1.5e-3, 30.0)]
SPEC_GRIDS = {"c0 grid 101-3001 eV": None, "100-10000 eV": (100.0, 10000.0)}


@pytest.mark.parametrize("grid", list(SPEC_GRIDS))
@pytest.mark.parametrize("obs", SPEC_OBSERVERS)
def test_spectrum_observers_and_10kev(abi, oracle, obs, grid):
    F, p, t, dt = cases.oracle_setup(oracle, "spectrum_c0_full")
    s = golden("spectrum_c0_full")["spectrum"]
    lim = SPEC_GRIDS[grid]
    es = s[:, 0].copy() if lim is None else lim[0] + (lim[1] - lim[0]) / 1999 * np.arange(2000)
    f64 = abi.spectrum(t, dt, obs, es, p.q, p.current, abi.FP64)
    mix = abi.spectrum(t, dt, obs, es, p.q, p.current, abi.MIXED)
    sub = slice(3, None, 50)  # 40 energies through the FP64 oracle (the reference's arithmetic)
    ref = oracle.spectrum(t, dt, obs, es[sub], p.q, p.current, nthreads=8)
    e64 = normwise(f64[sub], ref)
    e1, e2 = normwise(mix, f64), rms_rel(mix, f64)
    say("spectrum %s at %s: fp64 vs oracle %.3e; mixed vs fp64 normwise %.3e rms %.3e" % (grid, obs, e64, e1, e2))
    assert e64 <= TOL_FP64
    assert e1 <= TOL_MIXED_SPECTRUM and e2 <= TOL_MIXED_SPECTRUM


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_spectrum_far_observer_and_nonuniform_grid(abi, oracle, prec):
    """Runs of 8 energies need |tau| |theta bound| within the correction's series range:
    an observer 3 km away (tau ~ 1e-5 s) makes the kernel report it and the call is
    evaluated one energy per lane; a geometric energy grid is not a uniform run and goes
    that way from the start.  Both against the oracle."""
    P = abi.FP64 if prec == "fp64" else abi.MIXED
    tol = TOL_FP64 if prec == "fp64" else TOL_MIXED_SPECTRUM
    F, p, t, dt = cases.oracle_setup(oracle, "spectrum_c0_full")
    s = golden("spectrum_c0_full")["spectrum"]
    for obs, es in (((0.0, 0.0, 3000.0), s[::10, 0].copy()), ((0.05, 0.0, 3000.0), s[::10, 0].copy()),
                    ((0.0, 0.0, 30.0), np.geomspace(150.0, 2900.0, 301))):
        ref = oracle.spectrum(t, dt, obs, es, p.q, p.current, nthreads=8)
        got = abi.spectrum(t, dt, obs, es, p.q, p.current, P)
        err = normwise(got, ref)
        say("spectrum %s at %s, %d energies (%s grid): normwise vs oracle %.3e" %
            (prec, obs, len(es), "geometric" if obs[2] == 30.0 else "uniform", err))
        assert err <= tol


def test_mixed_power_density_c2_within_1e5(abi, oracle):
    say("power density C2 mixed vs reference fp64")
    F, p, t, dt = cases.oracle_setup(oracle, "pd_c2_full")
    xyz, nrm, x12 = _rect(oracle, "pd_c2_full")
    ref = golden("pd_c2_full")["value"][:len(xyz)]  # threads-route container: N values then N duplicates
    got = abi.power_density(t, dt, xyz, nrm, False, p.q, p.current, abi.MIXED)
    e1, e2 = normwise(got, ref), rms_rel(got, ref)
    # integrated power over the screen (dx dy sum, plots_mpl.py:474-489)
    area = (0.06 / 300) ** 2 * 1e6
    tp_ref, tp_got = ref.sum() * area, got.sum() * area
    say("  mixed normwise %.3e rms %.3e  total %.6f W vs %.6f W" % (e1, e2, tp_got, tp_ref))
    assert e1 <= TOL_MIXED and e2 <= TOL_MIXED
    assert abs(tp_got - tp_ref) / tp_ref <= TOL_MIXED


def test_mixed_power_density_wide_angles_backward_and_directional(abi, oracle):
    """Mixed power density away from the axis: screens at large angles, observers
    upstream of the trajectory (n_z < 0), random normals, directional on and off;
    each group is held to 1e-5 of its own largest value against the FP64 kernel,
    which is pinned to the reference at 1e-10.  The reference's screen inside the
    undulator (directional; part of the trajectory lies behind each point) is checked
    directly."""
    F, p, t, dt = cases.oracle_setup(oracle, "traj_c0_20001")
    rng = np.random.default_rng(11)
    groups = {
        "near axis": np.stack([rng.uniform(-0.02, 0.02, 300), rng.uniform(-0.02, 0.02, 300), np.full(300, 30.0)], 1),
        "wide": np.stack([rng.uniform(0.5, 3, 300) * rng.choice([-1, 1], 300),
                          rng.uniform(-3, 3, 300), rng.uniform(-6, 6, 300)], 1),
        "upstream": np.stack([rng.uniform(-0.5, 0.5, 300), rng.uniform(-0.5, 0.5, 300), rng.uniform(-30, -5, 300)], 1),
    }
    xyz = np.concatenate(list(groups.values()))
    nrm = rng.normal(size=xyz.shape)
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    nrm[:300] = (0, 0, 1)
    for directional in (False, True):
        ref = abi.power_density(t, dt, xyz, nrm, directional, p.q, p.current, abi.FP64)
        got = abi.power_density(t, dt, xyz, nrm, directional, p.q, p.current, abi.MIXED)
        for k, name in enumerate(groups):
            sl = slice(300 * k, 300 * (k + 1))
            err = normwise(got[sl], ref[sl])
            say("  %-9s directional=%d: mixed vs fp64 normwise %.3e" % (name, directional, err))
            assert err <= TOL_MIXED
    name = "pd_directional_inside"
    F, p, t, dt = cases.oracle_setup(oracle, name)
    xyz, nrm, x12 = _rect(oracle, name)
    ref = golden(name)["container"][:len(xyz), 3]
    got = abi.power_density(t, dt, xyz, nrm, True, p.q, p.current, abi.MIXED)
    say("  %s: mixed vs reference normwise %.3e" % (name, normwise(got, ref)))
    assert normwise(got, ref) <= TOL_MIXED


def test_abi_rejects_bad_arguments(abi, oracle):
    F, p, t, dt = cases.oracle_setup(oracle, "traj_c0_2001")
    with pytest.raises(abi.SrsError, match="precision"):
        abi.flux(t, dt, [[0, 0, 30]], [100.0], p.q, p.current, precision=7)
    with pytest.raises(abi.SrsError, match="no points in trajectory"):
        abi.flux(t[:, :0], dt, [[0, 0, 30]], [100.0], p.q, p.current)


def test_mixed_flux_paraxial_and_general_paths(abi, oracle):
    """The mixed flux takes the paraxial terms for blocks of points on one plane whose
    expansion bounds hold, and the plain mixed terms otherwise; both stay within 1e-5
    of the FP64 form.  Cases: a tilted screen (no common plane), a close screen
    (eps ~ 1e-4 fails the bound), a 300-point list whose first block mixes two planes
    and whose second block is flat, a flat screen just inside the eps bound, and a
    42 mm screen at 30 m whose corner blocks pass the eps bound for trajectory chunks
    far from the screen (rz ~ 31.5 m) and fail it for near ones (rz ~ 28.5 m), so their
    chunk sums come from both forms and must share one phase (the plane's omega Z/c)."""
    F, p, t, dt = cases.oracle_setup(oracle, "traj_c0_20001")
    es = [143.52176088044024, 431.2]
    tilted, _, _ = oracle.rect_surface("XY", 17, 15, 0.012, 0.01, (0.3, 0.1, 0), (0, 0, 30), 0)
    close, _, _ = oracle.rect_surface("XY", 17, 15, 0.02, 0.02, (0, 0, 0), (0, 0, 3.0), 0)
    rng = np.random.default_rng(11)
    two = np.column_stack([rng.uniform(-0.005, 0.005, 300), rng.uniform(-0.005, 0.005, 300),
                           np.where(np.arange(300) % 2 == 0, 30.0, 30.5)])
    two[256:, 2] = 30.0
    edge, _, _ = oracle.rect_surface("XY", 23, 23, 0.023, 0.023, (0, 0, 0), (0, 0, 20.0), 0)
    cross, _, _ = oracle.rect_surface("XY", 41, 41, 0.042, 0.042, (0, 0, 0), (0, 0, 30.0), 0)
    for name, xyz in (("tilted", tilted), ("close", close), ("two planes", two), ("edge of bound", edge),
                      ("bound crossed", cross)):
        f64 = abi.flux(t, dt, xyz, es, p.q, p.current, abi.FP64)
        mix = abi.flux(t, dt, xyz, es, p.q, p.current, abi.MIXED)
        for k, e in enumerate(es):
            e1, e2 = normwise(mix[k], f64[k]), rms_rel(mix[k], f64[k])
            say("  %-14s %7.2f eV: mixed vs fp64 normwise %.3e rms %.3e" % (name, e, e1, e2))
            assert e1 <= TOL_MIXED and e2 <= TOL_MIXED


@pytest.mark.parametrize("obs", [(0.0, 0.0, 2.0), (1.0e-3, 0.0, 5.0)])
def test_mixed_spectrum_close_observer_keeps_near_field(abi, oracle, obs):
    """Observers 0.5-6.5 m from the trajectory: at the low energies of the C0 grid
    (c/omega)/D exceeds 1e-9 on the tiles near the observer, so the mixed kernel keeps
    the x, y near-field terms there and leaves them out elsewhere; both against FP64,
    which the oracle pins."""
    F, p, t, dt = cases.oracle_setup(oracle, "spectrum_c0_full")
    es = golden("spectrum_c0_full")["spectrum"][:, 0].copy()
    f64 = abi.spectrum(t, dt, obs, es, p.q, p.current, abi.FP64)
    mix = abi.spectrum(t, dt, obs, es, p.q, p.current, abi.MIXED)
    sub = slice(1, None, 100)
    ref = oracle.spectrum(t, dt, obs, es[sub], p.q, p.current, nthreads=8)
    e64 = normwise(f64[sub], ref)
    e1, e2 = normwise(mix, f64), rms_rel(mix, f64)
    say("spectrum at %s: fp64 vs oracle %.3e; mixed vs fp64 normwise %.3e rms %.3e" % (obs, e64, e1, e2))
    assert e64 <= TOL_FP64
    assert e1 <= TOL_MIXED_SPECTRUM and e2 <= TOL_MIXED_SPECTRUM
