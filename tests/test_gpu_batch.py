"""GPU tests of the batched multi-particle pipeline (device RK4 + persistent
work-queue flux kernel + particle reduction), through the C ABI.

The device RK4 uses the device sine, so trajectories agree with the host
(reference-order) RK4 to rounding, not bit for bit; the FP64 batched flux is
therefore compared with the reference at 1e-10 (measured margin printed), the
mixed form at 1e-5 against the FP64 batched form."""
import numpy as np
import pytest

import cases
from conftest import golden
from gpu_util import normwise, rms_rel, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def abi():
    from srsoriginal_amd import abi as A
    assert A.device_count() >= 1
    return A


def _c3_particles(oracle, name, count, ct=(0.0, 3.0), npoints=None):
    a = cases.args(name)
    b = cases.beam(a)
    oracle.rng_seed(int(a["seed"]), reset_normal=True)
    rows = []
    for _ in range(count):
        p = oracle.random_electron(b["energy"], float(a["sigE"]), b["x0"], b["d0"], cases.vec(a, "hdir"),
                                   cases.vec(a, "beta"), cases.vec(a, "emit"), cases.vec(a, "latref"), b["current"])
        rows.append(dict(x0=list(p.x0), beta0=list(p.beta0), qmg=p.qmg, t0=0.0, ct_start=ct[0], ct_stop=ct[1],
                         npoints=npoints or int(a["npoints"])))
    return rows, p


def test_device_rk4_matches_host_trajectory(abi, oracle):
    say("device RK4 vs host (reference-order) RK4, C0 undulator, 20001 points")
    F, p, t, dt = cases.oracle_setup(oracle, "traj_c0_20001")
    parts = abi.particles([dict(x0=list(p.x0), beta0=list(p.beta0), qmg=p.qmg, t0=0.0, ct_start=0.0, ct_stop=3.0,
                                npoints=20001)])
    tr, offs, dts = abi.trajectories([dict(field=(0, 1, 0), period=(0, 0, 0.05), nperiods=41)], parts)
    assert dts[0] == dt and offs[1] == t.shape[1]
    for k, scale in ((0, 1e-5), (2, 1.0), (3, 1e-3), (5, 1.0)):
        err = np.max(np.abs(tr[k] - t[k])) / scale
        say("  component %d max deviation / scale %.3e" % (k, err))
        assert err < 1e-12
    assert np.array_equal(tr[6:, 0], np.zeros(3))


def test_batch_flux_fp64_matches_reference_multiparticle(abi, oracle):
    name = "multiflux_c3_small"
    say("batched fp64 flux vs the reference's nparticles=4 loop (9x9, 431.2 eV)")
    a = cases.args(name)
    rows, p = _c3_particles(oracle, name, int(a["nparticles"]))
    r = cases.rect(a)
    xyz, nrm, x12 = oracle.rect_surface(r["plane"], r["nx1"], r["nx2"], r["w1"], r["w2"], r["rot"], r["trans"], r["normal"])
    ref = golden(name)["container"][:len(xyz), 3]
    und = [dict(field=u["field"], period=u["period"], nperiods=u["nperiods"]) for u in cases.undulators(a)]
    got = abi.batch_flux(und, abi.particles(rows), p.q, p.current, xyz, [float(a["energy_ev"])], 1.0 / len(rows),
                         abi.FP64)[0]
    err = normwise(got, ref)
    say("  normwise %.3e" % err)
    assert err <= 1e-10


def test_batch_graph_replays_match_eager_and_autotune(abi, oracle):
    name = "multiflux_c3_small"
    say("batched mixed flux: eager vs captured-graph replays vs autotuned; ragged ct ranges")
    a = cases.args(name)
    rows, p = _c3_particles(oracle, name, 24)
    rng = np.random.default_rng(3)
    for rrow in rows:  # ragged trajectories: ct ranges 3 m +- 10 %, ~6667 points per metre
        stop = 3.0 * (1 + rng.uniform(-0.1, 0.1))
        rrow["ct_stop"] = stop
        rrow["npoints"] = int(6667 * stop)
    parts = abi.particles(rows)
    xyz, nrm, x12 = oracle.rect_surface("XY", 23, 19, 0.01, 0.008, (0, 0, 0), (0, 0, 30), 0)
    und = [dict(field=(0, 1, 0), period=(0, 0, 0.05), nperiods=41)]
    es = [430.0, 431.2, 432.4]
    w = 1.0 / len(rows)
    eager = abi.batch_flux(und, parts, p.q, p.current, xyz, es, w, abi.MIXED, 0)
    for rep in range(3):
        g = abi.batch_flux(und, parts, p.q, p.current, xyz, es, w, abi.MIXED, abi.BATCH_GRAPH)
        assert np.array_equal(g, eager), "graph replay %d differs from eager" % rep
        _ = abi.flux(np.zeros((6, 8)) + 1e-9, 1e-12, [[0, 0, 30]], [100.0], p.q, p.current)  # eager work between replays
    tuned = abi.batch_flux(und, parts, p.q, p.current, xyz, es, w, abi.MIXED, abi.BATCH_AUTOTUNE)
    say("  autotuned config (energies/item, tile) = %s" % (abi.batch_config(),))
    assert normwise(tuned, eager) <= 1e-6
    f64 = abi.batch_flux(und, parts, p.q, p.current, xyz, es, w, abi.FP64, 0)
    for k in range(3):
        say("  energy %.1f: mixed vs fp64 normwise %.3e rms %.3e" % (es[k], normwise(eager[k], f64[k]),
                                                                      rms_rel(eager[k], f64[k])))
        assert normwise(eager[k], f64[k]) <= 1e-5 and rms_rel(eager[k], f64[k]) <= 1e-5
    # single-particle consistency: particle 0 alone through the batch equals the single-particle kernels
    one = abi.batch_flux(und, abi.particles(rows[:1]), p.q, p.current, xyz, es[:1], 1.0, abi.FP64, 0)[0]
    tr, offs, dts = abi.trajectories(und, abi.particles(rows[:1]))
    single = abi.flux(tr, dts[0], xyz, es[:1], p.q, p.current, abi.FP64)[0]
    assert normwise(one, single) <= 1e-12


def test_batch_power_density_matches_per_particle_calls():
    """srs_hip_batch_power_density (device RK4 in time segments + pd_batch_kernel) against
    the single-trajectory entry point on each particle's device trajectory, weighted."""
    from srsoriginal_amd import abi
    from srsoriginal_amd.lib import _sr
    from srsoriginal_amd.sr import sr
    say("batched power density, 6 ragged particles (one > 4096 steps: segmented) vs per-particle srs_hip_power_density")
    osr = sr()
    osr.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.05], nperiods=41)
    osr.set_particle_beam(type="electron", name="b", energy_GeV=3, x0=[0, 0, -1.5], d0=[0, 0, 1], current=0.5,
                          sigma_energy_GeV=0.00267, beta=[1.5, 0.8], emittance=[0.9e-9, 8e-12],
                          horizontal_direction=[1, 0, 0], lattice_reference=[0, 0, 0])
    osr.set_seed(99)
    from srsoriginal_amd import batch
    rows, q, cur = batch.sample_particles(osr, 6, -0.2, 3.0, 3001, ct_jitter=0.3, points_per_meter=3000, rng_seed=5)
    rows[2]["npoints"] = 12001  # a long one: the pipeline runs in time segments
    surf = _sr.Surface.rectangle("XY", 9, 7, 0.04, 0.02, [0, 0, 0], [0, 0, 30], 0)
    parts = abi.particles(rows)
    w = 1.0 / len(rows)
    got = batch.power_density(osr, rows, q, cur, surf.xyz, surf.normals)
    ref = np.zeros(len(surf.xyz))
    for r in rows:
        t, off, dts = abi.trajectories(osr._e.undulators(), abi.particles([r]))
        ref += w * abi.power_density(t, dts[0], surf.xyz, surf.normals, False, q, cur, abi.FP64)
    err = normwise(got, ref)
    say("  normwise %.3e" % err)
    assert err <= 1e-12


def test_batch_spectrum_matches_per_particle_calls():
    """srs_hip_batch_spectrum (device RK4 in time segments + spectrum_batch_kernel) against
    the single-trajectory spectrum on each particle's device trajectory, weighted."""
    from srsoriginal_amd import abi, batch
    from srsoriginal_amd.sr import sr
    say("batched spectrum, 5 ragged particles (segmented), 300 energies, vs per-particle srs_hip_spectrum")
    osr = sr()
    osr.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.05], nperiods=41)
    osr.set_particle_beam(type="electron", name="b", energy_GeV=3, x0=[0, 0, -1.5], d0=[0, 0, 1], current=0.5,
                          sigma_energy_GeV=0.00267, beta=[1.5, 0.8], emittance=[0.9e-9, 8e-12],
                          horizontal_direction=[1, 0, 0], lattice_reference=[0, 0, 0])
    osr.set_seed(31)
    rows, q, cur = batch.sample_particles(osr, 5, -0.2, 3.0, 3001, ct_jitter=0.3, points_per_meter=3000, rng_seed=6)
    energies = np.linspace(100.0, 1500.0, 300)
    obs = [0.0005, -0.0002, 30.0]
    got = batch.spectrum(osr, rows, q, cur, obs, energies)
    ref = np.zeros(len(energies))
    w = 1.0 / len(rows)
    for r in rows:
        t, off, dts = abi.trajectories(osr._e.undulators(), abi.particles([r]))
        ref += w * abi.spectrum(t, dts[0], obs, energies, q, cur, abi.FP64)
    err = normwise(got, ref)
    say("  normwise %.3e" % err)
    assert err <= 1e-12


def test_batch_graph_follows_a_changed_field_list(abi, oracle):
    """A captured graph is keyed by the field list's length and kinds: the same shapes
    with an undulator, then a uniform box of the same count (general field evaluation),
    then two undulators, each replay equal to the eager run of its own list."""
    name = "multiflux_c3_small"
    say("batched flux graph: undulator -> box -> two undulators, each vs eager")
    a = cases.args(name)
    rows, p = _c3_particles(oracle, name, 4)
    parts = abi.particles(rows)
    xyz, nrm, x12 = oracle.rect_surface("XY", 9, 9, 0.01, 0.008, (0, 0, 0), (0, 0, 30), 0)
    lists = [[dict(field=(0, 1, 0), period=(0, 0, 0.05), nperiods=41)],
             [dict(kind=abi.FIELD_UNIFORM_BOX, field=(0, 0.05, 0), width=(0, 0, 2.0))],
             [dict(field=(0, 1, 0), period=(0, 0, 0.05), nperiods=41),
              dict(field=(0.1, 0, 0), period=(0, 0, 0.05), nperiods=41, phase=1.0)]]
    for und in lists:
        eager = abi.batch_flux(und, parts, p.q, p.current, xyz, [431.2], 0.25, abi.FP64, 0)
        graph = abi.batch_flux(und, parts, p.q, p.current, xyz, [431.2], 0.25, abi.FP64, abi.BATCH_GRAPH)
        assert np.array_equal(graph, eager)


def test_batch_graph_follows_current_charge_and_particle_count(abi, oracle):
    """The captured graph holds the flux prefactor c2 (from charge and current) and the
    particle buffer's address by value: replays with another current, another charge, or
    enough particles to move that buffer must each equal the eager run of their own
    inputs, never an earlier capture."""
    name = "multiflux_c3_small"
    say("batched flux graph: current 0.5 -> 0.25 A, charge x2, buffers regrown by a 700-particle call, each vs eager")
    a = cases.args(name)
    rows, p = _c3_particles(oracle, name, 4, npoints=2001)
    xyz, nrm, x12 = oracle.rect_surface("XY", 9, 9, 0.01, 0.008, (0, 0, 0), (0, 0, 30), 0)
    und = [dict(field=(0, 1, 0), period=(0, 0, 0.05), nperiods=41)]
    pp = abi.particles(rows)
    w = 1.0 / len(rows)
    prev = None
    for q, cur, regrow in ((p.q, 0.5, False), (p.q, 0.25, False), (2 * p.q, 0.25, False), (2 * p.q, 0.25, True)):
        if regrow:  # an eager call with many particles reallocates the batch buffers under the cached graph
            many = abi.particles((rows * 175)[:700])
            abi.batch_flux(und, many, q, cur, xyz, [431.2], 1.0 / 700, abi.FP64, 0)
        graph = abi.batch_flux(und, pp, q, cur, xyz, [431.2], w, abi.FP64, abi.BATCH_GRAPH)
        eager = abi.batch_flux(und, pp, q, cur, xyz, [431.2], w, abi.FP64, 0)
        say("  q=%+.3e I=%.2f regrown=%d: max %.6e" % (q, cur, regrow, float(np.max(eager))))
        assert np.array_equal(graph, eager)
        if prev is not None and not regrow:
            assert not np.array_equal(eager, prev)  # c2 changed, so must the flux
        prev = eager


def test_batch_power_density_mixed_within_1e5():
    """Mixed batched power density (pd_batch_kernel's FP32 terms) against the FP64 batch,
    16 ragged particles, one segmented, directional and not, normal and tilted screen."""
    from srsoriginal_amd import abi, batch
    from srsoriginal_amd.lib import _sr
    from srsoriginal_amd.sr import sr
    say("batched power density mixed vs fp64: 16 particles, 101x51 over 40 mm and a tilted screen")
    osr = sr()
    osr.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.05], nperiods=41)
    osr.set_particle_beam(type="electron", name="b", energy_GeV=3, x0=[0, 0, -1.5], d0=[0, 0, 1], current=0.5,
                          sigma_energy_GeV=0.00267, beta=[1.5, 0.8], emittance=[0.9e-9, 8e-12],
                          horizontal_direction=[1, 0, 0], lattice_reference=[0, 0, 0])
    osr.set_seed(17)
    rows, q, cur = batch.sample_particles(osr, 16, -0.2, 3.0, 3001, ct_jitter=0.3, points_per_meter=3000, rng_seed=8)
    rows[3]["npoints"] = 12001  # segmented pipeline
    for rot, directional in (((0, 0, 0), False), ((0.4, 0.3, 0), True)):
        surf = _sr.Surface.rectangle("XY", 101, 51, 0.04, 0.02, list(rot), [0, 0, 30], 0)
        f64 = batch.power_density(osr, rows, q, cur, surf.xyz, surf.normals, directional)
        mix = batch.power_density(osr, rows, q, cur, surf.xyz, surf.normals, directional, precision="mixed")
        err = normwise(mix, f64)
        say("  rotations %s directional=%d: normwise %.3e" % (rot, directional, err))
        assert err <= 1e-5


def test_batch_spectrum_mixed_within_1e5():
    """Mixed batched spectrum (spectrum_batch_kernel's FP32 terms, FP64 phase) against the
    FP64 batch: 12 ragged particles, one segmented, 2000 energies on and off axis."""
    from srsoriginal_amd import batch
    from srsoriginal_amd.sr import sr
    say("batched spectrum mixed vs fp64: 12 particles, 2000 energies")
    osr = sr()
    osr.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.05], nperiods=41)
    osr.set_particle_beam(type="electron", name="b", energy_GeV=3, x0=[0, 0, -1.5], d0=[0, 0, 1], current=0.5,
                          sigma_energy_GeV=0.00267, beta=[1.5, 0.8], emittance=[0.9e-9, 8e-12],
                          horizontal_direction=[1, 0, 0], lattice_reference=[0, 0, 0])
    osr.set_seed(23)
    rows, q, cur = batch.sample_particles(osr, 12, -0.2, 3.0, 3001, ct_jitter=0.3, points_per_meter=3000, rng_seed=9)
    rows[5]["npoints"] = 12001
    energies = 100.0 + (3000.0 - 100.0) / 1999 * (np.arange(2000) + 1)
    for obs in ([0.0, 0.0, 30.0], [0.0012, -0.0004, 30.0]):
        f64 = batch.spectrum(osr, rows, q, cur, obs, energies)
        mix = batch.spectrum(osr, rows, q, cur, obs, energies, precision="mixed")
        err = normwise(mix, f64)
        say("  obs %s: normwise %.3e" % (obs, err))
        assert err <= 1e-5


def test_batch_routes_use_each_particles_charge_and_current(abi, oracle):
    """Particles with their own charge and current (two beams of the reference's
    per-particle loop, OSCARSSR.cc:2465-2479): flux, power density and spectrum through the
    batch entry points against each particle's single-trajectory result with its own
    charge and current, weighted."""
    name = "multiflux_c3_small"
    say("batched flux / power density / spectrum, 4 particles with currents 0.5, 0.2, 0.5, 1.1 A and charge x2 on one")
    rows, p = _c3_particles(oracle, name, 4, npoints=3001)
    setup = [(p.q, 0.5), (p.q, 0.2), (0.0, 0.0), (2 * p.q, 1.1)]  # (0, 0): the call's charge and current
    for r, (q, cur) in zip(rows, setup):
        r["charge"], r["current"] = q, cur
    parts = abi.particles(rows)
    und = [dict(field=(0, 1, 0), period=(0, 0, 0.05), nperiods=41)]
    xyz, nrm, x12 = oracle.rect_surface("XY", 9, 7, 0.01, 0.008, (0, 0, 0), (0, 0, 30), 0)
    w = 0.25
    eff = [(q, cur) if q != 0 else (p.q, 0.5) for q, cur in setup]
    refs = {"flux": 0.0, "pd": 0.0, "spec": 0.0}
    es = np.linspace(100.0, 600.0, 37)
    for r, (q, cur) in zip(rows, eff):
        t, off, dts = abi.trajectories(und, abi.particles([r]))
        refs["flux"] = refs["flux"] + w * abi.flux(t[:6], dts[0], xyz, [431.2], q, cur, abi.FP64)[0]
        refs["pd"] = refs["pd"] + w * abi.power_density(t, dts[0], xyz, nrm, False, q, cur, abi.FP64)
        refs["spec"] = refs["spec"] + w * abi.spectrum(t[:6], dts[0], (0.0005, 0, 30), es, q, cur, abi.FP64)
    got = {"flux": abi.batch_flux(und, parts, p.q, 0.5, xyz, [431.2], w, abi.FP64, 0)[0],
           "pd": abi.batch_power_density(und, parts, p.q, 0.5, xyz, nrm, False, w, abi.FP64),
           "spec": abi.batch_spectrum(und, parts, p.q, 0.5, (0.0005, 0, 30), es, w, abi.FP64)}
    for k in refs:
        err = normwise(got[k], refs[k])
        say("  %s normwise %.3e" % (k, err))
        assert err <= 1e-12
    # the sampler's particles carry their beam's charge and current
    bp = dict(e0=3.0, sigma_e=0.001, mass=9.10938356e-31, charge=p.q, current=0.37, x0=[0, 0, -1.5], u0=[0, 0, 1],
              hdir=[1, 0, 0], vdir=[0, 1, 0], sigma_at=[0, 0, 0], sigma_u=[1e-5, 1e-6], sigma_up=[1e-5, 1e-6])
    s = abi.sample_beam(bp, 3, 5, ct_start=0.0, ct_stop=3.0, npoints=301)
    assert all(r["charge"] == p.q and r["current"] == 0.37 for r in s)
    bp["current"] = 0.0  # TParticleA: current 0 means I = q
    assert all(r["current"] == p.q for r in abi.sample_beam(bp, 2, 5, ct_start=0.0, ct_stop=3.0, npoints=301))


def test_batch_mixed_segments_across_the_paraxial_bound(abi, oracle):
    """20001-point trajectories (integrated in time segments) on a 42 mm screen at 30 m:
    the corner tiles pass the paraxial eps bound in the segments far from the screen and
    fail it in the near ones, so one point's field sums come from both forms and must
    share one phase.  Mixed vs FP64 batch, one and three energies, the one-energy shapes
    with one and two points per lane."""
    name = "multiflux_c3_small"
    say("batched flux across the paraxial bound: mixed vs fp64 (<=1e-5)")
    rows, p = _c3_particles(oracle, name, 3, npoints=20001)
    parts = abi.particles(rows)
    xyz, nrm, x12 = oracle.rect_surface("XY", 41, 41, 0.042, 0.042, (0, 0, 0), (0, 0, 30), 0)
    und = [dict(field=(0, 1, 0), period=(0, 0, 0.05), nperiods=41)]
    w = 1.0 / len(rows)
    try:
        for es, shapes in (([143.52176088044024], [(1, 128, 1), (1, 256, 2)]), ([430.0, 431.2, 432.4], [(3, 128, 1)])):
            abi.batch_set_config(shapes[0][0], 128, 1)
            f64 = np.atleast_2d(abi.batch_flux(und, parts, p.q, p.current, xyz, es, w, abi.FP64, 0))
            for sh in shapes:
                abi.batch_set_config(*sh)
                m = np.atleast_2d(abi.batch_flux(und, parts, p.q, p.current, xyz, es, w, abi.MIXED, 0))
                errs = [normwise(m[k], f64[k]) for k in range(len(es))]
                say("  %s shape %s: mixed vs fp64 normwise %s" % (es, sh, " ".join("%.2e" % e for e in errs)))
                assert max(errs) <= 1e-5 and max(rms_rel(m[k], f64[k]) for k in range(len(es))) <= 1e-5
    finally:
        abi.batch_set_config(0)


def test_batch_every_tile_shape_matches_fp64(abi, oracle):
    """Each tile shape of the persistent kernel, fixed with srs_hip_batch_set_config (two
    observation points per lane included, on an odd point count so the last 512-point
    tile is ragged), against the FP64 batch; the two-point shape also through replayed
    graphs, bit for bit against its eager run."""
    name = "multiflux_c3_small"
    say("batched flux, every tile shape: mixed vs fp64 (<=1e-5), fp64 shapes vs each other (<=1e-12)")
    rows, p = _c3_particles(oracle, name, 6)
    parts = abi.particles(rows)
    xyz, nrm, x12 = oracle.rect_surface("XY", 27, 23, 0.01, 0.008, (0, 0, 0), (0, 0, 30), 0)
    und = [dict(field=(0, 1, 0), period=(0, 0, 0.05), nperiods=41)]
    es = [430.0, 431.2, 432.4]
    w = 1.0 / len(rows)
    shapes = [(1, 64, 1), (1, 128, 1), (1, 256, 1), (3, 64, 1), (3, 128, 1), (3, 256, 1), (1, 128, 2), (1, 256, 2)]
    try:
        abi.batch_set_config(1, 128, 1)
        f64 = abi.batch_flux(und, parts, p.q, p.current, xyz, es, w, abi.FP64, 0)
        for sh in shapes:
            abi.batch_set_config(*sh)
            if sh[2] == 1:
                g = abi.batch_flux(und, parts, p.q, p.current, xyz, es, w, abi.FP64, 0)
                assert abi.batch_config() == sh
                assert normwise(g, f64) <= 1e-12, sh
            m = abi.batch_flux(und, parts, p.q, p.current, xyz, es, w, abi.MIXED, 0)
            assert abi.batch_config() == sh
            errs = [normwise(m[k], f64[k]) for k in range(3)]
            say("  shape %s: mixed vs fp64 normwise %s" % (sh, " ".join("%.2e" % e for e in errs)))
            assert max(errs) <= 1e-5 and max(rms_rel(m[k], f64[k]) for k in range(3)) <= 1e-5
            if sh[2] == 2:
                for rep in range(3):
                    gr = abi.batch_flux(und, parts, p.q, p.current, xyz, es, w, abi.MIXED, abi.BATCH_GRAPH)
                    assert np.array_equal(gr, m), "graph replay %d of %s differs from eager" % (rep, sh)
                    _ = abi.flux(np.zeros((6, 8)) + 1e-9, 1e-12, [[0, 0, 30]], [100.0], p.q, p.current)
        abi.batch_set_config(1, 256, 2)
        with pytest.raises(abi.SrsError, match="mixed precision"):
            abi.batch_flux(und, parts, p.q, p.current, xyz, es, w, abi.FP64, 0)
        with pytest.raises(abi.SrsError, match="no batch kernel"):
            abi.batch_set_config(3, 256, 2)
    finally:
        abi.batch_set_config(0)
