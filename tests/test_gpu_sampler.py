"""GPU tests of the device beam sampler (srs_hip_sample_beam): every particle against
the CPU restatement (oracle_sample_beam: Philox4x32-10, Box-Muller, the reference's
particle transform), independence of the draws from the batch size and offset, and the
sample moments of a large draw against the beam's rms values."""
import numpy as np
import pytest

from gpu_util import say

pytestmark = pytest.mark.gpu


def _beam_engine():
    from srsoriginal_amd.sr import sr
    osr = sr()
    osr.set_particle_beam(type="electron", name="b", energy_GeV=3, x0=[0, 0, -1.5], d0=[0, 0, 1], current=0.5,
                          sigma_energy_GeV=0.00267, beta=[1.5, 0.8], emittance=[0.9e-9, 8e-12],
                          horizontal_direction=[1, 0, 0], lattice_reference=[0, 0, 0])
    return osr


def test_device_sampler_matches_oracle_and_is_counter_based():
    from oracle import pyoracle as O
    from srsoriginal_amd import abi, batch
    say("device beam sampler: 512 particles vs oracle_sample_beam; offsets; moments of 200000")
    osr = _beam_engine()
    bp = osr._e.beam_params("b")
    rows, q, cur = batch.sample_particles_device(osr, 512, 0.0, 3.0, 20001, seed=2024)
    worst = 0.0
    for i, r in enumerate(rows):
        o = O.sample_beam(bp, 2024, i)
        for a, b in ((r["x0"], o.x0), (r["beta0"], o.beta0)):
            a, b = np.array(a), np.array(list(b))
            worst = max(worst, float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-12))))
        worst = max(worst, abs(r["qmg"] - o.qmg) / abs(o.qmg))
        assert r["ct_start"] == 0.0 and r["ct_stop"] == 3.0 and r["npoints"] == 20001
    say("  max relative deviation from the oracle %.3e" % worst)
    assert worst <= 1e-14
    assert q == bp["charge"] and cur == bp["current"]
    a = abi.sample_beam(bp, 300, 7)
    b = abi.sample_beam(bp, 100, 7, first=50)
    assert a[50:150] == b
    big = abi.sample_beam(bp, 200000, 11)
    x = np.array([r["x0"][0] for r in big])
    xp = np.array([r["beta0"][0] / r["beta0"][2] for r in big])
    sx, sxp = bp["sigma_u"][0] * 2.5, bp["sigma_up"][0]  # 1 + (sigma_at - x0).u0 = 2.5
    say("  std x / expected %.4f, std x' / expected %.4f" % (x.std() / sx, xp.std() / sxp))
    assert abs(x.std() / sx - 1) < 0.01 and abs(xp.std() / sxp - 1) < 0.01


def test_device_sampler_weighted_beams_match_oracle():
    """srs_hip_sample_beams with three beams (distinct starts and energies) and weights
    1 : 3 : 2: each particle's beam and values against the CPU restatement; the engine's
    beams and weights through batch.sample_particles_device."""
    from oracle import pyoracle as O
    from srsoriginal_amd import abi, batch
    from srsoriginal_amd.sr import sr
    say("device sampler, three weighted beams vs oracle_beam_index + oracle_sample_beam")
    osr = sr()
    for k, (e, w) in enumerate(((3.0, 1.0), (2.5, 3.0), (3.2, 2.0))):
        osr.add_particle_beam(type="electron", name="b%d" % k, energy_GeV=e, x0=[0.001 * k, 0, -1.5], d0=[0, 0, 1],
                              current=0.5, weight=w, sigma_energy_GeV=0.001, beta=[1.5, 0.8],
                              emittance=[0.9e-9, 8e-12], horizontal_direction=[1, 0, 0], lattice_reference=[0, 0, 0])
    bps = [osr._e.beam_params("b%d" % k) for k in range(3)]
    ws = [b["weight"] for b in bps]
    assert ws == [1.0, 3.0, 2.0]
    rows, q, cur = batch.sample_particles_device(osr, 3000, 0.0, 3.0, 2001, seed=5)
    counts = [0, 0, 0]
    worst = 0.0
    for i, r in enumerate(rows):
        k = O.beam_index(ws, 5, i)
        counts[k] += 1
        o = O.sample_beam(bps[k], 5, i)
        for a, b in ((r["x0"], o.x0), (r["beta0"], o.beta0)):
            a, b = np.array(a), np.array(list(b))
            worst = max(worst, float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-12))))
        worst = max(worst, abs(r["qmg"] - o.qmg) / abs(o.qmg))
    say("  beams drawn %s, max relative deviation %.3e" % (counts, worst))
    assert worst <= 1e-14
    assert all(abs(c / 3000 - w / 6.0) < 0.03 for c, w in zip(counts, ws))
    with pytest.raises(abi.SrsError, match="weights"):
        abi.sample_beams(bps, [0.0, 0.0, 0.0], 10, 1)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_device_sampled_batch_flux_equals_host_rows(prec):
    """srs_hip_batch_flux_beams draws the particles into the pipeline's records on the
    device; the same draws fetched to the host (srs_hip_sample_beams) and passed through
    srs_hip_batch_flux must give the same flux bit for bit (one beam, and two weighted
    beams whose trajectories differ in length).  The pipeline itself is pinned against the
    reference's per-particle route elsewhere."""
    from srsoriginal_amd import abi, batch
    from srsoriginal_amd.sr import sr
    P = abi.FP64 if prec == "fp64" else abi.MIXED
    say("device-sampled batch flux vs host rows, %s" % prec)
    osr = sr()
    osr.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.050], nperiods=41)
    for k, (e, w) in enumerate(((3.0, 1.0), (2.9, 2.0))):
        osr.add_particle_beam(type="electron", name="b%d" % k, energy_GeV=e, x0=[0, 0, -1.5], d0=[0, 0, 1],
                              current=0.5, weight=w, sigma_energy_GeV=0.00267, beta=[1.5, 0.8],
                              emittance=[0.9e-9, 8e-12], horizontal_direction=[1, 0, 0], lattice_reference=[0, 0, 0])
    xs = np.linspace(-0.004, 0.004, 21)
    obs = np.array([[x, y, 30.0] for y in xs for x in xs])
    es = [143.5, 431.0]
    bps = [osr._e.beam_params("b%d" % k) for k in range(2)]
    for beams, grids in (([bps[0]], [(0.0, 3.0, 4001)]), (bps, [(0.0, 3.0, 4001), (0.0, 2.5, 3001)])):
        for b, (c0, c1, n) in zip(beams, grids):
            b.update(ct_start=c0, ct_stop=c1, npoints=n)
        np_ = 96
        dev = abi.batch_flux_beams(osr._e.undulators(), beams, [b["weight"] for b in beams], np_, 77, obs, es,
                                   1.0 / np_, P, abi.BATCH_GRAPH, first=5)
        rows = []
        for k, b in enumerate(beams):
            # the host route gets the same draws; each row carries its own beam's grid
            r = abi.sample_beams(beams, [bb["weight"] for bb in beams], np_, 77, 5,
                                 b["ct_start"], b["ct_stop"], b["npoints"])
            rows.append(r)
        # every row takes the grid of the beam it was drawn from: match rows by their energy
        # (beta0) against the one-beam-grid draws, which only differ in the grid fields
        host_rows = []
        for i in range(np_):
            cand = [rows[k][i] for k in range(len(beams))]
            assert all(c["x0"] == cand[0]["x0"] and c["beta0"] == cand[0]["beta0"] for c in cand)
            kb = 0
            if len(beams) > 1:
                # the beam a particle came from: its energy is within 10 sigma of that beam's
                g = 1.0 / np.sqrt(1.0 - np.sum(np.square(cand[0]["beta0"])))
                egev = g * 0.51099895e-3
                kb = int(np.argmin([abs(egev - b["e0"]) for b in beams]))
            host_rows.append(cand[kb])
        q, cur = beams[0]["charge"], beams[0]["current"]
        host = abi.batch_flux(osr._e.undulators(), abi.particles(host_rows), q, cur, obs, es, 1.0 / np_, P,
                              abi.BATCH_GRAPH)
        say("  %d beam(s): max |device - host| = %.3e (max flux %.4e)" %
            (len(beams), float(np.max(np.abs(dev - host))), float(np.max(host))))
        assert np.array_equal(dev, host)
    # the entry point rejects a beam without charge
    bad = dict(bps[0], charge=0.0)
    with pytest.raises(abi.SrsError, match="charge"):
        abi.batch_flux_beams(osr._e.undulators(), [bad], [1.0], 4, 1, obs[:4], es, 0.25, P)
