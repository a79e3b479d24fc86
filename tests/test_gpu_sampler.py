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
