"""The counter-based beam sampler's CPU restatement (oracle_philox4x32_10,
oracle_counter_normals, oracle_sample_beam), which the device sampler
(srs_hip_sample_beam) is checked against on the GPU.

Pins: Philox4x32-10 against the known-answer vectors published with the generator
(Random123's kat_vectors: zero, all-ones and pi-digit counters and keys); the
Box-Muller normals by their distribution (moments and a Kolmogorov-Smirnov test on
120000 draws); the particle transform is TParticleBeam::GetNewParticle's, shared with
oracle_random_electron (pinned bit for bit by the beam_c3_seed1234 fixture), and here
by its closed forms: zero spreads give the ideal particle, and the sample moments
match the beam's rms values."""
import math

import numpy as np
import scipy.stats

import cases


def test_philox_known_answers():
    from oracle import pyoracle as O
    assert O.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert O.philox4x32_10([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert O.philox4x32_10([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_counter_normals_are_standard_normal():
    from oracle import pyoracle as O
    z = np.array([O.counter_normals(12345, i) for i in range(20000)]).ravel()
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    assert scipy.stats.kstest(z, "norm").pvalue > 1e-3
    # the six normals of one counter are not correlated with each other
    m = np.array([O.counter_normals(777, i) for i in range(20000)])
    c = np.corrcoef(m.T)
    assert np.max(np.abs(c - np.eye(6))) < 0.03


def _beam(sig_e=0.00267, su=(3.7e-5, 2.5e-6), sp=(2.4e-5, 3.2e-6)):
    me = 9.10938356e-31
    return dict(e0=3.0, sigma_e=sig_e, mass=me, charge=-1.602176462e-19, current=0.5, x0=[0, 0, -1.5], u0=[0, 0, 1],
                hdir=[1, 0, 0], vdir=[0, 1, 0], sigma_at=[0, 0, 0], sigma_u=list(su), sigma_up=list(sp), t0=0.0)


def test_sample_beam_closed_forms():
    from oracle import pyoracle as O
    # no spreads: the ideal particle (TParticleBeam::GetNewParticle("ideal"))
    p = O.sample_beam(_beam(0.0, (0, 0), (0, 0)), 1, 5)
    ideal = O.ideal_electron(3.0, [0, 0, -1.5], [0, 0, 1], 0.5)
    assert list(p.x0) == list(ideal.x0) and list(p.beta0) == list(ideal.beta0) and p.qmg == ideal.qmg
    # moments: position spread sigma_u (1 + d), d = (sigma_at - x0).u0 = 1.5; angles sigma_up;
    # energy from gamma
    b = _beam()
    ps = [O.sample_beam(b, 99, i) for i in range(20000)]
    x = np.array([q.x0[0] for q in ps])
    y = np.array([q.x0[1] for q in ps])
    bx = np.array([q.beta0[0] / q.beta0[2] for q in ps])
    by = np.array([q.beta0[1] / q.beta0[2] for q in ps])
    e = np.array([q.gamma for q in ps]) * 0.51099895e-3
    for v, s in ((x, b["sigma_u"][0] * 2.5), (y, b["sigma_u"][1] * 2.5), (bx, b["sigma_up"][0]),
                 (by, b["sigma_up"][1]), (e - 3.0, b["sigma_e"])):
        assert abs(v.std() / s - 1) < 0.03 and abs(v.mean()) < 0.03 * s
    assert math.isclose(e.mean(), 3.0, rel_tol=1e-4)


def test_beam_index_by_weight():
    """Weighted beam choice (TParticleBeamContainer::GetRandomBeamIndexByWeight): the
    frequencies follow the weights, a zero-weight beam is never chosen, one beam needs no
    draw."""
    from oracle import pyoracle as O
    w = [1.0, 3.0, 0.0, 2.0]
    idx = np.array([O.beam_index(w, 31, i) for i in range(30000)])
    f = np.bincount(idx, minlength=4) / len(idx)
    assert np.all(np.abs(f - np.array(w) / sum(w)) < 0.01) and f[2] == 0
    assert O.beam_index([5.0], 31, 7) == 0
