"""Batched multi-particle radiation: one device pipeline instead of the reference's
per-particle loop (OSCARSSR.cc:2465-2479).

Particles are sampled on the host by the engine's generator in the reference's
draw order (TParticleBeam::GetNewParticle, TParticleBeam.cc:228-265), so a seeded
run draws the same particles as the reference; trajectories (device RK4), the
flux integrals (persistent work-queue kernel) and the 1/NP-weighted Kahan sum over
particles run on the GPU (include/srs_hip.h, srs_hip_batch_flux).  The power
density and the spectrum have the same route (srs_hip_batch_power_density,
srs_hip_batch_spectrum: the nparticles loops of OSCARSSR.cc:1655-1669 and
:963-982, FP64).
"""
import numpy as np

from . import abi


def sample_particles_device(osr, n, ct_start, ct_stop, npoints, seed, first=0, beam=None):
    """Draw n random particles on the device (srs_hip_sample_beams): counter-based
    (Philox4x32-10), so particle i is the same whatever n and call order; each particle
    from a beam chosen by the beams' weights (TParticleBeamContainer.cc:100-142), then the
    reference's transform (TParticleBeam.cc:228-265).  beam = one beam's name to draw
    from it alone.  Returns (rows, charge, current): every row carries its beam's charge
    and current; the pair returned is the first beam's."""
    names = osr._e.beam_names()
    if not names:
        raise ValueError("No particle beam defined")
    bps = [osr._e.beam_params(beam)] if beam is not None else [osr._e.beam_params(nm) for nm in names]
    rows = abi.sample_beams(bps, [b["weight"] for b in bps], n, seed, first, ct_start, ct_stop, npoints)
    return rows, bps[0]["charge"], bps[0]["current"]


def sample_particles(osr, n, ct_start, ct_stop, npoints, ct_jitter=0.0, points_per_meter=None, rng_seed=0):
    """Draw n random particles from the engine's beam (reference order).

    ct_jitter > 0 gives each particle its own ct range ct_start + (ct_stop - ct_start)(1 + U(-j, j))
    with npoints scaled by points_per_meter (ragged trajectories); U comes from a separate
    NumPy generator so the beam draws are untouched."""
    rows = []
    rng = np.random.default_rng(rng_seed)
    q = cur = None
    for _ in range(n):
        osr._e.set_new_particle()
        st = osr._e.particle_state()
        q, cur = st["charge"], st["current"]
        stop, npts = ct_stop, npoints
        if ct_jitter > 0:
            stop = ct_start + (ct_stop - ct_start) * (1 + rng.uniform(-ct_jitter, ct_jitter))
            npts = int((points_per_meter or (npoints / (ct_stop - ct_start))) * (stop - ct_start))
        rows.append(dict(x0=st["x0"], beta0=st["beta0"], qmg=st["qmg"], t0=st["t0"], ct_start=ct_start, ct_stop=stop,
                         npoints=npts, charge=q, current=cur))
    return rows, q, cur


def flux_device_sampled(osr, n, ct_start, ct_stop, npoints, seed, obs_xyz, energies_ev, precision="double",
                        first=0, beam=None, graph=True, autotune=True):
    """Sum over n particles of flux/n on obs_xyz (N x 3) at each energy -> array (ne, N),
    the particles drawn on the device as sample_particles_device draws them and written
    straight into the pipeline's particle records (srs_hip_batch_flux_beams): no host
    round trip between the sampler and the work queue."""
    names = osr._e.beam_names()
    if not names:
        raise ValueError("No particle beam defined")
    bps = [osr._e.beam_params(beam)] if beam is not None else [osr._e.beam_params(nm) for nm in names]
    prec = {"double": abi.FP64, "fp64": abi.FP64, "mixed": abi.MIXED}[precision]
    flags = (abi.BATCH_GRAPH if graph else 0) | (abi.BATCH_AUTOTUNE if autotune else 0)
    return abi.batch_flux_beams(osr._e.undulators(), bps, [b["weight"] for b in bps], n, seed, obs_xyz, energies_ev,
                                1.0 / n, prec, flags, first, ct_start, ct_stop, npoints)


def flux(osr, rows, charge, current, obs_xyz, energies_ev, precision="double", graph=True, autotune=True):
    """Sum over particles of flux/NP on obs_xyz (N x 3) at each energy -> array (ne, N)."""
    prec = {"double": abi.FP64, "fp64": abi.FP64, "mixed": abi.MIXED}[precision]
    flags = (abi.BATCH_GRAPH if graph else 0) | (abi.BATCH_AUTOTUNE if autotune else 0)
    return abi.batch_flux(osr._e.undulators(), abi.particles(rows), charge, current, obs_xyz, energies_ev,
                          1.0 / len(rows), prec, flags)


def spectrum(osr, rows, charge, current, obs, energies_ev, precision="double"):
    """Sum over particles of spectrum/NP at obs (3,) for each energy -> array (ne,)."""
    prec = {"double": abi.FP64, "fp64": abi.FP64, "mixed": abi.MIXED}[precision]
    return abi.batch_spectrum(osr._e.undulators(), abi.particles(rows), charge, current, obs, energies_ev,
                              1.0 / len(rows), prec)


def power_density(osr, rows, charge, current, obs_xyz, normals_xyz, directional=False, precision="double"):
    """Sum over particles of power density/NP [W/mm^2] on obs_xyz (N x 3) -> array (N,)."""
    prec = {"double": abi.FP64, "fp64": abi.FP64, "mixed": abi.MIXED}[precision]
    return abi.batch_power_density(osr._e.undulators(), abi.particles(rows), charge, current, obs_xyz, normals_xyz,
                                   directional, 1.0 / len(rows), prec)
