"""Benchmark of the radiation hot path on one MI355X.

Workload (the reference's headline configuration on one GPU): the flux map of
an ideal planar undulator -- 3 GeV, 0.5 A electron, By = 1 T sin(2 pi z/50 mm),
41 periods, RK4 trajectory of 20001 points over ct = 0..3 m -- on a 201 x 201
rectangular grid, 20 mm x 20 mm at z = 30 m, at the first-harmonic energy
E* = 143.5218 eV.  One step = one evaluation of the whole map (40401 points x
20001 trajectory steps = 8.08e8 terms) with inputs resident in HBM.  The mixed
form (FP64 phase, FP32 trig/amplitudes/sums) is the headline value; the plain
FP64 form is timed in the same run, and the mixed map is compared against it.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line.  The roofline entry is for the dominant kernel (the mixed
flux kernel), timed with HIP events on its launch stream; the CPU baseline is
the FP64 CPU path (the oracle's restatement of the reference, one std::thread
per output point as the reference does) on a bounded sample of the same map.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

ENERGY_EV = 143.52176088044024   # first-harmonic argmax of the reference's C0 spectrum grid
F_ALG_FLUX = 53                  # algorithmic FLOP per flux term
F_ALG_SPEC = 30                  # per spectrum term
F_ALG_PD = 57                    # per power-density term
F_ALG_FLUX1 = 77                 # per CalculateFlux1 (Lienard-Wiechert) term
PEAK_FP32_VALU = 157.3e12        # MI355X FP32 vector peak (published)
PEAK_FP64_VALU = 78.6e12         # MI355X FP64 vector peak (published)
# Issue model of the dominant loop (flux_mixed2_kernel<256, 256>, paraxial terms), from the
# measured cycles per wave64 instruction on gfx950 (a one-off rocprofv3 measurement per class):
# FP32 with VGPR operands 2, FP64 (add, fma, fract, cvt) 4, sin/cos 8.  Per term: 8 FP64
# (phase 6 with fract, two conversions) + 9 FP32 (8 accumulations, 1 amplitude) + sin + cos
# = 66 cycles, plus ~3 for the 32-term sub-tile epilogue.
ISSUE_CYCLES_PER_WAVE_TERM = 69
SIMDS, CLOCK_HZ = 1024, 2.4e9


def _frac(flop_per_term, terms, fp64_ms, mixed_ms):
    """Algorithmic rate of a kernel pair against the published vector peaks (FP64 form
    against 78.6 TF/s, mixed form against 157.3 TF/s FP32)."""
    a64, amx = flop_per_term * terms / (fp64_ms * 1e-3), flop_per_term * terms / (mixed_ms * 1e-3)
    return {"flop_per_term": flop_per_term, "unit": "TFLOP/s",
            "fp64": {"achieved": a64 / 1e12, "peak": PEAK_FP64_VALU / 1e12, "frac": a64 / PEAK_FP64_VALU},
            "mixed": {"achieved": amx / 1e12, "peak": PEAK_FP32_VALU / 1e12, "frac": amx / PEAK_FP32_VALU}}
def _pipeline_frac(flop_per_term, terms, fp64_ms, mixed_ms):
    """_frac over a batched pipeline's whole time (device RK4, term kernel and reductions):
    a lower bound on its term kernel's own fraction."""
    r = _frac(flop_per_term, terms, fp64_ms, mixed_ms)
    r["over"] = "whole pipeline time (RK4, term kernel, reductions)"
    return r


# HBM bytes per launch of flux_mixed2_kernel<256, 256> on this workload, from rocprofv3
# FETCH_SIZE (x2.0) and WRITE_SIZE (x1.0) runs of `bench.py --profile`, the factors being
# what the same counters read for a 256 MiB device copy (profiles/flux_c1_201x201_pmc.md).
TRAFFIC_MIXED_BYTES = 61667.3 * 1024


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def setup_problem():
    from srsoriginal_amd.sr import sr
    osr = sr()
    osr.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.050], nperiods=41)
    osr.set_particle_beam(type="electron", name="beam", energy_GeV=3, x0=[0, 0, -1.5], d0=[0, 0, 1], current=0.5)
    osr.set_ctstartstop(0, 3)
    osr.set_npoints_trajectory(20001)
    osr.calculate_trajectory()
    traj, dt = osr.trajectory_array()
    from srsoriginal_amd.lib import _sr
    surf = _sr.Surface.rectangle("XY", 201, 201, 0.02, 0.02, [0, 0, 0], [0, 0, 30], 0)
    q, cur = osr._e.particle_charge_current()
    return traj, dt, surf.xyz, q, cur


def batched_c3(nparticles=4096, reps=1):
    """Config 3 at its stated size on one GPU: 4096 Gaussian-beam particles, ragged 18k-22k
    point trajectories, 101x101 points over 10 mm at 30 m, 3 energies around the 3rd
    harmonic; FP64 and mixed, each the whole captured pipeline."""
    from srsoriginal_amd import abi, batch
    from srsoriginal_amd.lib import _sr
    from srsoriginal_amd.sr import sr
    osr = sr()
    osr.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.050], nperiods=41)
    osr.set_particle_beam(type="electron", name="beam", energy_GeV=3, x0=[0, 0, -1.5], d0=[0, 0, 1], current=0.5,
                          sigma_energy_GeV=8.9e-4 * 3, beta=[1.5, 0.8], emittance=[0.9e-9, 8e-12],
                          horizontal_direction=[1, 0, 0], lattice_reference=[0, 0, 0])
    osr.set_seed(4096)
    t0 = time.perf_counter()
    rows, q, cur = batch.sample_particles(osr, nparticles, 0.0, 3.0, 20001, ct_jitter=0.1, points_per_meter=6667,
                                          rng_seed=4096)
    host_sampling_s = time.perf_counter() - t0
    xyz = _sr.Surface.rectangle("XY", 101, 101, 0.01, 0.01, [0, 0, 0], [0, 0, 30], 0).xyz
    es = [430.0, 431.2, 432.4]
    npts = abi.trajectory_points(abi.particles(rows))
    lens = [r["npoints"] for r in rows]
    terms = float(npts) * len(xyz) * len(es)
    res = {}
    for prec in ("double", "mixed"):
        out = batch.flux(osr, rows, q, cur, xyz, es, prec)  # autotune + graph capture
        abi.timing_enable(True)
        t0 = time.perf_counter()
        for _ in range(reps):
            out = batch.flux(osr, rows, q, cur, xyz, es, prec)
        wall = (time.perf_counter() - t0) / reps
        kms, nl = abi.timing_read()
        abi.timing_enable(False)
        res[prec] = dict(out=out, wall_s=wall, graph_ms=kms / max(nl, 1), cfg=abi.batch_config())
        log("batched C3 %s: %.3f s/eval (graph %.1f ms), config %s" % (prec, wall, res[prec]["graph_ms"],
                                                                        res[prec]["cfg"]))
    f64, mix = res["double"]["out"], res["mixed"]["out"]
    # the same beam drawn on the device straight into the pipeline (uniform 20001-point
    # trajectories: the device sampler gives every particle its beam's ct range)
    dev = batch.flux_device_sampled(osr, nparticles, 0.0, 3.0, 20001, 4096, xyz, es, "mixed")  # warm-up + capture
    abi.timing_enable(True)
    t0 = time.perf_counter()
    dev = batch.flux_device_sampled(osr, nparticles, 0.0, 3.0, 20001, 4096, xyz, es, "mixed")
    dev_wall = time.perf_counter() - t0
    kms, nl = abi.timing_read()
    abi.timing_enable(False)
    log("batched C3 device-sampled mixed: %.3f s/eval (graph %.1f ms); host sampling of the ragged set %.3f s"
        % (dev_wall, kms / max(nl, 1), host_sampling_s))
    return {"workload": "%d particles x %d traj points (ragged %d-%d) x 101x101 x 3 energies" %
                        (nparticles, npts, min(lens), max(lens)),
            "terms": terms,
            "fp64_ms": res["double"]["graph_ms"], "mixed_ms": res["mixed"]["graph_ms"],
            "fp64_terms_per_s": terms / (res["double"]["graph_ms"] * 1e-3),
            "mixed_terms_per_s": terms / (res["mixed"]["graph_ms"] * 1e-3),
            "speedup": res["double"]["graph_ms"] / res["mixed"]["graph_ms"],
            "roofline": _pipeline_frac(F_ALG_FLUX, terms, res["double"]["graph_ms"], res["mixed"]["graph_ms"]),
            "mixed_vs_fp64_max_rel_err": float(np.max(np.abs(mix - f64)) / np.max(np.abs(f64))),
            "mixed_vs_fp64_rms_rel_err": float(np.sqrt(np.mean((mix - f64) ** 2)) / np.sqrt(np.mean(f64 ** 2))),
            "tile_config_fp64": res["double"]["cfg"], "tile_config_mixed": res["mixed"]["cfg"],
            "host_sampling_s": host_sampling_s,
            "device_sampled_mixed": {"particles": nparticles, "traj_points": 20001, "wall_s": dev_wall,
                                     "graph_ms": kms / max(nl, 1), "peak_flux": float(np.max(dev)),
                                     "path": "srs_hip_batch_flux_beams: device sampler -> particle records -> "
                                             "RK4 + flux + reduction, no host round trip"},
            "timed": "whole hipGraph (device RK4 + persistent flux + particle reduction), HIP events; "
                     "%d timed evaluation(s) per form after one warm-up" % reps}


def flux_general_c1(traj, dt, q, cur, reps=5):
    """C1's map on a tilted screen (201x201 over 20 mm at 30 m, rotated 0.3 rad about y): the
    points do not lie on one plane z = Z, so the mixed kernel takes its general terms (the
    reference's D and tau in FP64, FP32 amplitudes) instead of the paraxial ones."""
    import torch
    from srsoriginal_amd import abi
    from srsoriginal_amd.lib import _sr
    xyz = _sr.Surface.rectangle("XY", 201, 201, 0.02, 0.02, [0, 0.3, 0], [0, 0, 30], 0).xyz
    n, ns = traj.shape[1], len(xyz)
    dev = torch.device("cuda:0")
    d_traj = torch.from_numpy(np.ascontiguousarray(traj)).to(dev)
    d_obs = torch.from_numpy(np.ascontiguousarray(xyz.T)).to(dev)
    ptrs = [d_traj[k].data_ptr() for k in range(9)]
    stream = torch.cuda.current_stream()
    res = {}
    for prec in (abi.FP64, abi.MIXED):
        out = torch.zeros(ns, dtype=torch.float64, device=dev)
        for _ in range(2):
            abi.flux_device(ptrs, n, dt, d_obs.data_ptr(), ns, [ENERGY_EV], q, cur, prec, out.data_ptr(),
                            stream.cuda_stream)
        torch.cuda.synchronize()
        abi.timing_enable(True)
        for _ in range(reps):
            abi.flux_device(ptrs, n, dt, d_obs.data_ptr(), ns, [ENERGY_EV], q, cur, prec, out.data_ptr(),
                            stream.cuda_stream)
        torch.cuda.synchronize()
        kms, nl = abi.timing_read()
        abi.timing_enable(False)
        res[prec] = dict(ms=kms / max(nl, 1), out=out.cpu().numpy())
    terms = float(n) * ns
    f64, mix = res[abi.FP64]["out"], res[abi.MIXED]["out"]
    log("general-form C1 (tilted screen): fp64 %.3f ms, mixed %.3f ms" % (res[abi.FP64]["ms"], res[abi.MIXED]["ms"]))
    return {"mixed_general": {
        "workload": "flux map 201x201, 20 mm at 30 m tilted 0.3 rad about y, E=143.52 eV, NT=20001 (general mixed terms)",
        "terms": terms, "fp64_kernel_ms": res[abi.FP64]["ms"], "mixed_kernel_ms": res[abi.MIXED]["ms"],
        "fp64_terms_per_s": terms / (res[abi.FP64]["ms"] * 1e-3),
        "mixed_terms_per_s": terms / (res[abi.MIXED]["ms"] * 1e-3),
        "speedup": res[abi.FP64]["ms"] / res[abi.MIXED]["ms"],
        "mixed_vs_fp64_max_rel_err": float(np.max(np.abs(mix - f64)) / np.max(np.abs(f64))),
        "mixed_vs_fp64_rms_rel_err": float(np.sqrt(np.mean((mix - f64) ** 2)) / np.sqrt(np.mean(f64 ** 2))),
        "roofline": _frac(F_ALG_FLUX, terms, res[abi.FP64]["ms"], res[abi.MIXED]["ms"])}}


def spectrum_c0_setup():
    """Config 0: the 41 x 50 mm, 1 T undulator, 3 GeV 0.5 A, NT=20001, and the reference's
    2000-point spectrum grid 100-3000 eV (TSpectrumContainer.cc:82)."""
    from srsoriginal_amd.sr import sr
    osr = sr()
    osr.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.050], nperiods=41)
    osr.set_particle_beam(type="electron", name="beam", energy_GeV=3, x0=[0, 0, -1.5], d0=[0, 0, 1], current=0.5)
    osr.set_ctstartstop(0, 3)
    osr.set_npoints_trajectory(20001)
    osr.calculate_trajectory()
    traj, dt = osr.trajectory_array()
    q, cur = osr._e.particle_charge_current()
    energies = 100.0 + (3000.0 - 100.0) / 1999 * (np.arange(2000) + 1)
    return traj, dt, q, cur, energies


def spectrum_and_power_density(reps=5):
    """Config 0 (FP64 spectrum, 2000 energies at (0,0,30 m)) and config 2 (EPU power
    density on 301x301 over 60 mm at 30 m, NT=40001; mixed, checked against FP64 and
    by the integrated power), device entry points, inputs resident in HBM."""
    import torch
    from srsoriginal_amd import abi
    from srsoriginal_amd.lib import _sr
    from srsoriginal_amd.sr import sr
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    res = {}
    # config 0
    traj, dt, q, cur, energies = spectrum_c0_setup()
    d_traj = torch.from_numpy(np.ascontiguousarray(traj)).to(dev)
    ptrs = [d_traj[k].data_ptr() for k in range(9)]
    terms0 = float(traj.shape[1]) * len(energies)
    spec = {}
    for prec in (abi.FP64, abi.MIXED):
        out = torch.zeros(len(energies), dtype=torch.float64, device=dev)
        for _ in range(2):
            abi.spectrum_device(ptrs, traj.shape[1], dt, (0, 0, 30), energies, q, cur, prec, out.data_ptr(),
                                stream.cuda_stream)
        torch.cuda.synchronize()
        abi.timing_enable(True)
        for _ in range(reps):
            abi.spectrum_device(ptrs, traj.shape[1], dt, (0, 0, 30), energies, q, cur, prec, out.data_ptr(),
                                stream.cuda_stream)
        kms, nl = abi.timing_read()
        abi.timing_enable(False)
        spec[prec] = (kms / max(nl, 1), out.cpu().numpy())
    s64, smx = spec[abi.FP64][1], spec[abi.MIXED][1]
    res["spectrum_c0"] = {"workload": "spectrum at (0,0,30 m), 2000 energies 101.45-3001.45 eV, NT=20001",
                          "fp64_terms_per_s": terms0 / (spec[abi.FP64][0] * 1e-3),
                          "mixed_terms_per_s": terms0 / (spec[abi.MIXED][0] * 1e-3),
                          "fp64_kernel_ms": spec[abi.FP64][0], "mixed_kernel_ms": spec[abi.MIXED][0],
                          "mixed_vs_fp64_max_rel_err": float(np.max(np.abs(smx - s64)) / np.max(np.abs(s64))),
                          "roofline": _frac(F_ALG_SPEC, terms0, spec[abi.FP64][0], spec[abi.MIXED][0])}
    # config 2
    osr = sr()
    osr.add_bfield_undulator(bfield=[0, 0.6, 0], period=[0, 0, 0.057], nperiods=30)
    osr.add_bfield_undulator(bfield=[0.6, 0, 0], period=[0, 0, 0.057], nperiods=30, phase=-np.pi / 2)
    osr.set_particle_beam(type="electron", name="beam", energy_GeV=3, x0=[0, 0, -1.5], d0=[0, 0, 1], current=0.5)
    osr.set_ctstartstop(0, 3)
    osr.set_npoints_trajectory(40001)
    osr.calculate_trajectory()
    traj, dt = osr.trajectory_array()
    q, cur = osr._e.particle_charge_current()
    surf = _sr.Surface.rectangle("XY", 301, 301, 0.06, 0.06, [0, 0, 0], [0, 0, 30], 0)
    xyz, nrm = surf.xyz, surf.normals
    d_traj = torch.from_numpy(np.ascontiguousarray(traj)).to(dev)
    ptrs = [d_traj[k].data_ptr() for k in range(9)]
    d_obs = torch.from_numpy(np.ascontiguousarray(xyz.T)).to(dev)
    d_nrm = torch.from_numpy(np.ascontiguousarray(nrm.T)).to(dev)
    terms2 = float(traj.shape[1]) * len(xyz)
    pdr = {}
    for prec in (abi.FP64, abi.MIXED):
        out = torch.zeros(len(xyz), dtype=torch.float64, device=dev)
        for _ in range(2):
            abi.power_density_device(ptrs, traj.shape[1], dt, d_obs.data_ptr(), d_nrm.data_ptr(), len(xyz), False, q,
                                     cur, prec, out.data_ptr(), stream.cuda_stream)
        torch.cuda.synchronize()
        abi.timing_enable(True)
        for _ in range(reps):
            abi.power_density_device(ptrs, traj.shape[1], dt, d_obs.data_ptr(), d_nrm.data_ptr(), len(xyz), False, q,
                                     cur, prec, out.data_ptr(), stream.cuda_stream)
        kms, nl = abi.timing_read()
        abi.timing_enable(False)
        pdr[prec] = (kms / max(nl, 1), out.cpu().numpy())
    p64, pmx = pdr[abi.FP64][1], pdr[abi.MIXED][1]
    area = (0.06 / 300) ** 2 * 1e6  # mm^2 per point (plots_mpl.py:474-489 sums the map this way)
    res["power_density_c2"] = {"workload": "EPU 2x30 periods 57 mm 0.6 T, NT=40001, 301x301 over 60 mm at 30 m",
                               "fp64_terms_per_s": terms2 / (pdr[abi.FP64][0] * 1e-3),
                               "mixed_terms_per_s": terms2 / (pdr[abi.MIXED][0] * 1e-3),
                               "fp64_kernel_ms": pdr[abi.FP64][0], "mixed_kernel_ms": pdr[abi.MIXED][0],
                               "mixed_vs_fp64_max_rel_err": float(np.max(np.abs(pmx - p64)) / np.max(np.abs(p64))),
                               "mixed_vs_fp64_rms_rel_err": float(np.sqrt(np.mean((pmx - p64) ** 2))
                                                                  / np.sqrt(np.mean(p64 ** 2))),
                               "integrated_power_W": {"fp64": float(p64.sum() * area), "mixed": float(pmx.sum() * area)},
                               "roofline": _frac(F_ALG_PD, terms2, pdr[abi.FP64][0], pdr[abi.MIXED][0])}
    log("spectrum C0: fp64 %.3f ms, mixed %.3f ms; power density C2: fp64 %.3f ms, mixed %.3f ms"
        % (spec[abi.FP64][0], spec[abi.MIXED][0], pdr[abi.FP64][0], pdr[abi.MIXED][0]))
    return res


def flux1_c1(reps=5):
    """CalculateFlux1's near- plus far-field flux (srs_hip_flux1, FP64) on the flux
    map's geometry: 201x201 over 20 mm at 30 m, 20001 trajectory points, 143.52 eV;
    kernel time from HIP events around flux1_fp64_kernel."""
    from srsoriginal_amd import abi
    traj, dt, xyz, q, cur = setup_problem()
    e = 143.52
    for _ in range(2):
        out = abi.flux1(traj, dt, xyz, e, q, cur)
    abi.timing_enable(True)
    for _ in range(reps):
        out = abi.flux1(traj, dt, xyz, e, q, cur)
    kms, nl = abi.timing_read()
    abi.timing_enable(False)
    kms /= max(nl, 1)
    terms = float(traj.shape[1]) * len(xyz)
    ach = F_ALG_FLUX1 * terms / (kms * 1e-3)
    log("flux1 C1 geometry: kernel %.3f ms" % kms)
    return {"flux1_c1": {"workload": "CalculateFlux1 (Lienard-Wiechert with A/c) 201x201, 20 mm at 30 m, "
                                     "143.52 eV, NT=20001, FP64",
                         "kernel": "flux1_fp64_kernel", "kernel_ms": kms, "terms_per_s": terms / (kms * 1e-3),
                         "peak_flux": float(out.max()),
                         "roofline": {"bound": "valu", "flop_per_term": F_ALG_FLUX1, "unit": "TFLOP/s",
                                      "achieved": ach / 1e12, "peak": PEAK_FP64_VALU / 1e12,
                                      "frac": ach / PEAK_FP64_VALU}}}


def batched_power_density(nparticles=256, reps=2):
    """Multi-particle power density (the C3 Gaussian beam on the C0 undulator, ragged
    18k-22k point trajectories) on 101x101 points over 40 mm at 30 m, FP64 and mixed, one
    device pipeline for all particles; timed with HIP events."""
    from srsoriginal_amd import abi, batch
    from srsoriginal_amd.lib import _sr
    from srsoriginal_amd.sr import sr
    osr = sr()
    osr.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.050], nperiods=41)
    osr.set_particle_beam(type="electron", name="beam", energy_GeV=3, x0=[0, 0, -1.5], d0=[0, 0, 1], current=0.5,
                          sigma_energy_GeV=8.9e-4 * 3, beta=[1.5, 0.8], emittance=[0.9e-9, 8e-12],
                          horizontal_direction=[1, 0, 0], lattice_reference=[0, 0, 0])
    osr.set_seed(4096)
    rows, q, cur = batch.sample_particles(osr, nparticles, 0.0, 3.0, 20001, ct_jitter=0.1, points_per_meter=6667,
                                          rng_seed=4096)
    surf = _sr.Surface.rectangle("XY", 101, 101, 0.04, 0.04, [0, 0, 0], [0, 0, 30], 0)
    npts = sum(abi.trajectory_points(abi.particles([r])) for r in rows)
    terms = float(npts) * len(surf.xyz)
    res = {}
    for prec in ("double", "mixed"):
        out = batch.power_density(osr, rows, q, cur, surf.xyz, surf.normals, precision=prec)  # warm-up
        ms = []
        for _ in range(reps):
            abi.timing_enable(True)
            out = batch.power_density(osr, rows, q, cur, surf.xyz, surf.normals, precision=prec)
            kms, _ = abi.timing_read()
            abi.timing_enable(False)
            ms.append(kms)
        res[prec] = (min(ms), out)
    t, out = res["double"]
    tm, om = res["mixed"]
    err = float(np.max(np.abs(om - out)) / np.max(np.abs(out)))
    log("batched power density: fp64 %.1f ms, mixed %.1f ms (%d particles)" % (t, tm, nparticles))
    return {"workload": "%d particles x %d traj points (ragged 18k-22k) x 101x101 over 40 mm at 30 m" % (nparticles, npts),
            "terms": terms, "fp64_ms": t, "fp64_terms_per_s": terms / (t * 1e-3),
            "mixed_ms": tm, "mixed_terms_per_s": terms / (tm * 1e-3), "mixed_vs_fp64_max_rel_err": err,
            "roofline": _pipeline_frac(F_ALG_PD, terms, t, tm),
            "peak_w_per_mm2": float(np.max(out)),
            "timed": "device RK4 in time segments overlapped with pd_batch_kernel, reduction (HIP events)"}


def batched_spectrum(nparticles=4096, reps=1):
    """Multi-particle spectrum (the C3 Gaussian beam on the C0 undulator, ragged 18k-22k
    point trajectories) at (0, 0, 30 m) over the C0 grid of 2000 energies, FP64 and mixed,
    one device pipeline; timed with HIP events."""
    from srsoriginal_amd import abi, batch
    from srsoriginal_amd.sr import sr
    osr = sr()
    osr.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.050], nperiods=41)
    osr.set_particle_beam(type="electron", name="beam", energy_GeV=3, x0=[0, 0, -1.5], d0=[0, 0, 1], current=0.5,
                          sigma_energy_GeV=8.9e-4 * 3, beta=[1.5, 0.8], emittance=[0.9e-9, 8e-12],
                          horizontal_direction=[1, 0, 0], lattice_reference=[0, 0, 0])
    osr.set_seed(4096)
    rows, q, cur = batch.sample_particles(osr, nparticles, 0.0, 3.0, 20001, ct_jitter=0.1, points_per_meter=6667,
                                          rng_seed=4096)
    energies = 100.0 + (3000.0 - 100.0) / 1999.0 * (np.arange(2000) + 1)  # TSpectrumContainer grid
    npts = sum(abi.trajectory_points(abi.particles([r])) for r in rows)
    terms = float(npts) * len(energies)
    res = {}
    for prec in ("double", "mixed"):
        out = batch.spectrum(osr, rows, q, cur, [0.0, 0.0, 30.0], energies, precision=prec)  # warm-up
        ms = []
        for _ in range(reps):
            abi.timing_enable(True)
            out = batch.spectrum(osr, rows, q, cur, [0.0, 0.0, 30.0], energies, precision=prec)
            kms, _ = abi.timing_read()
            abi.timing_enable(False)
            ms.append(kms)
        res[prec] = (min(ms), out)
    t, out = res["double"]
    tm, om = res["mixed"]
    log("batched spectrum: fp64 %.1f ms, mixed %.1f ms (%d particles)" % (t, tm, nparticles))
    return {"workload": "%d particles x %d traj points (ragged 18k-22k) x 2000 energies at (0,0,30 m)"
                        % (nparticles, npts),
            "terms": terms, "fp64_ms": t, "fp64_terms_per_s": terms / (t * 1e-3),
            "mixed_ms": tm, "mixed_terms_per_s": terms / (tm * 1e-3),
            "roofline": _pipeline_frac(F_ALG_SPEC, terms, t, tm),
            "mixed_vs_fp64_max_rel_err": float(np.max(np.abs(om - out)) / np.max(np.abs(out))),
            "peak_flux": float(np.max(out)),
            "timed": "device RK4 in time segments overlapped with spectrum_batch_kernel, reductions (HIP events)"}


def gradient_c4(nreal=256, reps=2):
    """Config-4 gradient: 256 pole-error realisations (82 poles, delta ~ N(0, 0.005^2),
    seeded) of the 41-period undulator, loss = sum over 51x51 points (5 mm at 30 m, E*)
    of (flux - ideal flux)^2; forward (RK4 with checkpoints every 1000 steps + FP64
    flux) and backward (flux adjoint + RK4 adjoint) timed with HIP events."""
    from srsoriginal_amd import abi, grad
    from srsoriginal_amd.lib import _sr
    from srsoriginal_amd.sr import sr
    osr = sr()
    osr.set_particle_beam(type="electron", name="beam", energy_GeV=3, x0=[0, 0, -1.5], d0=[0, 0, 1], current=0.5)
    st = osr._e.particle_state()
    q, cur = osr._e.particle_charge_current()
    row = dict(x0=st["x0"], beta0=st["beta0"], qmg=st["qmg"], t0=st["t0"], ct_start=0.0, ct_stop=3.0, npoints=20001)
    xyz = _sr.Surface.rectangle("XY", 51, 51, 0.005, 0.005, [0, 0, 0], [0, 0, 30], 0).xyz
    und = [dict(field=(0, 1, 0), period=(0, 0, 0.05), nperiods=41)]
    prob = grad.PoleProblem(und, 0, row, xyz, ENERGY_EV, q, cur, checkpoint_every=1000)
    prob.target = prob.flux(np.zeros((1, 82)))[0]
    deltas = np.random.default_rng(2024).normal(0.0, 0.005, (nreal, 82))
    loss, _, tape = prob.forward(deltas)  # warm-up (buffers)
    prob.backward(tape, nreal)
    fwd, bwd = [], []
    for _ in range(reps):
        abi.timing_enable(True)
        loss, _, tape = prob.forward(deltas)
        f_ms, _ = abi.timing_read()
        g = prob.backward(tape, nreal)
        b_ms, _ = abi.timing_read()
        abi.timing_enable(False)
        fwd.append(f_ms)
        bwd.append(b_ms)
    f_ms, b_ms = min(fwd), min(bwd)
    terms = float(nreal) * len(xyz) * 20001
    log("gradient C4: forward %.1f ms, backward %.1f ms" % (f_ms, b_ms))
    # the same in the mixed form (mixed flux terms forward, mixed adjoint), against FP64
    prob.precision = "mixed"
    lmx, _, tape = prob.forward(deltas)
    prob.backward(tape, nreal)
    mfwd, mbwd = [], []
    for _ in range(reps):
        abi.timing_enable(True)
        lmx, _, tape = prob.forward(deltas)
        mf, _ = abi.timing_read()
        gmx = prob.backward(tape, nreal)
        mb, _ = abi.timing_read()
        abi.timing_enable(False)
        mfwd.append(mf)
        mbwd.append(mb)
    mf, mb = min(mfwd), min(mbwd)
    gerr = np.max(np.abs(gmx - g), axis=1) / np.max(np.abs(g), axis=1)
    log("gradient C4 mixed: forward %.1f ms, backward %.1f ms, gradient vs FP64 max normwise %.2e"
        % (mf, mb, float(np.max(gerr))))
    return {"workload": "%d realisations x 82 poles, 51x51 over 5 mm at 30 m, E=%.2f eV, NT=20001, checkpoints "
                        "every 1000 steps" % (nreal, ENERGY_EV),
            "forward_ms": f_ms, "backward_ms": b_ms, "bwd_over_fwd": b_ms / f_ms,
            "gradient_over_forward": (f_ms + b_ms) / f_ms,
            "forward_terms_per_s": terms / (f_ms * 1e-3), "backward_terms_per_s": terms / (b_ms * 1e-3),
            "loss_mean": float(np.mean(loss)), "grad_rms": float(np.sqrt(np.mean(g ** 2))),
            "mixed": {"forward_ms": mf, "backward_ms": mb, "bwd_over_fwd": mb / mf,
                      "gradient_over_forward": (mf + mb) / mf, "gradient_speedup_vs_fp64": (f_ms + b_ms) / (mf + mb),
                      "gradient_vs_fp64_max_normwise": float(np.max(gerr)),
                      "loss_vs_fp64_max_rel": float(np.max(np.abs(lmx - loss) / np.abs(loss)))},
            "timed": "sum of kernel durations (HIP events): forward = RK4 + flux, backward = flux adjoint + RK4 adjoint"}


def gradient_pd(nreal=256, reps=2):
    """Power-density gradient on the config-4 field: 256 pole-error realisations, loss
    = sum over 51x51 points (20 mm at 30 m, the central radiation fan) of
    (P - ideal P)^2; forward (RK4 with checkpoints, storing A/c, + FP64 power density)
    and backward (power-density adjoint + RK4 adjoint) timed with HIP events."""
    from srsoriginal_amd import abi, grad
    from srsoriginal_amd.lib import _sr
    from srsoriginal_amd.sr import sr
    osr = sr()
    osr.set_particle_beam(type="electron", name="beam", energy_GeV=3, x0=[0, 0, -1.5], d0=[0, 0, 1], current=0.5)
    st = osr._e.particle_state()
    q, cur = osr._e.particle_charge_current()
    row = dict(x0=st["x0"], beta0=st["beta0"], qmg=st["qmg"], t0=st["t0"], ct_start=0.0, ct_stop=3.0, npoints=20001)
    surf = _sr.Surface.rectangle("XY", 51, 51, 0.02, 0.02, [0, 0, 0], [0, 0, 30], 0)
    xyz, nrm = surf.xyz, surf.normals
    und = [dict(field=(0, 1, 0), period=(0, 0, 0.05), nperiods=41)]
    prob = grad.PoleProblem(und, 0, row, xyz, None, q, cur, target=np.zeros(len(xyz)), checkpoint_every=1000,
                            quantity="power_density", normals=nrm)
    prob.target = prob.flux(np.zeros((1, 82)))[0]
    deltas = np.random.default_rng(2025).normal(0.0, 0.005, (nreal, 82))
    loss, _, tape = prob.forward(deltas)  # warm-up (buffers)
    prob.backward(tape, nreal)
    fwd, bwd = [], []
    for _ in range(reps):
        abi.timing_enable(True)
        loss, _, tape = prob.forward(deltas)
        f_ms, _ = abi.timing_read()
        g = prob.backward(tape, nreal)
        b_ms, _ = abi.timing_read()
        abi.timing_enable(False)
        fwd.append(f_ms)
        bwd.append(b_ms)
    f_ms, b_ms = min(fwd), min(bwd)
    terms = float(nreal) * len(xyz) * 20001
    log("gradient power density: forward %.1f ms, backward %.1f ms" % (f_ms, b_ms))
    return {"workload": "%d realisations x 82 poles, power density 51x51 over 20 mm at 30 m, NT=20001, checkpoints "
                        "every 1000 steps" % nreal,
            "forward_ms": f_ms, "backward_ms": b_ms, "bwd_over_fwd": b_ms / f_ms,
            "gradient_over_forward": (f_ms + b_ms) / f_ms,
            "forward_terms_per_s": terms / (f_ms * 1e-3), "backward_terms_per_s": terms / (b_ms * 1e-3),
            "loss_mean": float(np.mean(loss)), "grad_rms": float(np.sqrt(np.mean(g ** 2))),
            "timed": "sum of kernel durations (HIP events): forward = RK4 + power density, backward = power-density "
                     "adjoint + RK4 adjoint"}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


REF_HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
UND_C0 = "und1=0,1,0,0,0,0.05,41"
BEAM_C0 = ["x0=0,0,-1.5", "d0=0,0,1", "energy=3", "current=0.5", "ct=0,3"]
BEAM_C3 = BEAM_C0 + ["sigE=0.00267", "beta=1.5,0.8", "emit=0.9e-9,8e-12", "hdir=1,0,0", "latref=0,0,0", "seed=4096"]


def _ref_time(mode, args):
    """Wall time of one reference-engine run (oracle/_ref/ref_harness: the reference's own
    src/*.cc, its -O3 flags, one std::thread per output point as it threads), CPU only."""
    import subprocess
    import tempfile
    with tempfile.NamedTemporaryFile(suffix=".bin") as f:
        t0 = time.perf_counter()
        subprocess.run([REF_HARNESS, mode] + args + ["out=" + f.name], check=True, stdout=subprocess.DEVNULL)
        return time.perf_counter() - t0


def cpu_baselines():
    """The reference CPU path timed on this host for every config, with the host's thread
    share (16 on a one-GPU box) and single-threaded, each on a bounded sample whose terms
    are stated; C3 and C4 are extrapolated from their samples by the terms count.  Uses
    the reference engine itself when oracle/_ref/ref_harness was built (kind "reference"),
    else the oracle's restatement (kind "port", measured 2-2.6x faster than the reference
    engine on the same samples)."""
    threads = max(1, min(16, os.cpu_count() or 1))
    out = {"cpu_model": _cpu_model(), "threads": threads}
    if not os.path.exists(REF_HARNESS):
        out["kind"] = "port"
        return out
    out["kind"] = "reference"
    # (name, mode, args at `threads`, terms, args single-threaded, terms, full-config terms)
    E1 = "energy_ev=143.52176088044024"
    pole_scales = ",".join("%.17g" % x for x in np.random.default_rng(7).normal(0, 0.005, 82))
    plan = [
        ("c0_spectrum", "spectrum", [UND_C0] + BEAM_C0 + ["npoints=20001", "ne=2000", "erange=100,3000", "obs=0,0,30"],
         2000 * 20001.0, [UND_C0] + BEAM_C0 + ["npoints=20001", "ne=200", "erange=100,3000", "obs=0,0,30"],
         200 * 20001.0, 2000 * 20001.0),
        ("c1_flux", "flux", [UND_C0] + BEAM_C0 + ["npoints=20001", "plane=XY", "width=0.02,0.02", "nx=101,51",
                                                   "trans=0,0,30", E1], 101 * 51 * 20001.0,
         [UND_C0] + BEAM_C0 + ["npoints=20001", "plane=XY", "width=0.02,0.02", "nx=21,11", "trans=0,0,30", E1],
         21 * 11 * 20001.0, 201 * 201 * 20001.0),
        ("c2_power_density", "pd", ["und1=0,0.6,0,0,0,0.057,30", "und2=0.6,0,0,0,0,0.057,30,-1.5707963267948966"] +
         BEAM_C0 + ["npoints=40001", "plane=XY", "width=0.06,0.06", "nx=61,61", "trans=0,0,30"], 61 * 61 * 40001.0,
         ["und1=0,0.6,0,0,0,0.057,30", "und2=0.6,0,0,0,0,0.057,30,-1.5707963267948966"] + BEAM_C0 +
         ["npoints=40001", "plane=XY", "width=0.06,0.06", "nx=15,11", "trans=0,0,30"], 15 * 11 * 40001.0,
         301 * 301 * 40001.0),
        ("c3_batched_flux", "flux", [UND_C0] + BEAM_C3 + ["npoints=20001", "plane=XY", "width=0.01,0.01", "nx=11,11",
                                                           "trans=0,0,30", "energy_ev=431.2", "nparticles=16"],
         16 * 121 * 20001.0, [UND_C0] + BEAM_C3 + ["npoints=20001", "plane=XY", "width=0.01,0.01", "nx=5,5",
                                                   "trans=0,0,30", "energy_ev=431.2", "nparticles=4"],
         4 * 25 * 20001.0, 2.51e12),
        ("c4_gradient_forward", "flux", ["poles=0,1,0,0,0,0.05,41", "scales=" + pole_scales] + BEAM_C0 +
         ["npoints=20001", "plane=XY", "width=0.005,0.005", "nx=51,51", "trans=0,0,30", E1], 51 * 51 * 20001.0,
         ["poles=0,1,0,0,0,0.05,41", "scales=" + pole_scales] + BEAM_C0 +
         ["npoints=20001", "plane=XY", "width=0.005,0.005", "nx=13,13", "trans=0,0,30", E1], 13 * 13 * 20001.0,
         256 * 51 * 51 * 20001.0),
    ]
    for name, mode, a_mt, t_mt, a_1, t_1, t_full in plan:
        row = {}
        for label, args, terms, nt in (("threads", a_mt, t_mt, threads), ("single", a_1, t_1, 1)):
            el = _ref_time(mode, args + ["nthreads=%d" % nt])
            row[label] = {"terms_per_s": terms / el, "sample_terms": terms, "seconds": el, "threads": nt}
        row["full_config_terms"] = t_full
        row["full_config_s_at_threads"] = t_full / row["threads"]["terms_per_s"]
        row["extrapolated"] = t_full != row["threads"]["sample_terms"]
        out[name] = row
        log("cpu %s: %.3e terms/s on %d threads, %.3e single" % (name, row["threads"]["terms_per_s"], threads,
                                                                  row["single"]["terms_per_s"]))
    return out


def cpu_baseline(traj, dt, xyz, q, cur, budget_s=12.0, table=None):
    """The headline cpu_baseline object: the C1 flux rate of the reference engine (table from
    cpu_baselines) or, without it, of the oracle's restatement on a strided sample."""
    if table and table.get("kind") == "reference":
        r = table["c1_flux"]["threads"]
        return {"value": r["terms_per_s"], "unit": "terms/s", "cores": r["threads"], "kind": "reference",
                "sample": "101x51 of the 201x201 map points (same 20 mm screen) x 20001 trajectory steps, the "
                          "reference engine (oracle/_ref, its own sources and flags), %d threads, %.1f s; %s"
                          % (r["threads"], r["seconds"], table["cpu_model"])}
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads = max(1, min(16, os.cpu_count() or 1))
    # calibrate on a few points, then time a sample sized to ~budget_s
    idx = np.linspace(0, len(xyz) - 1, 2 * threads).astype(int)
    t0 = time.perf_counter()
    pyoracle.flux(traj, dt, xyz[idx], ENERGY_EV, q, cur, nthreads=threads)
    per_pt = (time.perf_counter() - t0) / len(idx)
    npts = int(min(len(xyz), max(threads, budget_s / max(per_pt, 1e-9))))
    idx = np.linspace(0, len(xyz) - 1, npts).astype(int)
    t0 = time.perf_counter()
    pyoracle.flux(traj, dt, xyz[idx], ENERGY_EV, q, cur, nthreads=threads)
    el = time.perf_counter() - t0
    terms = npts * traj.shape[1]
    return {"value": terms / el, "unit": "terms/s", "cores": threads, "kind": "port",
            "sample": "%d of the 40401 map points x %d trajectory steps, FP64 CPU path, %d threads, %.1f s"
                      % (npts, traj.shape[1], threads, el)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--no-other", action="store_true", help="skip the config-0 spectrum, config-2 power density and CalculateFlux1 lines")
    ap.add_argument("--profile", action="store_true",
                    help="one launch of each form plus a 256 MiB device copy (the HBM counter scale), for rocprofv3")
    ap.add_argument("--profile-spectrum", action="store_true",
                    help="one launch of each form of the config-0 spectrum, for rocprofv3")
    a = ap.parse_args()
    if a.gpus != 1:
        sys.exit("bench.py measures one GPU: this build runs on a single MI355X (multi-GPU sharding is design-only)")

    import torch
    from srsoriginal_amd import abi

    if a.profile_spectrum:
        traj, dt, q, cur, energies = spectrum_c0_setup()
        d_traj = torch.from_numpy(np.ascontiguousarray(traj)).to("cuda:0")
        ptrs = [d_traj[k].data_ptr() for k in range(9)]
        out = torch.zeros(len(energies), dtype=torch.float64, device="cuda:0")
        s = torch.cuda.current_stream().cuda_stream
        for prec in (abi.FP64, abi.MIXED):
            abi.spectrum_device(ptrs, traj.shape[1], dt, (0, 0, 30), energies, q, cur, prec, out.data_ptr(), s)
        torch.cuda.synchronize()
        log("profile run: one fp64 and one mixed config-0 spectrum launch")
        return

    traj, dt, xyz, q, cur = setup_problem()
    n, ns = traj.shape[1], len(xyz)
    dev = torch.device("cuda:0")
    d_traj = torch.from_numpy(np.ascontiguousarray(traj)).to(dev)
    d_obs = torch.from_numpy(np.ascontiguousarray(xyz.T)).to(dev)
    d_out = {p: torch.zeros(ns, dtype=torch.float64, device=dev) for p in (abi.FP64, abi.MIXED)}
    ptrs = [d_traj[k].data_ptr() for k in range(9)]
    stream = torch.cuda.current_stream()
    terms = float(n) * ns

    def step(prec):
        abi.flux_device(ptrs, n, dt, d_obs.data_ptr(), ns, [ENERGY_EV], q, cur, prec, d_out[prec].data_ptr(),
                        stream.cuda_stream)

    if a.profile:
        for prec in (abi.FP64, abi.MIXED):
            step(prec)
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("profile run: one fp64 and one mixed C1 launch, one 256 MiB copy")
        return

    res = {}
    for prec in (abi.FP64, abi.MIXED):
        for _ in range(a.warmup):
            step(prec)
        torch.cuda.synchronize()
        abi.timing_enable(True)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(a.steps):
            step(prec)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / a.steps
        kms, launches = abi.timing_read()
        abi.timing_enable(False)
        res[prec] = dict(ms=ms, kernel_ms=kms / max(launches, 1), out=d_out[prec].cpu().numpy())
        log("%s: %.3f ms/step, kernel %.3f ms" % ("fp64" if prec == 0 else "mixed", ms, res[prec]["kernel_ms"]))

    ref, mix = res[abi.FP64]["out"], res[abi.MIXED]["out"]
    max_rel = float(np.max(np.abs(mix - ref)) / np.max(np.abs(ref)))
    rms_rel = float(np.sqrt(np.mean((mix - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))
    value = terms / (res[abi.MIXED]["ms"] * 1e-3)
    kernel_s = res[abi.MIXED]["kernel_ms"] * 1e-3
    achieved = F_ALG_FLUX * terms / kernel_s
    out = {
        "metric": "obs*energy*traj-step terms/s (mixed precision, <=1e-5 rel err vs FP64)",
        "value": value,
        "unit": "terms/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": res[abi.MIXED]["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp64 phase + fp32 trig/amplitude/sums",
        "data": "synthetic (analytic undulator field, RK4 trajectory)",
        "config": {"workload": "flux map 201x201, 20 mm at 30 m, E=143.52 eV, ideal undulator 41x50 mm By=1 T, "
                               "3 GeV 0.5 A, RK4 NT=20001", "obs_points": ns, "energies": 1, "traj_points": n,
                   "terms_per_step": terms},
        "fp64": {"value": terms / (res[abi.FP64]["ms"] * 1e-3), "ms_per_step": res[abi.FP64]["ms"],
                 "kernel_ms": res[abi.FP64]["kernel_ms"],
                 "valu_frac": F_ALG_FLUX * terms / (res[abi.FP64]["kernel_ms"] * 1e-3) / PEAK_FP64_VALU},
        "mixed_vs_fp64": {"speedup": res[abi.FP64]["ms"] / res[abi.MIXED]["ms"], "max_rel_err": max_rel,
                          "rms_rel_err": rms_rel},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32_VALU / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_VALU, "traffic": TRAFFIC_MIXED_BYTES,
                     "issue_model": {"cycles_per_wave_term": ISSUE_CYCLES_PER_WAVE_TERM,
                                     "bound_ms_at_2p4GHz": terms / 64 * ISSUE_CYCLES_PER_WAVE_TERM / (SIMDS * CLOCK_HZ)
                                     * 1e3,
                                     "frac": terms / 64 * ISSUE_CYCLES_PER_WAVE_TERM / (SIMDS * CLOCK_HZ) / kernel_s},
                     "kernel": "flux_mixed2_kernel<256, 256>", "kernel_ms": res[abi.MIXED]["kernel_ms"],
                     "flop_per_term": F_ALG_FLUX},
    }
    if not a.no_batch:
        out["batched_c3"] = batched_c3()
        out["batched_power_density"] = batched_power_density()
        out["batched_spectrum"] = batched_spectrum()
    if not a.no_grad:
        out["gradient_c4"] = gradient_c4()
        out["gradient_power_density"] = gradient_pd()
    if not a.no_other:
        out.update(spectrum_and_power_density())
        out.update(flux1_c1())
        out.update(flux_general_c1(traj, dt, q, cur))
    if not a.no_cpu:
        log("cpu baselines ...")
        table = cpu_baselines()
        out["cpu_baseline"] = cpu_baseline(traj, dt, xyz, q, cur, table=table)
        out["cpu_baselines"] = table
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
