"""ctypes binding of the C ABI in include/srs_hip.h (libsrs_hip.so).

This is the thin Python view of the drop-in boundary: the same entry points a
cgo/JNI/N-API binding would declare.  Loading fails loudly when the library is
missing; there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libsrs_hip.so")

SRS_OK = 0
SRS_ERR_NO_DEVICE = 1
SRS_ERR_INVALID = 2
SRS_ERR_HIP = 3
FP64 = 0
MIXED = 1

_D = C.c_double
_P = C.POINTER(C.c_double)
_I64 = C.c_int64


class SrsTraj(C.Structure):
    _fields_ = [("n", _I64), ("dt", _D), ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("bx", C.c_void_p),
                ("by", C.c_void_p), ("bz", C.c_void_p), ("ax", C.c_void_p), ("ay", C.c_void_p), ("az", C.c_void_p)]


class SrsError(ValueError):
    """Raised for a non-zero status from libsrs_hip (maps to the reference's ValueError)."""


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libsrs_hip.so not built (%s): run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        L.srs_hip_last_error.restype = C.c_char_p
        for name, args in {
            "srs_hip_spectrum": [C.POINTER(SrsTraj), _P, _P, _I64, _D, _D, C.c_int, _P],
            "srs_hip_flux": [C.POINTER(SrsTraj), _P, _I64, _P, _I64, _D, _D, C.c_int, _P],
            "srs_hip_power_density": [C.POINTER(SrsTraj), _P, _P, _I64, C.c_int, _D, _D, C.c_int, _P],
            "srs_hip_flux1": [C.POINTER(SrsTraj), _P, _I64, _D, _D, _D, _P],
            "srs_hip_flux_device": [C.POINTER(SrsTraj), C.c_void_p, _I64, _P, _I64, _D, _D, C.c_int, C.c_void_p, C.c_void_p],
            "srs_hip_flux_device_noalloc": [C.POINTER(SrsTraj), C.c_void_p, _I64, _P, _I64, _D, _D, C.c_int,
                                            C.c_void_p, C.c_void_p],
            "srs_hip_spectrum_device": [C.POINTER(SrsTraj), _P, _P, _I64, _D, _D, C.c_int, C.c_void_p, C.c_void_p],
            "srs_hip_power_density_device": [C.POINTER(SrsTraj), C.c_void_p, C.c_void_p, _I64, C.c_int, _D, _D, C.c_int,
                                             C.c_void_p, C.c_void_p],
            "srs_hip_synchronize": [C.c_void_p],
            "srs_hip_timing_enable": [C.c_int],
            "srs_hip_timing_read": [C.POINTER(_D), C.POINTER(_I64)],
        }.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.srs_hip_device_count.restype = C.c_int
        _lib = L
    return _lib


def _check(rc):
    if rc != SRS_OK:
        raise SrsError(lib().srs_hip_last_error().decode())


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def traj_struct(traj, dt, keep):
    """srs_traj over a host [9][n] (or [6][n]) float64 array; `keep` holds references."""
    t = _f64(traj)
    keep.append(t)
    n = t.shape[1]
    ptr = [t[k].ctypes.data if k < t.shape[0] else None for k in range(9)]
    return SrsTraj(n, dt, *ptr)


def traj_struct_device(ptrs, n, dt):
    """srs_traj over nine device pointers (ints; ax..az may be None)."""
    return SrsTraj(n, dt, *ptrs)


def device_count():
    return lib().srs_hip_device_count()


def spectrum(traj, dt, obs, energies_ev, charge, current, precision=FP64):
    keep = []
    t = traj_struct(traj, dt, keep)
    e = _f64(energies_ev)
    o = _f64(obs)
    out = np.zeros(len(e))
    _check(lib().srs_hip_spectrum(C.byref(t), o.ctypes.data_as(_P), e.ctypes.data_as(_P), len(e), charge, current,
                                  precision, out.ctypes.data_as(_P)))
    return out


def flux(traj, dt, obs_xyz, energies_ev, charge, current, precision=FP64):
    keep = []
    t = traj_struct(traj, dt, keep)
    o = _f64(obs_xyz).reshape(-1, 3)
    e = _f64(np.atleast_1d(energies_ev))
    out = np.zeros(len(e) * len(o))
    _check(lib().srs_hip_flux(C.byref(t), o.ctypes.data_as(_P), len(o), e.ctypes.data_as(_P), len(e), charge, current,
                              precision, out.ctypes.data_as(_P)))
    return out.reshape(len(e), len(o))


def flux1(traj, dt, obs_xyz, energy_ev, charge, current):
    """CalculateFlux1's near- plus far-field flux (srs_hip_flux1); traj needs A/c."""
    keep = []
    t = traj_struct(traj, dt, keep)
    o = _f64(obs_xyz).reshape(-1, 3)
    out = np.zeros(len(o))
    _check(lib().srs_hip_flux1(C.byref(t), o.ctypes.data_as(_P), len(o), energy_ev, charge, current,
                               out.ctypes.data_as(_P)))
    return out


def power_density(traj, dt, obs_xyz, normals_xyz, directional, charge, current, precision=FP64):
    keep = []
    t = traj_struct(traj, dt, keep)
    o = _f64(obs_xyz).reshape(-1, 3)
    nn = _f64(normals_xyz).reshape(-1, 3)
    out = np.zeros(len(o))
    _check(lib().srs_hip_power_density(C.byref(t), o.ctypes.data_as(_P), nn.ctypes.data_as(_P), len(o),
                                       1 if directional else 0, charge, current, precision, out.ctypes.data_as(_P)))
    return out


def timing_enable(on=True):
    _check(lib().srs_hip_timing_enable(1 if on else 0))


def timing_read():
    """(total ms, launches) of the main integral kernels since timing_enable, from HIP events."""
    ms = _D()
    n = _I64()
    _check(lib().srs_hip_timing_read(C.byref(ms), C.byref(n)))
    return ms.value, n.value


def flux_device(traj_ptrs, n, dt, obs_soa_ptr, ns, energies_ev, charge, current, precision, out_ptr, stream):
    """Queue a flux evaluation on device-resident data (pointers as ints)."""
    t = SrsTraj(n, dt, *traj_ptrs)
    e = _f64(np.atleast_1d(energies_ev))
    _check(lib().srs_hip_flux_device(C.byref(t), obs_soa_ptr, ns, e.ctypes.data_as(_P), len(e), charge, current,
                                     precision, out_ptr, stream))


def flux_device_noalloc(traj_ptrs, n, dt, obs_soa_ptr, ns, energies_ev, charge, current, precision, out_ptr,
                        stream):
    """flux_device with the library's static device scratch (srs_hip_flux_device_noalloc)."""
    t = SrsTraj(n, dt, *traj_ptrs)
    e = _f64(np.atleast_1d(energies_ev))
    _check(lib().srs_hip_flux_device_noalloc(C.byref(t), obs_soa_ptr, ns, e.ctypes.data_as(_P), len(e), charge,
                                             current, precision, out_ptr, stream))


def power_density_device(traj_ptrs, n, dt, obs_soa_ptr, nrm_soa_ptr, ns, directional, charge, current, precision,
                         out_ptr, stream):
    t = SrsTraj(n, dt, *traj_ptrs)
    _check(lib().srs_hip_power_density_device(C.byref(t), obs_soa_ptr, nrm_soa_ptr, ns, 1 if directional else 0,
                                              charge, current, precision, out_ptr, stream))


def spectrum_device(traj_ptrs, n, dt, obs, energies_ev, charge, current, precision, out_ptr, stream):
    t = SrsTraj(n, dt, *traj_ptrs)
    e = _f64(energies_ev)
    o = _f64(obs)
    _check(lib().srs_hip_spectrum_device(C.byref(t), o.ctypes.data_as(_P), e.ctypes.data_as(_P), len(e), charge,
                                         current, precision, out_ptr, stream))


# ---- batched multi-particle path -------------------------------------------------

BATCH_GRAPH = 1
BATCH_AUTOTUNE = 2


FIELD_UNDULATOR, FIELD_GAUSSIAN, FIELD_UNIFORM_BOX, FIELD_GRID = 0, 1, 2, 3
FIELD_MAGNETIC, FIELD_ELECTRIC = 0, 1


class SrsField(C.Structure):
    _fields_ = [("field", _D * 3), ("period", _D * 3), ("nperiods", C.c_int32), ("phase", _D), ("taper", _D),
                ("center", _D * 3), ("pole_delta", C.c_void_p), ("kind", C.c_int32), ("width", _D * 3),
                ("rotations", _D * 3), ("grid", C.c_void_p), ("grid_n", C.c_int32 * 3), ("grid_start", _D * 3),
                ("grid_step", _D * 3), ("grid_stop", _D * 3), ("target", C.c_int32)]


class SrsParticle(C.Structure):
    _fields_ = [("x0", _D * 3), ("beta0", _D * 3), ("qmg", _D), ("t0", _D), ("ct_start", _D), ("ct_stop", _D),
                ("npoints", _I64), ("charge", _D), ("current", _D)]


def _bind_batch():
    L = lib()
    if getattr(L, "_batch_bound", False):
        return L
    L.srs_hip_trajectory_points.argtypes = [C.POINTER(SrsParticle), _I64, C.POINTER(_I64)]
    L.srs_hip_trajectories.argtypes = [C.POINTER(SrsField), C.c_int, C.POINTER(SrsParticle), _I64, _P, _I64,
                                       C.POINTER(_I64), _P]
    L.srs_hip_batch_flux.argtypes = [C.POINTER(SrsField), C.c_int, C.POINTER(SrsParticle), _I64, _D, _D, _P, _I64,
                                     _P, _I64, _D, C.c_int, C.c_int, _P]
    L.srs_hip_batch_power_density.argtypes = [C.POINTER(SrsField), C.c_int, C.POINTER(SrsParticle), _I64, _D, _D,
                                              _P, _P, _I64, C.c_int, _D, C.c_int, _P]
    L.srs_hip_batch_spectrum.argtypes = [C.POINTER(SrsField), C.c_int, C.POINTER(SrsParticle), _I64, _D, _D, _P,
                                         _P, _I64, _D, C.c_int, _P]
    L.srs_hip_batch_config.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.srs_hip_batch_points_per_lane.argtypes = [C.POINTER(C.c_int)]
    L.srs_hip_batch_set_config.argtypes = [C.c_int, C.c_int, C.c_int]
    L.srs_hip_pole_forward.argtypes = [C.POINTER(SrsField), C.c_int, C.c_int, C.POINTER(SrsParticle), _I64, _P, _D,
                                       _D, _P, _I64, _D, _P, C.c_int, _P, _P, C.POINTER(_I64)]
    L.srs_hip_pole_forward_ex.argtypes = [C.POINTER(SrsField), C.c_int, C.c_int, C.POINTER(SrsParticle), _I64, _P,
                                          _D, _D, _P, _I64, _D, _P, C.c_int, C.c_int, _P, _P, C.POINTER(_I64)]
    L.srs_hip_pole_pd_forward.argtypes = [C.POINTER(SrsField), C.c_int, C.c_int, C.POINTER(SrsParticle), _I64, _P,
                                          _D, _D, _P, _P, _I64, C.c_int, _P, C.c_int, _P, _P, C.POINTER(_I64)]
    L.srs_hip_pole_backward.argtypes = [_I64, _P, _P]
    for n in ("srs_hip_trajectory_points", "srs_hip_trajectories", "srs_hip_batch_flux", "srs_hip_batch_config",
              "srs_hip_batch_points_per_lane", "srs_hip_batch_set_config",
              "srs_hip_batch_power_density", "srs_hip_batch_spectrum",
              "srs_hip_pole_forward", "srs_hip_pole_forward_ex", "srs_hip_pole_pd_forward", "srs_hip_pole_backward"):
        getattr(L, n).restype = C.c_int
    L._batch_bound = True
    return L


def undulators(specs, keep):
    """srs_field array from dicts: undulators (field, period, nperiods[, phase, taper,
    center, pole_delta]) or, with kind FIELD_GAUSSIAN / FIELD_UNIFORM_BOX, (field,
    width[, center, rotations]), or with kind FIELD_GRID (grid [n0*n1*n2*3], grid_n,
    grid_start, grid_step, grid_stop[, center, rotations]); target FIELD_ELECTRIC for an
    electric field (default magnetic)."""
    arr = (SrsField * len(specs))()
    for i, u in enumerate(specs):
        pd = None
        if u.get("pole_delta") is not None:
            d = _f64(u["pole_delta"])
            keep.append(d)
            pd = d.ctypes.data
        grid = None
        if u.get("grid") is not None:
            g = _f64(u["grid"])
            keep.append(g)
            grid = g.ctypes.data
        arr[i] = SrsField((_D * 3)(*u.get("field", (0, 0, 0))), (_D * 3)(*u.get("period", (0, 0, 0))),
                          int(u.get("nperiods", 0)), float(u.get("phase", 0)), float(u.get("taper", 0)),
                          (_D * 3)(*u.get("center", (0, 0, 0))), pd, int(u.get("kind", FIELD_UNDULATOR)),
                          (_D * 3)(*u.get("width", (0, 0, 0))), (_D * 3)(*u.get("rotations", (0, 0, 0))), grid,
                          (C.c_int32 * 3)(*u.get("grid_n", (1, 1, 1))), (_D * 3)(*u.get("grid_start", (0, 0, 0))),
                          (_D * 3)(*u.get("grid_step", (0, 0, 0))), (_D * 3)(*u.get("grid_stop", (0, 0, 0))),
                          int(u.get("target", FIELD_MAGNETIC)))
    return arr


class SrsBeam(C.Structure):
    _fields_ = [("e0", _D), ("sigma_e", _D), ("mass", _D), ("charge", _D), ("x0", _D * 3), ("u0", _D * 3),
                ("hdir", _D * 3), ("vdir", _D * 3), ("sigma_at", _D * 3), ("sigma_u", _D * 2), ("sigma_up", _D * 2),
                ("t0", _D), ("ct_start", _D), ("ct_stop", _D), ("npoints", _I64), ("current", _D)]


def _srs_beam(beam, ct_start, ct_stop, npoints):
    return SrsBeam(float(beam["e0"]), float(beam["sigma_e"]), float(beam["mass"]), float(beam["charge"]),
                   (_D * 3)(*beam["x0"]), (_D * 3)(*beam["u0"]), (_D * 3)(*beam["hdir"]), (_D * 3)(*beam["vdir"]),
                   (_D * 3)(*beam["sigma_at"]), (_D * 2)(*beam["sigma_u"]), (_D * 2)(*beam["sigma_up"]),
                   float(beam.get("t0", 0.0)), float(ct_start), float(ct_stop), int(npoints),
                   float(beam.get("current", 0.0)))


def _rows(out):
    return [dict(x0=list(r.x0), beta0=list(r.beta0), qmg=r.qmg, t0=r.t0, ct_start=r.ct_start, ct_stop=r.ct_stop,
                 npoints=r.npoints, charge=r.charge, current=r.current) for r in out]


def sample_beam(beam, np_, seed, first=0, ct_start=0.0, ct_stop=1.0, npoints=2):
    """srs_hip_sample_beam: np_ particles of beam (a dict as sr._e.beam_params returns)
    drawn on the device from counters first .. first + np_ - 1 under key seed; returns
    particle rows (dicts of x0, beta0, qmg, t0, ct_start, ct_stop, npoints)."""
    L = lib()
    L.srs_hip_sample_beam.argtypes = [C.POINTER(SrsBeam), _I64, C.c_uint64, C.c_uint64, C.POINTER(SrsParticle)]
    L.srs_hip_sample_beam.restype = C.c_int
    b = _srs_beam(beam, ct_start, ct_stop, npoints)
    out = (SrsParticle * int(np_))()
    _check(L.srs_hip_sample_beam(C.byref(b), int(np_), C.c_uint64(int(seed)), C.c_uint64(int(first)), out))
    return _rows(out)


def sample_beams(beams, weights, np_, seed, first=0, ct_start=0.0, ct_stop=1.0, npoints=2):
    """srs_hip_sample_beams: as sample_beam, each particle from a beam chosen by weight."""
    L = lib()
    L.srs_hip_sample_beams.argtypes = [C.POINTER(SrsBeam), _P, C.c_int, _I64, C.c_uint64, C.c_uint64,
                                       C.POINTER(SrsParticle)]
    L.srs_hip_sample_beams.restype = C.c_int
    arr = (SrsBeam * len(beams))(*[_srs_beam(b, ct_start, ct_stop, npoints) for b in beams])
    w = _f64(weights)
    out = (SrsParticle * int(np_))()
    _check(L.srs_hip_sample_beams(arr, w.ctypes.data_as(_P), len(beams), int(np_), C.c_uint64(int(seed)),
                                  C.c_uint64(int(first)), out))
    return _rows(out)


def particles(rows):
    """srs_particle array from dicts (x0, beta0, qmg, t0, ct_start, ct_stop, npoints[, charge,
    current]); without a charge (or with 0) a particle takes the batch call's charge and current."""
    arr = (SrsParticle * len(rows))()
    for i, p in enumerate(rows):
        arr[i] = SrsParticle((_D * 3)(*p["x0"]), (_D * 3)(*p["beta0"]), float(p["qmg"]), float(p.get("t0", 0.0)),
                             float(p["ct_start"]), float(p["ct_stop"]), int(p["npoints"]), float(p.get("charge", 0.0)),
                             float(p.get("current", 0.0)))
    return arr


def trajectory_points(parts):
    L = _bind_batch()
    tot = _I64()
    _check(L.srs_hip_trajectory_points(parts, len(parts), C.byref(tot)))
    return tot.value


def trajectories(und_specs, parts):
    """Device RK4: returns (traj [9][total], offsets [np+1], dt [np])."""
    L = _bind_batch()
    keep = []
    U = undulators(und_specs, keep)
    tot = trajectory_points(parts)
    out = np.zeros((9, tot))
    offs = np.zeros(len(parts) + 1, dtype=np.int64)
    dts = np.zeros(len(parts))
    _check(L.srs_hip_trajectories(U, len(U), parts, len(parts), out.ctypes.data_as(_P), tot,
                                  offs.ctypes.data_as(C.POINTER(_I64)), dts.ctypes.data_as(_P)))
    return out, offs, dts


def batch_flux(und_specs, parts, charge, current, obs_xyz, energies_ev, weight, precision=FP64, flags=0):
    L = _bind_batch()
    keep = []
    U = undulators(und_specs, keep)
    o = _f64(obs_xyz).reshape(-1, 3)
    e = _f64(np.atleast_1d(energies_ev))
    out = np.zeros(len(e) * len(o))
    _check(L.srs_hip_batch_flux(U, len(U), parts, len(parts), charge, current, o.ctypes.data_as(_P), len(o),
                                e.ctypes.data_as(_P), len(e), weight, precision, flags, out.ctypes.data_as(_P)))
    return out.reshape(len(e), len(o))


def batch_flux_beams(und_specs, beams, weights, np_, seed, obs_xyz, energies_ev, weight, precision=FP64, flags=0,
                     first=0, ct_start=0.0, ct_stop=1.0, npoints=2):
    """srs_hip_batch_flux_beams: the batch flux of np_ particles drawn on the device (the
    draws of sample_beams, counters first .. first + np_ - 1 under key seed) straight into
    the pipeline; beams as sample_beams takes them, every particle with the beam's ct range
    and point count (ct_start, ct_stop, npoints, or per beam from its dict)."""
    L = lib()
    L.srs_hip_batch_flux_beams.argtypes = [C.POINTER(SrsField), C.c_int, C.POINTER(SrsBeam), _P, C.c_int, _I64,
                                           C.c_uint64, C.c_uint64, _P, _I64, _P, _I64, _D, C.c_int, C.c_int, _P]
    L.srs_hip_batch_flux_beams.restype = C.c_int
    keep = []
    U = undulators(und_specs, keep)
    arr = (SrsBeam * len(beams))(*[_srs_beam(b, b.get("ct_start", ct_start), b.get("ct_stop", ct_stop),
                                             b.get("npoints", npoints)) for b in beams])
    w = _f64(weights)
    o = _f64(obs_xyz).reshape(-1, 3)
    e = _f64(np.atleast_1d(energies_ev))
    out = np.zeros(len(e) * len(o))
    _check(L.srs_hip_batch_flux_beams(U, len(U), arr, w.ctypes.data_as(_P), len(beams), int(np_),
                                      C.c_uint64(int(seed)), C.c_uint64(int(first)), o.ctypes.data_as(_P), len(o),
                                      e.ctypes.data_as(_P), len(e), weight, precision, flags, out.ctypes.data_as(_P)))
    return out.reshape(len(e), len(o))


def batch_power_density(und_specs, parts, charge, current, obs_xyz, normals_xyz, directional, weight, precision=FP64):
    L = _bind_batch()
    keep = []
    U = undulators(und_specs, keep)
    o = _f64(obs_xyz).reshape(-1, 3)
    nrm = _f64(normals_xyz).reshape(-1, 3)
    if len(nrm) != len(o):
        raise SrsError("normals need one row per observation point")
    out = np.zeros(len(o))
    _check(L.srs_hip_batch_power_density(U, len(U), parts, len(parts), charge, current, o.ctypes.data_as(_P),
                                         nrm.ctypes.data_as(_P), len(o), 1 if directional else 0, weight, precision,
                                         out.ctypes.data_as(_P)))
    return out


def batch_spectrum(und_specs, parts, charge, current, obs, energies_ev, weight, precision=FP64):
    L = _bind_batch()
    keep = []
    U = undulators(und_specs, keep)
    o = _f64(obs)
    e = _f64(np.atleast_1d(energies_ev))
    out = np.zeros(len(e))
    _check(L.srs_hip_batch_spectrum(U, len(U), parts, len(parts), charge, current, o.ctypes.data_as(_P),
                                    e.ctypes.data_as(_P), len(e), weight, precision, out.ctypes.data_as(_P)))
    return out


def batch_config():
    """(energies per item, trajectory tile, observation points per lane) of the last batch call."""
    L = _bind_batch()
    a, b, p = C.c_int(), C.c_int(), C.c_int()
    _check(L.srs_hip_batch_config(C.byref(a), C.byref(b)))
    _check(L.srs_hip_batch_points_per_lane(C.byref(p)))
    return a.value, b.value, p.value


def batch_set_config(ne_per_item=0, tile=0, points_per_lane=1):
    """Fix the tile shape of later non-autotuned batch calls (ne_per_item=0: the default)."""
    L = _bind_batch()
    _check(L.srs_hip_batch_set_config(int(ne_per_item), int(tile), int(points_per_lane)))


def pole_forward(und_specs, poled, particle, deltas, charge, current, obs_xyz, energy_ev, target, checkpoint_every=1000,
                 want_flux=True, precision=FP64):
    """Forward of the pole-error flux loss for nreal realisations (deltas [nreal, 2N]),
    in FP64 or in the mixed form (MIXED: mixed flux terms forward and mixed adjoint).
    Returns (loss [nreal], flux [nreal, ns] or None, tape id)."""
    L = _bind_batch()
    keep = []
    U = undulators(und_specs, keep)
    d = _f64(np.atleast_2d(deltas))
    o = _f64(obs_xyz).reshape(-1, 3)
    t = _f64(target)
    if len(t) != len(o):
        raise SrsError("target needs one value per observation point")
    nreal = d.shape[0]
    loss = np.zeros(nreal)
    flux = np.zeros((nreal, len(o))) if want_flux else None
    tape = _I64()
    _check(L.srs_hip_pole_forward_ex(U, len(U), int(poled), C.byref(particle), nreal, d.ctypes.data_as(_P), charge,
                                     current, o.ctypes.data_as(_P), len(o), float(energy_ev), t.ctypes.data_as(_P),
                                     int(checkpoint_every), int(precision),
                                     flux.ctypes.data_as(_P) if want_flux else None, loss.ctypes.data_as(_P),
                                     C.byref(tape)))
    return loss, flux, tape.value


def pole_pd_forward(und_specs, poled, particle, deltas, charge, current, obs_xyz, normals_xyz, target, directional=False,
                    checkpoint_every=1000, want_map=True):
    """Forward of the pole-error power-density loss for nreal realisations (deltas
    [nreal, 2N]).  Returns (loss [nreal], power density [nreal, ns] or None, tape id)."""
    L = _bind_batch()
    keep = []
    U = undulators(und_specs, keep)
    d = _f64(np.atleast_2d(deltas))
    o = _f64(obs_xyz).reshape(-1, 3)
    nrm = _f64(normals_xyz).reshape(-1, 3)
    t = _f64(target)
    if len(t) != len(o) or len(nrm) != len(o):
        raise SrsError("target and normals need one value per observation point")
    nreal = d.shape[0]
    loss = np.zeros(nreal)
    pd = np.zeros((nreal, len(o))) if want_map else None
    tape = _I64()
    _check(L.srs_hip_pole_pd_forward(U, len(U), int(poled), C.byref(particle), nreal, d.ctypes.data_as(_P), charge,
                                     current, o.ctypes.data_as(_P), nrm.ctypes.data_as(_P), len(o),
                                     1 if directional else 0, t.ctypes.data_as(_P), int(checkpoint_every),
                                     pd.ctypes.data_as(_P) if want_map else None, loss.ctypes.data_as(_P),
                                     C.byref(tape)))
    return loss, pd, tape.value


def pole_backward(tape, nreal, npoles, dloss=None):
    """dL_r/d delta_p * dloss[r] for the realisations of forward call `tape`: [nreal, npoles]."""
    L = _bind_batch()
    g = np.zeros((nreal, npoles))
    dl = None
    if dloss is not None:
        dl = _f64(dloss)
        if len(dl) != nreal:
            raise SrsError("dloss needs one value per realisation")
    _check(L.srs_hip_pole_backward(int(tape), dl.ctypes.data_as(_P) if dl is not None else None, g.ctypes.data_as(_P)))
    return g
