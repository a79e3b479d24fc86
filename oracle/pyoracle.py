"""ctypes wrapper of liboracle.so (the CPU oracle).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

D = C.c_double
P = C.POINTER(C.c_double)
I64 = C.c_int64


class Undulator(C.Structure):
    _fields_ = [("field", D * 3), ("period", D * 3), ("nperiods", C.c_int), ("phase", D), ("taper", D), ("center", D * 3),
                ("kind", C.c_int), ("width", D * 3), ("rotations", D * 3)]


class Fields(C.Structure):
    _fields_ = [("und", C.POINTER(Undulator)), ("nund", C.c_int), ("pole_und", C.POINTER(Undulator)), ("pole_scales", P),
                ("eund", C.POINTER(Undulator)), ("neund", C.c_int)]


class Particle(C.Structure):
    _fields_ = [("q", D), ("m", D), ("current", D), ("x0", D * 3), ("beta0", D * 3), ("t0", D), ("gamma", D), ("qmg", D)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        L = C.CDLL(path)
        L.oracle_trajectory.restype = I64
        L.oracle_trajectory_npoints.restype = I64
        L.oracle_trajectory_npoints.argtypes = [D, D, I64, D]
        L.oracle_total_power.restype = D
        L.oracle_rng_normal.restype = D
        L.oracle_rng_uniform.restype = D
        L.oracle_rect_surface.restype = C.c_int
        L.oracle_rect3_surface.restype = C.c_int
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(P)


def _v3(x):
    return (D * 3)(*[float(v) for v in x])


def undulator(field, period, nperiods, phase=0.0, taper=0.0, center=(0, 0, 0)):
    return dict(field=field, period=period, nperiods=nperiods, phase=phase, taper=taper, center=center)


class FieldSet:
    """Keeps the ctypes structures alive for the duration of the calls."""

    def __init__(self, undulators=(), pole=None, pole_scales=None, efields=()):
        def records(lst):
            arr = (Undulator * max(1, len(lst)))()
            for i, u in enumerate(lst):
                # kind 1 (Gaussian) / 2 (uniform box): field (already rotated), width, center, rotations
                arr[i] = Undulator(_v3(u["field"]), _v3(u.get("period", (0, 0, 0))), int(u.get("nperiods", 0)),
                                   float(u.get("phase", 0)), float(u.get("taper", 0)),
                                   _v3(u.get("center", (0, 0, 0))), int(u.get("kind", 0)),
                                   _v3(u.get("width", (0, 0, 0))), _v3(u.get("rotations", (0, 0, 0))))
            return arr
        self._und = records(undulators)
        self._eund = records(efields)
        self.s = Fields()
        self.s.und = self._und
        self.s.nund = len(undulators)
        self.s.eund = self._eund
        self.s.neund = len(efields)
        if pole is not None:
            self._pole = Undulator(_v3(pole["field"]), _v3(pole["period"]), int(pole["nperiods"]), 0.0, 0.0, _v3((0, 0, 0)),
                                   0, _v3((0, 0, 0)), _v3((0, 0, 0)))
            self._scales = np.ascontiguousarray(pole_scales, dtype=np.float64)
            self.s.pole_und = C.pointer(self._pole)
            self.s.pole_scales = _p(self._scales)


def ideal_electron(energy_gev, x0, d0, current):
    p = Particle()
    lib().oracle_ideal_electron(D(energy_gev), _v3(x0), _v3(d0), D(current), C.byref(p))
    return p


def rng_seed(seed, reset_normal=True):
    lib().oracle_rng_seed(C.c_int(int(seed)), C.c_int(1 if reset_normal else 0))


def rng_normal():
    return lib().oracle_rng_normal()


def rng_uniform():
    """TRandomA::Uniform from the oracle's global generator."""
    return lib().oracle_rng_uniform()


def random_electron(energy_gev, sigma_e_gev, x0, d0, hdir, beta, emit, lattice_ref, current):
    p = Particle()
    lib().oracle_random_electron(D(energy_gev), D(sigma_e_gev), _v3(x0), _v3(d0), _v3(hdir), (D * 2)(*beta), (D * 2)(*emit),
                                 _v3(lattice_ref), D(current), C.byref(p))
    return p


class Beam(C.Structure):
    _fields_ = [("e0", D), ("sigma_e", D), ("mass", D), ("charge", D), ("current", D), ("x0", D * 3), ("u0", D * 3),
                ("hdir", D * 3), ("vdir", D * 3), ("sigma_at", D * 3), ("sigma_u", D * 2), ("sigma_up", D * 2),
                ("t0", D)]


def philox4x32_10(ctr, key):
    out = (C.c_uint32 * 4)()
    lib().oracle_philox4x32_10((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), out)
    return list(out)


def counter_normals(seed, idx):
    out = (D * 6)()
    lib().oracle_counter_normals(C.c_uint64(seed), C.c_uint64(idx), out)
    return list(out)


def sample_beam(b, seed, idx):
    """The device sampler's particle idx (oracle_sample_beam); b: dict as beam_params."""
    bm = Beam(b["e0"], b["sigma_e"], b["mass"], b["charge"], b.get("current", 0.0), (D * 3)(*b["x0"]),
              (D * 3)(*b["u0"]), (D * 3)(*b["hdir"]), (D * 3)(*b["vdir"]), (D * 3)(*b["sigma_at"]),
              (D * 2)(*b["sigma_u"]), (D * 2)(*b["sigma_up"]), b.get("t0", 0.0))
    p = Particle()
    lib().oracle_sample_beam(C.byref(bm), C.c_uint64(seed), C.c_uint64(idx), C.byref(p))
    return p


def beam_index(weights, seed, idx):
    w = (D * len(weights))(*weights)
    return lib().oracle_beam_index(w, C.c_int(len(weights)), C.c_uint64(seed), C.c_uint64(idx))


def trajectory(fields, particle, ct_start, ct_stop, npoints):
    n = lib().oracle_trajectory_npoints(ct_start, ct_stop, npoints, particle.t0)
    t = np.zeros((9, n))
    dt = D()
    lib().oracle_trajectory(C.byref(fields.s), C.byref(particle), D(ct_start), D(ct_stop), I64(npoints), _p(t), C.byref(dt))
    return t, dt.value


def efield(fields, x):
    out = (D * 3)()
    lib().oracle_efield(C.byref(fields.s), _v3(x), out)
    return np.array(out[:])


def bfield(fields, x):
    out = (D * 3)()
    lib().oracle_bfield(C.byref(fields.s), _v3(x), out)
    return np.array(out[:])


def energy_grid(n, e0, e1):
    out = np.zeros(n)
    lib().oracle_energy_grid(I64(n), D(e0), D(e1), _p(out))
    return out


def rect_surface(plane, nx1, nx2, w1, w2, rot=(0, 0, 0), trans=(0, 0, 0), normal=0):
    n = nx1 * nx2
    xyz = np.zeros((n, 3)); nrm = np.zeros((n, 3)); x12 = np.zeros((n, 2))
    rc = lib().oracle_rect_surface(plane.encode(), C.c_int(nx1), C.c_int(nx2), D(w1), D(w2), _v3(rot), _v3(trans),
                                   C.c_int(normal), _p(xyz), _p(nrm), _p(x12))
    if rc != 0:
        raise ValueError("bad plane or normal")
    return xyz, nrm, x12


def rect3_surface(nx1, nx2, x0, x1, x2, normal=0):
    """Rectangle from three corners, already rotated and translated."""
    n = nx1 * nx2
    xyz = np.zeros((n, 3)); nrm = np.zeros((n, 3)); x12 = np.zeros((n, 2))
    rc = lib().oracle_rect3_surface(C.c_int(nx1), C.c_int(nx2), _v3(x0), _v3(x1), _v3(x2), C.c_int(normal), _p(xyz),
                                    _p(nrm), _p(x12))
    if rc != 0:
        raise ValueError("bad normal")
    return xyz, nrm, x12


def spectrum(traj, dt, obs, energies, q, current, weight=1.0, nthreads=1):
    traj = np.ascontiguousarray(traj); e = np.ascontiguousarray(energies, dtype=np.float64)
    out = np.zeros(len(e))
    lib().oracle_spectrum(_p(traj), I64(traj.shape[1]), D(dt), _v3(obs), _p(e), I64(len(e)), D(q), D(current), D(weight),
                          C.c_int(nthreads), _p(out))
    return out


def flux(traj, dt, obs, energy_ev, q, current, weight=1.0, nthreads=1):
    traj = np.ascontiguousarray(traj); obs = np.ascontiguousarray(obs, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(len(obs))
    lib().oracle_flux(_p(traj), I64(traj.shape[1]), D(dt), _p(obs), I64(len(obs)), D(energy_ev), D(q), D(current), D(weight),
                      C.c_int(nthreads), _p(out))
    return out


def flux1(traj, dt, obs, energy_ev, q, current):
    """CalculateFlux1's near- plus far-field flux on the points obs [n, 3] (traj needs A/c)."""
    traj = np.ascontiguousarray(traj); obs = np.ascontiguousarray(obs, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(len(obs))
    lib().oracle_flux1(_p(traj), I64(traj.shape[1]), D(dt), _p(obs), I64(len(obs)), D(energy_ev), D(q), D(current),
                       _p(out))
    return out


def power_density(traj, dt, obs, normals, directional, q, current, nthreads=1):
    traj = np.ascontiguousarray(traj); obs = np.ascontiguousarray(obs, dtype=np.float64).reshape(-1, 3)
    nrm = np.ascontiguousarray(normals, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(len(obs))
    lib().oracle_power_density(_p(traj), I64(traj.shape[1]), D(dt), _p(obs), _p(nrm), I64(len(obs)), C.c_int(1 if directional else 0),
                               D(q), D(current), C.c_int(nthreads), _p(out))
    return out


def total_power(traj, dt, q, current, gamma):
    traj = np.ascontiguousarray(traj)
    return lib().oracle_total_power(_p(traj), I64(traj.shape[1]), D(dt), D(q), D(current), D(gamma))


def kahan_add(s, c, v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    lib().oracle_kahan_add(_p(s), _p(c), _p(v), I64(len(v)))


def pole_loss_gradient(fields, particle, ct_start, ct_stop, npoints, obs, energy_ev, target, nthreads=1):
    """(L, dL/d delta) of L = sum_j (F_j - target_j)^2 for the pole-scaled undulator of
    `fields` (forward-mode dual numbers, FP64)."""
    obs = np.ascontiguousarray(obs, dtype=np.float64).reshape(-1, 3)
    target = np.ascontiguousarray(target, dtype=np.float64)
    grad = np.zeros(len(fields._scales))
    loss = D()
    rc = lib().oracle_pole_loss_gradient(C.byref(fields.s), C.byref(particle), D(ct_start), D(ct_stop), I64(npoints),
                                         _p(obs), I64(len(obs)), D(energy_ev), _p(target), C.byref(loss), _p(grad),
                                         C.c_int(nthreads))
    if rc != 0:
        raise ValueError("no pole-scaled undulator")
    return loss.value, grad


def pole_pd_loss_gradient(fields, particle, ct_start, ct_stop, npoints, obs, normals, target, directional=False,
                          nthreads=1):
    """(L, dL/d delta, P) of L = sum_j (P_j - target_j)^2 for the power density P on the
    points obs with normals (forward-mode dual numbers, FP64)."""
    obs = np.ascontiguousarray(obs, dtype=np.float64).reshape(-1, 3)
    nrm = np.ascontiguousarray(normals, dtype=np.float64).reshape(-1, 3)
    target = np.ascontiguousarray(target, dtype=np.float64)
    grad = np.zeros(len(fields._scales))
    pd = np.zeros(len(obs))
    loss = D()
    rc = lib().oracle_pole_pd_loss_gradient(C.byref(fields.s), C.byref(particle), D(ct_start), D(ct_stop),
                                            I64(npoints), _p(obs), _p(nrm), I64(len(obs)),
                                            C.c_int(1 if directional else 0), _p(target), _p(pd), C.byref(loss),
                                            _p(grad), C.c_int(nthreads))
    if rc != 0:
        raise ValueError("no pole-scaled undulator")
    return loss.value, grad, pd


def efield_time(traj, dt, obs, q):
    """(E [n, 3], arrival time [n]) of the time-domain electric field at obs."""
    traj = np.ascontiguousarray(traj)
    n = traj.shape[1]
    e = np.zeros((n, 3))
    t = np.zeros(n)
    lib().oracle_efield_time(_p(traj), I64(n), D(dt), _p(np.ascontiguousarray(obs, dtype=np.float64)), D(q), _p(e),
                             _p(t))
    return e, t


# ---- the integrals in extended precision (x86 80-bit long double, 64-bit mantissa) ----
# The reference's FP64 phase omega * (dt i + D/c) reaches ~5e11 rad, where one FP64 ulp is
# ~6e-5 rad: its rounding is a pseudo-random phase error per term, and where the integral
# cancels (between harmonics) it shows in the result at ~1e-5.  These evaluate the same
# integrand (OSCARSSR.cc:2054-2092, on the same FP64 trajectory, t_i = i dt exactly) with
# every step in long double (phase spacing ~3e-8 rad): the yardstick for how close the
# FP64 and mixed results are to the integral they both approximate.
_LD = np.longdouble


def _ld_consts(q, current):
    pi = _LD("3.14159265358979323846264338327950288")
    c = _LD("2.99792458e8")
    h, hbar, qe = _LD("6.62606876e-34"), _LD("1.054571596e-34"), _LD("1.602176462e-19")
    eps0, mu0 = _LD("8.854187817e-12"), _LD("1.2566370614e-6")
    q = _LD(q)
    c0 = q / (4 * pi * c * eps0 * np.sqrt(2 * pi))
    c2 = 4 * pi * _LD(current) / (h * abs(q) * mu0 * c) * _LD("1e-6") * _LD("1e-3")
    return c, qe, hbar, c0, c2


def _ld_sum(traj, dt, obs, omega, c):
    n = traj.shape[1]
    X, B = traj[:3].astype(_LD), traj[3:6].astype(_LD)
    t = np.arange(n, dtype=_LD) * _LD(dt)
    R = np.asarray(obs, dtype=_LD).reshape(3, 1) - X
    D = np.sqrt(np.sum(R * R, axis=0))
    N = R / D
    ph = omega * (t + D / c)
    cs, sn = np.cos(ph), np.sin(ph)
    g = c / (omega * D)  # the i c/(omega D) near-field factor
    are = (B - N) / D    # real part of the amplitude
    aim = -(N * g) / D   # imaginary part
    # (are + i aim)(cos + i sin)
    return np.sum(are * cs - aim * sn, axis=1), np.sum(are * sn + aim * cs, axis=1)


def flux_extended(traj, dt, obs, energy_ev, q, current):
    """Flux on obs [n, 3] at one energy, the integral in long double (see above)."""
    c, qe, hbar, c0, c2 = _ld_consts(q, current)
    omega = _LD(energy_ev) * qe / hbar
    k = c0 * omega * _LD(dt)
    out = np.zeros(len(obs))
    for j, o in enumerate(np.asarray(obs, dtype=np.float64).reshape(-1, 3)):
        re, im = _ld_sum(traj, dt, o, omega, c)
        out[j] = float(c2 * k * k * np.sum(re * re + im * im))
    return out


def spectrum_extended(traj, dt, obs, energies_ev, q, current):
    """Spectrum at obs over energies, the integral in long double (see above)."""
    return np.array([flux_extended(traj, dt, [obs], e, q, current)[0] for e in energies_ev])
