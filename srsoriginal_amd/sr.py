"""`sr`: drop-in for the reference's Python type oscars.sr.sr on the radiation path.

Method names, keyword names, defaults, validation messages and the nested-list
return shapes follow src/OSCARSSR_Python.cc of the reference (method table
:3670-3747); each method cites the function it mirrors.  The engine underneath
(lib/_sr, srs_engine.cc) computes trajectories on the host in FP64 and runs
every radiation integral as a gfx950 HIP kernel through include/srs_hip.h.

Differences from the reference, by design:
  * the integrals always run on the GPU; `gpu=0/1` and `nthreads` keep their
    validation and select the reference's container layout per route;
  * `precision='double'` (default, matches the reference FP64 CPU results to
    1e-10) or `'mixed'` (FP64 phase, FP32 trig and sums, <= 1e-5);
  * `add_bfield_undulator(pole_scales=[...])` scales the 2N interior half-period
    poles by (1 + delta_p) (field-error studies and gradients);
  * `set_gpu_global` is stored and, like the reference's (OSCARSSR.cc:490-494),
    not consulted.
"""
import math
import sys

import numpy as np

from . import containers

try:
    from .lib import _sr
except ImportError as e:  # no silent fallback: the native engine is required
    raise ImportError("srsoriginal_amd native engine not built (lib/_sr*.so): run __graft_entry__.build()") from e

_PREC = {"double": 0, "fp64": 0, "mixed": 1}


def _vec3(v, name):
    if v is None or len(v) != 3:
        raise ValueError("Incorrect format in '%s'" % name)
    return [float(x) for x in v]


def _opt3(v, name, default=(0.0, 0.0, 0.0)):
    if v is None or len(v) == 0:
        return list(default)
    return _vec3(v, name)


_MAGNETIC, _ELECTRIC = 0, 1  # field container of an element (srs_field.target)


def _rotate_xyz(v, r):
    """TVector3D::RotateSelfXYZ (TVector3D.cc:183-207) for point/normal lists."""
    x, y, z = v
    for axis, a in enumerate(r):
        s, c = math.sin(a), math.cos(a)
        if axis == 0:
            y, z = c * y - s * z, s * y + c * z
        elif axis == 1:
            z, x = c * z - s * x, s * z + c * x
        else:
            x, y = c * x - s * y, s * x + c * y
    return [x, y, z]


def _prec(p):
    try:
        return _PREC[str(p).lower()]
    except KeyError:
        raise ValueError("'precision' must be 'double' or 'mixed'")


def _check_route(gpu, nthreads, nparticles=None):
    if nparticles is not None and nparticles < 0:
        raise ValueError("'nparticles' must be >= 1 (sort of)")
    if gpu != 0 and gpu != 1:
        raise ValueError("'gpu' must be 0 or 1")
    if nthreads < 0:
        raise ValueError("'nthreads' must be > 0")
    if nthreads > 0 and gpu == 1:
        raise ValueError("gpu is 1 and nthreads > 0.  Both are not currently allowed.")


def _container_list(xyz, v):
    return [[[float(a), float(b), float(c)], float(val)] for (a, b, c), val in zip(xyz.tolist(), v.tolist())]


class sr:
    """OSCARS SR radiation engine on MI355X (reference type oscars.sr.sr)."""

    def __init__(self):
        self._e = _sr.Engine()
        self._spectrum = None
        self._flux = None
        self._power_density = None
        self._numpy = False

    def set_output_format(self, fmt="list"):
        """Extension: 'list' (default) returns the reference's nested lists, built point by
        point as OSCARSSR_GetSpectrumAsList / OSCARSSR_GetT3DScalarAsList do
        (OSCARSSR_Python.cc:2046-2106); 'numpy' returns float64 arrays instead:
        (points [N, 3], values [N]) for flux and power-density maps, (energies [N],
        flux [N]) for spectra and (t [N], E [N, 3]) for calculate_efield_vs_time.  The
        add_to_* accumulators take either form."""
        if fmt not in ("list", "numpy"):
            raise ValueError("'fmt' must be 'list' or 'numpy'")
        self._numpy = fmt == "numpy"

    def _spec_out(self, e, f):
        if self._numpy:
            return np.array(e, dtype=np.float64), np.array(f, dtype=np.float64)
        return [[a, b] for a, b in zip(e.tolist(), f.tolist())]

    def _map_out(self, xyz, v):
        if self._numpy:
            return np.array(xyz, dtype=np.float64).reshape(-1, 3), np.array(v, dtype=np.float64)
        return _container_list(xyz, v)

    # ---- constants and globals (OSCARSSR_Python.cc:140-270) -------------------
    def pi(self):
        return 3.14159265358979323846

    def qe(self):
        return 1.602176462e-19

    def me(self):
        return 9.10938356e-31

    def rand(self):
        return self._e.rand()

    def norm(self):
        return self._e.norm()

    def set_seed(self, seed):
        """OSCARSSR_Python.cc:184-194 (float seed truncated to int)."""
        self._e.set_seed(int(float(seed)))

    def set_gpu_global(self, gpu):
        self._e.set_gpu_global(int(gpu))

    def check_gpu(self):
        """Number of gfx950 devices (OSCARSSR_Python.cc:216-230)."""
        return self._e.check_gpu()

    def set_nthreads_global(self, n):
        self._e.set_nthreads_global(int(n))

    def get_ctstart(self):
        return self._e.get_ctstart()

    def get_ctstop(self):
        return self._e.get_ctstop()

    def set_ctstartstop(self, start, stop):
        """OSCARSSR_Python.cc:282-299; resets the point count (OSCARSSR.cc:448-458)."""
        self._e.set_ctstartstop(float(start), float(stop))

    def set_npoints_trajectory(self, n):
        self._e.set_npoints_trajectory(int(n))

    def set_npoints_per_meter(self, n):
        self._e.set_npoints_per_meter(int(n))

    def get_npoints_trajectory(self):
        return self._e.get_npoints_trajectory()

    # ---- fields (OSCARSSR_Python.cc:736-818) ------------------------------------
    def add_bfield_undulator(self, bfield, period, nperiods, phase=0.0, rotations=None, translation=None, taper=0.0,
                             pole_scales=None):
        field = _vec3(bfield, "bfield")
        per = _vec3(period, "period")
        rot = _opt3(rotations, "rotations")
        trans = _opt3(translation, "translation")
        field = _rotate_xyz(field, rot)
        per = _rotate_xyz(per, rot)
        self._e.add_undulator(field, per, int(nperiods), float(phase), float(taper), trans,
                              [float(x) for x in (pole_scales if pole_scales is not None else [])], _MAGNETIC)

    def add_bfield_gaussian(self, bfield, sigma, rotations=None, translation=None):
        """OSCARSSR_AddMagneticFieldGaussian (OSCARSSR_Python.cc:561-636): bfield rotated by
        rotations, centred at translation; axes with sigma <= 0 are not bounded."""
        field = _vec3(bfield, "bfield")
        sig = _vec3(sigma, "sigma")
        rot = _opt3(rotations, "rotations")
        trans = _opt3(translation, "translation")
        self._e.add_field(1, _rotate_xyz(field, rot), sig, rot, trans, _MAGNETIC)

    def add_bfield_uniform(self, bfield, width=None, rotations=None, translation=None):
        """OSCARSSR_AddMagneticFieldUniform (OSCARSSR_Python.cc:640-718): bfield inside the
        box of the given full widths (axes with width <= 0, or no width, unbounded)."""
        field = _vec3(bfield, "bfield")
        w = _opt3(width, "width")
        rot = _opt3(rotations, "rotations")
        trans = _opt3(translation, "translation")
        self._e.add_field(2, _rotate_xyz(field, rot), w, rot, trans, _MAGNETIC)

    def add_bfield_file(self, ifile="", iformat="", rotations=None, translation=None, scale=None):
        """OSCARSSR_AddMagneticField (OSCARSSR_Python.cc:337-414): a field map read from
        ifile in iformat 'OSCARS', 'OSCARS1D <columns>', 'SRW' or 'SPECTRA'
        (TField3D_Grid.cc:318-1146); scale multiplies the steps (OSCARS: x, y, z then the
        three field components) or the columns in order (OSCARS1D)."""
        if not ifile or not iformat:
            raise ValueError("'ifile' or 'iformat' is blank")
        rot = _opt3(rotations, "rotations")
        trans = _opt3(translation, "translation")
        sc = [float(v) for v in (scale or [])]
        try:
            self._e.add_bfield_file(str(ifile), str(iformat), rot, trans, sc, _MAGNETIC)
        except Exception:
            raise ValueError("Could not import magnetic field.  Check 'ifile' and 'iformat' are correct")

    def add_bfield_interpolated(self, mapping, iformat, parameter, rotations=None, translation=None, scale=None):
        """OSCARSSR_AddMagneticFieldInterpolated (OSCARSSR_Python.cc:419-521): mapping is a
        list of [parameter value, file] pairs of OSCARS-format maps with one header; every
        grid value is interpolated linearly in the parameter between the two maps that
        bracket it (TField3D_Grid.cc:1150-1445)."""
        pairs = []
        for m in mapping or []:
            if len(m) != 2:
                raise ValueError("Incorrect format in 'mapping'")
            f = m[1].decode() if isinstance(m[1], bytes) else str(m[1])
            pairs.append((float(m[0]), f))
        if not iformat:
            raise ValueError("'iformat' is blank")
        rot = _opt3(rotations, "rotations")
        trans = _opt3(translation, "translation")
        sc = [float(v) for v in (scale or [])]
        try:
            self._e.add_bfield_interpolated(pairs, str(iformat), float(parameter), rot, trans, sc)
        except Exception:
            raise ValueError("Could not import magnetic field.  Check filenames and 'iformat' are correct")

    def get_bfield(self, x):
        return self._e.get_bfield(_vec3(x, "x"))

    def clear_bfields(self):
        self._e.clear_bfields()

    # ---- electric fields (OSCARSSR_Python.cc:876-1290, 1306-1400) -----------------
    # The same field elements in the electric container (OSCARSSR::AddElectricField); the
    # trajectory's force is then q/(m gamma) (E + v x B) (OSCARSSR::DerivativesE / EB,
    # OSCARSSR.cc:690-794), on the host and in the device RK4.
    def add_efield_undulator(self, efield, period, nperiods, phase=0.0, rotations=None, translation=None):
        """OSCARSSR_AddElectricFieldIdealUndulator (OSCARSSR_Python.cc:1156-1240): field and
        period rotated, centred at translation, no taper."""
        field = _rotate_xyz(_vec3(efield, "efield"), _opt3(rotations, "rotations"))
        per = _rotate_xyz(_vec3(period, "period"), _opt3(rotations, "rotations"))
        self._e.add_undulator(field, per, int(nperiods), float(phase), 0.0, _opt3(translation, "translation"), [],
                              _ELECTRIC)

    def add_efield_gaussian(self, efield, sigma, rotations=None, translation=None):
        """OSCARSSR_AddElectricFieldGaussian (OSCARSSR_Python.cc:994-1069)."""
        field = _vec3(efield, "efield")
        sig = _vec3(sigma, "sigma")
        rot = _opt3(rotations, "rotations")
        self._e.add_field(1, _rotate_xyz(field, rot), sig, rot, _opt3(translation, "translation"), _ELECTRIC)

    def add_efield_uniform(self, efield, width=None, rotations=None, translation=None):
        """OSCARSSR_AddElectricFieldUniform (OSCARSSR_Python.cc:1073-1150)."""
        field = _vec3(efield, "efield")
        rot = _opt3(rotations, "rotations")
        self._e.add_field(2, _rotate_xyz(field, rot), _opt3(width, "width"), rot, _opt3(translation, "translation"),
                          _ELECTRIC)

    def add_efield_file(self, ifile="", iformat="", rotations=None, translation=None, scale=None):
        """OSCARSSR_AddElectricField (OSCARSSR_Python.cc:876-954): a field map, as add_bfield_file."""
        if not ifile or not iformat:
            raise ValueError("'ifile' or 'iformat' is blank")
        rot = _opt3(rotations, "rotations")
        trans = _opt3(translation, "translation")
        sc = [float(v) for v in (scale or [])]
        try:
            self._e.add_bfield_file(str(ifile), str(iformat), rot, trans, sc, _ELECTRIC)
        except Exception:
            raise ValueError("Could not import electric field.  Check 'ifile' and 'iformat' are correct")

    def add_field_gaussian(self, sigma, bfield=None, efield=None, rotations=None, translation=None):
        """OSCARSSR_AddFieldGaussian (OSCARSSR_Python.cc:1306-1400): at most one of bfield and
        efield.  As in the reference, a bfield adds nothing (that branch is commented out,
        :1392-1396) and otherwise an electric Gaussian is added (field 0 without efield)."""
        has_b = bfield is not None and len(bfield) > 0
        has_e = efield is not None and len(efield) > 0
        if has_b and has_e:
            raise ValueError("You cannot define both 'bfield' and 'efield'")
        field = _vec3(bfield, "bfield") if has_b else (_vec3(efield, "efield") if has_e else [0.0, 0.0, 0.0])
        sig = _vec3(sigma, "sigma")
        rot = _opt3(rotations, "rotations")
        trans = _opt3(translation, "translation")
        if not has_b:
            self._e.add_field(1, _rotate_xyz(field, rot), sig, rot, trans, _ELECTRIC)

    def get_efield(self, x):
        """OSCARSSR_GetEField (OSCARSSR_Python.cc:1244-1278): the summed electric field."""
        return self._e.get_efield(_vec3(x, "x"))

    def clear_efields(self):
        self._e.clear_efields()

    # ---- beams (OSCARSSR_Python.cc:1641-1938) -----------------------------------
    def add_particle_beam(self, type="", name="", energy_GeV=0.0, d0=None, x0=None, sigma_energy_GeV=0.0, t0=0.0,
                          current=0.0, weight=1.0, rotations=None, translation=None, horizontal_direction=None,
                          beta=None, emittance=None, lattice_reference=None, mass=0.0, charge=0.0):
        """OSCARSSR_AddParticleBeam, OSCARSSR_Python.cc:1661-1870."""
        if not type or not name:
            raise ValueError("'type' or 'name' is blank")
        pos = _opt3(x0, "x0")
        direction = _opt3(d0, "d0")
        if sigma_energy_GeV < 0:
            raise ValueError("'sigma_energy_GeV' cannot be less than zero")
        rot = _opt3(rotations, "rotations")
        trans = _opt3(translation, "translation")
        if horizontal_direction is not None and len(horizontal_direction) != 0:
            hdir = _vec3(horizontal_direction, "horizontal_direction")
        else:
            a = [abs(v) for v in direction]
            dx, dy, dz = direction
            if a[0] < a[1]:
                hdir = [0.0, dz, -dy] if a[0] < a[2] else [dy, -dx, 0.0]
            else:
                hdir = [-dz, 0.0, dx] if a[1] < a[2] else [dy, -dx, 0.0]
        su, sup = [0.0, 0.0], [0.0, 0.0]
        if beta is not None and len(beta) and emittance is not None and len(emittance):
            if len(beta) != 2:
                raise ValueError("Incorrect format in 'beta'")
            if len(emittance) != 2:
                raise ValueError("Incorrect format in 'emittance'")
            su = [math.sqrt(emittance[0] * beta[0]), math.sqrt(emittance[1] * beta[1])]
            sup = [math.sqrt(emittance[0] / beta[0]), math.sqrt(emittance[1] / beta[1])]
        latref = _opt3(lattice_reference, "lattice_reference")
        pos = _rotate_xyz(pos, rot)
        pos = [pos[i] + trans[i] for i in range(3)]
        if type == "custom" and (mass == 0 or charge == 0):
            raise ValueError("'mass' or 'charge' is zero")
        try:
            self._e.add_particle_beam(type, name, pos, direction, float(energy_GeV), float(t0), float(current),
                                      float(weight))
        except ValueError:
            raise ValueError("invalid argument in adding particle beam.  possibly 'name' already exists")
        n = math.sqrt(hdir[0] ** 2 + hdir[1] ** 2 + hdir[2] ** 2)
        hdir = [h / n for h in hdir]
        self._e.set_sigma(name, hdir, su, sup, latref, float(sigma_energy_GeV))

    def set_particle_beam(self, **kw):
        """OSCARSSR_SetParticleBeam, OSCARSSR_Python.cc:1636-1655: clear, add, ideal particle."""
        self._e.clear_particle_beams()
        self.add_particle_beam(**kw)
        self._e.set_new_particle_from("", "ideal")

    def clear_particle_beams(self):
        self._e.clear_particle_beams()

    def set_new_particle(self, beam="", particle=""):
        """OSCARSSR_SetNewParticle, OSCARSSR_Python.cc:1891-1938."""
        if self._e.n_beams() < 1:
            raise ValueError("No particle beam defined")
        b, p = beam.lower(), particle.lower()
        if p != "" and p not in ("ideal", "random"):
            raise ValueError("'particle' must be 'ideal' or 'random'")
        try:
            if b == "" and p == "":
                self._e.set_new_particle()
            else:
                self._e.set_new_particle_from(beam, p)
        except ValueError:
            raise ValueError("'beam' name not found")

    def get_particle_x0(self):
        return self._e.particle_x0()

    def get_particle_beta0(self):
        return self._e.particle_beta0()

    def get_particle_e0(self):
        return self._e.particle_e0()

    # ---- trajectory (OSCARSSR_Python.cc:1988-2037) ------------------------------
    def calculate_trajectory(self, gpu=0):
        """Extension: gpu=1 integrates the RK4 on the device (srs_hip_trajectories);
        the result agrees with the host integration to ~1e-14 relative."""
        if gpu != 0 and gpu != 1:
            raise ValueError("'gpu' must be 0 or 1")
        self._e.calculate_trajectory(int(gpu))
        return self.get_trajectory()

    def get_trajectory(self):
        t, _ = self._e.get_trajectory()
        x = t[0:3].T.tolist()
        b = t[3:6].T.tolist()
        return [[xi, bi] for xi, bi in zip(x, b)]

    def trajectory_array(self):
        """Extension: the last trajectory as a (9, n) float64 array (x, beta, A/c) and dt."""
        return self._e.get_trajectory()

    # ---- spectrum (OSCARSSR_Python.cc:2215-2351) --------------------------------
    def calculate_spectrum(self, obs=None, npoints=0, energy_range_eV=None, points_eV=None, nparticles=0, nthreads=0,
                           gpu=0, ofile="", bofile="", precision="double", batched=1):
        """Extension: with gpu=1 and nparticles > 1, batched=1 (default) draws all
        particles first and integrates them in one device pipeline (FP64 or mixed);
        batched=0 keeps the reference's per-particle loop."""
        if nparticles < 0:
            raise ValueError("'nparticles' must be >= 1 (sort of)")
        pts = [float(v) for v in (points_eV or [])]
        e0 = e1 = 0.0
        if energy_range_eV is not None and len(energy_range_eV) != 0:
            if len(energy_range_eV) != 2:
                raise ValueError("'energy_range_eV' must be a list of length 2")
            e0, e1 = float(energy_range_eV[0]), float(energy_range_eV[1])
        o = _vec3(obs, "obs")
        _check_route(gpu, nthreads)
        if not pts and npoints < 1:
            raise ValueError("'npoints' must be > 0")
        e, f = self._e.calculate_spectrum(o, pts, int(npoints), e0, e1, int(nparticles), int(nthreads), int(gpu),
                                          _prec(precision), bool(batched))
        if ofile:
            containers.write_spectrum_text(ofile, e, f)
        if bofile:
            containers.write_spectrum_binary(bofile, e, f)
        return self._spec_out(e, f)

    # ---- flux (OSCARSSR_Python.cc:2829-2996, :3009-3233) --------------------------
    def calculate_flux_rectangle(self, energy_eV, npoints, plane="", normal=0, dim=2, width=None, rotations=None,
                                 translation=None, x0x1x2=None, nparticles=0, polarization="", nthreads=0, gpu=0,
                                 ofile="", precision="double", batched=1):
        """Extension: with gpu=1 and nparticles > 1, batched=1 (default) draws all
        particles first and integrates them in one device pipeline (trajectories on the
        device); batched=0 keeps the reference's per-particle loop."""
        if self._e.n_beams() < 1:
            raise ValueError("No particle beam defined")
        if dim != 2 and dim != 3:
            raise ValueError("'dim' must be 2 or 3")
        surf = self._rectangle(npoints, plane, width, rotations, translation, x0x1x2, normal)
        self._check_common(nparticles, gpu, nthreads)
        xyz, v = self._e.calculate_flux(surf, float(energy_eV), int(nparticles), int(nthreads), int(gpu), int(dim),
                                        _prec(precision), bool(batched))
        return self._out(xyz, v, ofile, dim)

    def calculate_flux(self, energy_eV, points, normal=0, rotations=None, translation=None, nparticles=0, nthreads=0,
                       gpu=0, ofile="", precision="double", batched=1):
        if self._e.n_beams() < 1:
            raise ValueError("No particle beam defined")
        surf = self._points(points, rotations, translation, normal, flip=False)
        self._check_common(nparticles, gpu, nthreads)
        xyz, v = self._e.calculate_flux(surf, float(energy_eV), int(nparticles), int(nthreads), int(gpu), 3,
                                        _prec(precision), bool(batched))
        return self._out(xyz, v, ofile, 3)

    def calculate_flux1_rectangle(self, energy_eV, npoints, plane="", normal=0, width=None, rotations=None,
                                  translation=None, x0x1x2=None, ofile=""):
        """OSCARSSR::CalculateFlux1 (OSCARSSR.cc:2209-2346), which the reference keeps in
        C++ only: the flux of the current particle from the Lienard-Wiechert field with
        the stored A/c (velocity and acceleration terms, no integration by parts;
        Omega = 2 pi E / 4.1357e-15 as written there), summed on the device in FP64.
        Returns [[[x1, x2, 0], flux], ...]; with ofile, writes its "X1 X2 flux" lines
        (std::scientific) there instead and the values stay 0, as the reference does."""
        if self._e.n_beams() < 1:
            raise ValueError("No particle beam defined")
        surf = self._rectangle(npoints, plane, width, rotations, translation, x0x1x2, normal)
        xyz, v = self._e.calculate_flux1(surf, float(energy_eV), str(ofile))
        return self._out(xyz, v, "", 2)

    def calculate_flux1(self, energy_eV, points, normal=0, rotations=None, translation=None, ofile=""):
        """calculate_flux1_rectangle on a list of points ([[x, y, z], [nx, ny, nz]] as
        calculate_flux takes them); returns [[[x, y, z], flux], ...]; ofile lines carry
        X1 = X2 = 0, as TSurfacePoints_3D::GetX1/GetX2 return for point lists."""
        if self._e.n_beams() < 1:
            raise ValueError("No particle beam defined")
        surf = self._points(points, rotations, translation, normal, flip=False)
        xyz, v = self._e.calculate_flux1(surf, float(energy_eV), str(ofile))
        return self._map_out(xyz, v)

    # ---- power density (OSCARSSR_Python.cc:2403-2577, :2587-2807) ----------------
    def calculate_power_density_rectangle(self, npoints, plane="", width=None, x0x1x2=None, rotations=None,
                                          translation=None, ofile="", normal=0, nparticles=0, gpu=0, nthreads=0,
                                          dim=2, precision="double", batched=1):
        """Extension: with gpu=1 and nparticles > 1, batched=1 (default) draws all
        particles first and integrates them in one device pipeline (FP64 or mixed);
        batched=0 keeps the reference's per-particle loop."""
        if self._e.n_beams() < 1:
            raise ValueError("No particle beam defined")
        if dim != 2 and dim != 3:
            raise ValueError("'dim' must be 2 or 3")
        surf = self._rectangle(npoints, plane, width, rotations, translation, x0x1x2, normal)
        self._check_common(nparticles, gpu, nthreads)
        xyz, v = self._e.calculate_power_density(surf, int(dim), normal != 0, int(nparticles), int(nthreads), int(gpu),
                                                 _prec(precision), bool(batched))
        return self._out(xyz, v, ofile, dim)

    def calculate_power_density(self, points, normal=0, rotations=None, translation=None, nparticles=0, gpu=0,
                                nthreads=0, ofile="", precision="double", batched=1):
        if self._e.n_beams() < 1:
            raise ValueError("No particle beam defined")
        surf = self._points(points, rotations, translation, normal, flip=True)
        self._check_common(nparticles, gpu, nthreads)
        xyz, v = self._e.calculate_power_density(surf, 3, normal != 0, int(nparticles), int(nthreads), int(gpu),
                                                 _prec(precision), bool(batched))
        return self._out(xyz, v, ofile, 3)

    def calculate_efield_vs_time(self, obs, ofile=""):
        """OSCARSSR_CalculateElectricFieldTimeDomain (OSCARSSR_Python.cc:3573-3630):
        [[t, [Ex, Ey, Ez]], ...] at unevenly spaced arrival times t (trajectory order);
        ofile writes rows "Ex Ey Ez t" (T3DScalarContainer::WriteToFileText, dim 3)."""
        if self._e.n_beams() < 1:
            raise ValueError("No particle beam defined")
        o = _vec3(obs, "obs")
        E, t = self._e.calculate_efield_vs_time(o)
        if ofile:
            containers.write_scalar_text(ofile, E, t, 3)
        if self._numpy:
            return np.array(t, dtype=np.float64), np.array(E, dtype=np.float64).reshape(-1, 3)
        return [[float(a), [float(x), float(y), float(z)]] for a, (x, y, z) in zip(t.tolist(), E.tolist())]

    def calculate_total_power(self, gpu=0):
        """OSCARSSR_Python.cc:2369-2398: the host sum in the reference's order, or with
        gpu=1 (an extension) the device sum (srs_hip_total_power)."""
        if self._e.n_beams() < 1:
            raise ValueError("No particle beam defined")
        if gpu not in (0, 1):
            raise ValueError("'gpu' must be 0 or 1")
        return self._e.calculate_total_power_device() if gpu else self._e.calculate_total_power()

    # ---- container files (OSCARSSR_Python.cc:3252-3318, :3475-3563) -------------
    @staticmethod
    def _files(ifiles, bifiles):
        ti = list(ifiles or [])
        bi = list(bifiles or [])
        if ti and bi:
            raise ValueError("either text or binary files may be added, but not both.")
        if not ti and not bi:
            raise ValueError("No files given.  You need at least one file as input in a list.")
        return ti, bi

    def average_spectra(self, ifiles=None, bifiles=None, ofile="", bofile=""):
        """OSCARSSR_AverageSpectra: average of text spectrum files.  As in the reference,
        binary inputs are not read (the call returns an empty spectrum)."""
        ti, bi = self._files(ifiles, bifiles)
        if ti:
            e, f = containers.average_spectrum_files(ti)
        else:
            e, f = np.zeros(0), np.zeros(0)
        if ofile:
            containers.write_spectrum_text(ofile, e, f)
        if bofile:
            containers.write_spectrum_binary(bofile, e, f)
        return self._spec_out(e, f)

    def average_flux(self, ifiles=None, bifiles=None, ofile="", bofile="", dim=2):
        """OSCARSSR_AverageT3DScalars: average of flux (or power-density) files, text or
        binary, written as text (ofile) and binary (bofile)."""
        ti, bi = self._files(ifiles, bifiles)
        xyz, v = containers.average_scalar_files(ti or bi, binary=not ti, dim=int(dim))
        if ofile:
            containers.write_scalar_text(ofile, xyz, v, int(dim))
        if bofile:
            containers.write_scalar_binary(bofile, xyz, v, int(dim))
        return self._map_out(xyz, v)

    def average_power_density(self, ifiles=None, bifiles=None, ofile="", bofile="", dim=2):
        """OSCARSSR_AverageT3DScalars, as average_flux."""
        return self.average_flux(ifiles, bifiles, ofile, bofile, dim)

    # ---- engine-level accumulators (OSCARSSR_Python.cc:3330-3470) ----------------
    def add_to_spectrum(self, spectrum, weight=1.0):
        if isinstance(spectrum, tuple):  # (energies, flux) arrays of the numpy output format
            en = np.array(spectrum[0], dtype=np.float64)
            s = np.array(spectrum[1], dtype=np.float64) * weight
        else:
            en = np.array([p[0] for p in spectrum], dtype=np.float64)
            s = np.array([p[1] for p in spectrum], dtype=np.float64) * weight
        if self._spectrum is None:
            self._spectrum = [en, s, np.zeros_like(s)]
        elif len(s) == len(self._spectrum[1]):
            _kahan(self._spectrum[1], self._spectrum[2], s)
        else:
            raise ValueError("spectra dimensions do not match")

    def get_spectrum(self):
        e, f, _ = self._spectrum if self._spectrum is not None else (np.zeros(0), np.zeros(0), None)
        return self._spec_out(e, f)

    def add_to_flux(self, flux, weight=1.0):
        self._flux = _add_container(self._flux, flux, weight)

    def get_flux(self):
        return self._map_out(*self._flux[:2]) if self._flux is not None else self._map_out(np.zeros((0, 3)), np.zeros(0))

    def add_to_power_density(self, power_density, weight=1.0):
        self._power_density = _add_container(self._power_density, power_density, weight)

    def get_power_density(self):
        if self._power_density is None:
            return self._map_out(np.zeros((0, 3)), np.zeros(0))
        return self._map_out(*self._power_density[:2])

    # ---- helpers -------------------------------------------------------------------
    def _check_common(self, nparticles, gpu, nthreads):
        _check_route(gpu, nthreads, nparticles)

    def _rectangle(self, npoints, plane, width, rotations, translation, x0x1x2, normal):
        if npoints is None or len(npoints) != 2:
            raise ValueError("'npoints' must be [int, int]")
        nx1, nx2 = int(npoints[0]), int(npoints[1])
        if nx1 <= 0 or nx2 <= 0:
            raise ValueError("an entry in 'npoints' is <= 0")
        rot = _opt3(rotations, "rotations")
        trans = _opt3(translation, "translation")
        w1 = w2 = 0.0
        if width is not None and len(width) == 2:
            w1, w2 = float(width[0]), float(width[1])
        surf = None
        if plane and w1 > 0 and w2 > 0:
            surf = _sr.Surface.rectangle(plane, nx1, nx2, w1, w2, rot, trans, int(normal))
        if x0x1x2 is not None and len(x0x1x2) != 0:
            if len(x0x1x2) != 3:
                raise ValueError("'x0x1x2' must have 3 XYZ points defined correctly")
            pts = []
            for p in x0x1x2:
                try:
                    q = _vec3(p, "x0x1x2")
                except ValueError:
                    raise ValueError("Incorrect format in 'x0x1x2'")
                q = _rotate_xyz(q, rot)
                pts.append([q[i] + trans[i] for i in range(3)])
            surf = _sr.Surface.rectangle3(nx1, nx2, pts[0], pts[1], pts[2], int(normal))
        if surf is None:
            surf = _sr.Surface.points(np.zeros((0, 3)), np.zeros((0, 3)))
        return surf

    def _points(self, points, rotations, translation, normal, flip):
        rot = _opt3(rotations, "rotations")
        trans = _opt3(translation, "translation")
        xs, ns = [], []
        for p in points:
            if len(p) != 2:
                raise ValueError("Incorrect format in 'points'")
            if len(p[0]) != 3 or len(p[1]) != 3:
                raise ValueError("Incorrect format in 'points': Point or Normal does not have 3 elements")
            x = _rotate_xyz([float(v) for v in p[0]], rot)
            n = _rotate_xyz([float(v) for v in p[1]], rot)
            if flip and normal == -1:
                n = [-v for v in n]
            xs.append([x[i] + trans[i] for i in range(3)])
            ns.append(n)
        xs = np.array(xs, dtype=np.float64).reshape(-1, 3)
        ns = np.array(ns, dtype=np.float64).reshape(-1, 3)
        return _sr.Surface.points(xs, ns)

    def _out(self, xyz, v, ofile, dim):
        if ofile:
            containers.write_scalar_text(ofile, xyz, v, dim)
        return self._map_out(xyz, v)


def _kahan(s, c, v):
    for i in range(len(v)):
        y = v[i] - c[i]
        t = s[i] + y
        c[i] = (t - s[i]) - y
        s[i] = t


def _add_container(state, data, weight):
    if isinstance(data, tuple):  # (points, values) arrays of the numpy output format
        xyz = np.array(data[0], dtype=np.float64).reshape(-1, 3)
        v = np.array(data[1], dtype=np.float64) * weight
    else:
        xyz = np.array([p[0] for p in data], dtype=np.float64).reshape(-1, 3)
        v = np.array([p[1] for p in data], dtype=np.float64) * weight
    if state is None:
        return [xyz, v, np.zeros_like(v)]
    if len(v) != len(state[1]):
        raise ValueError("spectra dimensions do not match")
    _kahan(state[1], state[2], v)
    return state


if __name__ == "__main__":  # pragma: no cover
    sys.exit("import srsoriginal_amd.sr")
