// pybind11 binding of the host engine: srsoriginal_amd.lib._sr
//
// The user-facing `sr` type (srsoriginal_amd/sr.py) mirrors the reference's
// CPython type oscars.sr.sr (OSCARSSR_Python.cc:3670-3747) and calls into this
// module.  Arrays cross as NumPy float64; C++ exceptions map to ValueError like
// the reference's (std::length_error / out_of_range / invalid_argument ->
// PyExc_ValueError, e.g. OSCARSSR_Python.cc:2329-2338).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "srs_engine.h"

namespace py = pybind11;
using srs::Vec3;

namespace {

Vec3 v3(std::vector<double> const& v) {
  if (v.size() != 3) throw std::length_error("number of elements not 3");
  return Vec3(v[0], v[1], v[2]);
}

py::array_t<double> traj_array(srs::Trajectory const& T) {
  size_t const n = T.size();
  py::array_t<double> a({(py::ssize_t)9, (py::ssize_t)n});
  auto m = a.mutable_unchecked<2>();
  std::vector<double> const* src[9] = {&T.x, &T.y, &T.z, &T.bx, &T.by, &T.bz, &T.ax, &T.ay, &T.az};
  for (int k = 0; k < 9; ++k)
    for (size_t i = 0; i < n; ++i) m(k, i) = (*src[k])[i];
  return a;
}

py::tuple container_arrays(srs::ScalarContainer const& C) {
  size_t const n = C.size();
  py::array_t<double> x({(py::ssize_t)n, (py::ssize_t)3});
  py::array_t<double> v((py::ssize_t)n);
  auto mx = x.mutable_unchecked<2>();
  auto mv = v.mutable_unchecked<1>();
  for (size_t i = 0; i < n; ++i) {
    mx(i, 0) = C.xyz[3 * i];
    mx(i, 1) = C.xyz[3 * i + 1];
    mx(i, 2) = C.xyz[3 * i + 2];
    mv(i) = C.v[i];
  }
  return py::make_tuple(x, v);
}

srs::Precision prec_of(int p) {
  if (p != 0 && p != 1) throw std::invalid_argument("'precision' must be 'double' or 'mixed'");
  return static_cast<srs::Precision>(p);
}

}  // namespace

// A field element into the magnetic (target 0) or electric (target 1) container.
static void add_to(srs::OSCARSSR& e, srs::IdealUndulator const& u, int target) {
  if (target == 0) e.AddMagneticField(u);
  else if (target == 1) e.AddElectricField(u);
  else throw std::invalid_argument("field target must be 0 (magnetic) or 1 (electric)");
}

PYBIND11_MODULE(_sr, m) {
  m.doc() = "srsoriginal_amd host engine (OSCARSSR-compatible) over the gfx950 kernels";
  py::register_exception_translator([](std::exception_ptr p) {
    try {
      if (p) std::rethrow_exception(p);
    } catch (std::length_error const& e) {
      PyErr_SetString(PyExc_ValueError, e.what());
    } catch (std::out_of_range const& e) {
      PyErr_SetString(PyExc_ValueError, e.what());
    } catch (std::invalid_argument const& e) {
      PyErr_SetString(PyExc_ValueError, e.what());
    }
  });

  py::class_<srs::SurfacePoints>(m, "Surface")
      .def_static("rectangle",
                  [](std::string const& plane, int nx1, int nx2, double w1, double w2, std::vector<double> rot,
                     std::vector<double> trans, int normal) {
                    return srs::SurfacePoints::rectangle(plane, nx1, nx2, w1, w2, v3(rot), v3(trans), normal);
                  })
      .def_static("rectangle3",
                  [](int nx1, int nx2, std::vector<double> X0, std::vector<double> X1, std::vector<double> X2,
                     int normal) { return srs::SurfacePoints::rectangle3(nx1, nx2, v3(X0), v3(X1), v3(X2), normal); })
      .def_static("points",
                  [](py::array_t<double, py::array::c_style | py::array::forcecast> xyz,
                     py::array_t<double, py::array::c_style | py::array::forcecast> nrm) {
                    srs::SurfacePoints S;
                    auto a = xyz.unchecked<2>();
                    auto b = nrm.unchecked<2>();
                    for (py::ssize_t i = 0; i < a.shape(0); ++i)
                      S.add(Vec3(a(i, 0), a(i, 1), a(i, 2)), Vec3(b(i, 0), b(i, 1), b(i, 2)));
                    return S;
                  })
      .def_property_readonly("npoints", &srs::SurfacePoints::size)
      .def_property_readonly("xyz", [](srs::SurfacePoints const& S) {
        py::array_t<double> a({(py::ssize_t)S.size(), (py::ssize_t)3});
        std::copy(S.xyz.begin(), S.xyz.end(), a.mutable_data());
        return a;
      })
      .def_property_readonly("normals", [](srs::SurfacePoints const& S) {
        py::array_t<double> a({(py::ssize_t)S.size(), (py::ssize_t)3});
        std::copy(S.normals.begin(), S.normals.end(), a.mutable_data());
        return a;
      });

  py::class_<srs::OSCARSSR>(m, "Engine")
      .def(py::init<>())
      .def("add_undulator",
           [](srs::OSCARSSR& e, std::vector<double> field, std::vector<double> period, int nperiods, double phase,
              double taper, std::vector<double> center, std::vector<double> pole_deltas, int target) {
             srs::IdealUndulator u;
             u.field = v3(field);
             u.period = v3(period);
             u.nperiods = nperiods;
             u.phase = phase;
             u.taper = taper;
             u.center = v3(center);
             if (!pole_deltas.empty()) {
               if ((int)pole_deltas.size() != 2 * nperiods)
                 throw std::invalid_argument("'pole_scales' must have 2*nperiods entries");
               u.pole_delta = pole_deltas;
             }
             add_to(e, u, target);
           })
      .def("add_field",
           [](srs::OSCARSSR& e, int kind, std::vector<double> field, std::vector<double> width,
              std::vector<double> rot, std::vector<double> center, int target) {
             // TField3D_Gaussian / TField3D_UniformBox (field given already rotated)
             srs::IdealUndulator u;
             if (kind != 1 && kind != 2) throw std::invalid_argument("unknown field kind");
             u.kind = kind == 1 ? srs::FieldKind::Gaussian : srs::FieldKind::UniformBox;
             u.field = v3(field);
             u.width = v3(width);
             u.rot = v3(rot);
             u.center = v3(center);
             add_to(e, u, target);
           })
      .def("add_bfield_file",
           [](srs::OSCARSSR& e, std::string const& file, std::string const& format, std::vector<double> rot,
              std::vector<double> trans, std::vector<double> scale, int target) {
             // OSCARSSR::AddMagneticField(file, format, ...), OSCARSSR.cc:68-90
             srs::IdealUndulator u;
             u.kind = srs::FieldKind::Grid;
             u.rot = v3(rot);
             u.center = v3(trans);
             u.grid = std::make_shared<srs::FieldGrid>(srs::read_grid_file(file, format, u.rot, scale));
             add_to(e, u, target);
           })
      .def("add_bfield_interpolated",
           [](srs::OSCARSSR& e, std::vector<std::pair<double, std::string>> mapping, std::string const& format,
              double parameter, std::vector<double> rot, std::vector<double> trans, std::vector<double> scale) {
             // OSCARSSR::AddMagneticFieldInterpolated, OSCARSSR.cc:94-114
             srs::IdealUndulator u;
             u.kind = srs::FieldKind::Grid;
             u.rot = v3(rot);
             u.center = v3(trans);
             u.grid = std::make_shared<srs::FieldGrid>(
                 srs::read_grid_interpolated(mapping, format, parameter, u.rot, scale));
             e.AddMagneticField(u);
           })
      .def("clear_bfields", &srs::OSCARSSR::ClearMagneticFields)
      .def("n_bfields", [](srs::OSCARSSR const& e) { return e.Fields().size(); })
      .def("get_bfield",
           [](srs::OSCARSSR const& e, std::vector<double> x) {
             Vec3 const b = e.GetB(v3(x));
             return std::vector<double>{b.x, b.y, b.z};
           })
      .def("clear_efields", &srs::OSCARSSR::ClearElectricFields)
      .def("n_efields", [](srs::OSCARSSR const& e) { return e.EFields().size(); })
      .def("get_efield",
           [](srs::OSCARSSR const& e, std::vector<double> x) {
             Vec3 const f = e.GetE(v3(x));
             return std::vector<double>{f.x, f.y, f.z};
           })
      .def("add_particle_beam",
           [](srs::OSCARSSR& e, std::string const& type, std::string const& name, std::vector<double> x0,
              std::vector<double> d0, double energy, double t0, double current, double weight) {
             e.AddParticleBeam(type, name, v3(x0), v3(d0), energy, t0, current, weight);
           })
      .def("set_sigma",
           [](srs::OSCARSSR& e, std::string const& name, std::vector<double> h, std::vector<double> su,
              std::vector<double> sup, std::vector<double> at, double se) {
             double a[2] = {su.at(0), su.at(1)}, b[2] = {sup.at(0), sup.at(1)};
             e.GetParticleBeam(name).set_sigma(v3(h), a, b, v3(at), se);
           })
      .def("beam_params",
           [](srs::OSCARSSR& e, std::string const& name) {
             // the beam's definition, for the device sampler (srs_beam)
             auto const& b = e.GetParticleBeam(name);
             py::dict d;
             d["e0"] = b.e0;
             d["sigma_e"] = b.sigma_e;
             d["mass"] = b.proto.m;
             d["charge"] = b.proto.q;
             d["current"] = b.proto.current;
             d["x0"] = std::vector<double>{b.x0.x, b.x0.y, b.x0.z};
             d["u0"] = std::vector<double>{b.u0.x, b.u0.y, b.u0.z};
             d["hdir"] = std::vector<double>{b.hdir.x, b.hdir.y, b.hdir.z};
             d["vdir"] = std::vector<double>{b.vdir.x, b.vdir.y, b.vdir.z};
             d["sigma_at"] = std::vector<double>{b.sigma_at.x, b.sigma_at.y, b.sigma_at.z};
             d["sigma_u"] = std::vector<double>{b.sigma_u[0], b.sigma_u[1]};
             d["sigma_up"] = std::vector<double>{b.sigma_up[0], b.sigma_up[1]};
             d["t0"] = b.t0;
             d["weight"] = b.weight;
             return d;
           })
      .def("beam_names",
           [](srs::OSCARSSR const& e) { return e.BeamNames(); })
      .def("clear_particle_beams", &srs::OSCARSSR::ClearParticleBeams)
      .def("n_beams", &srs::OSCARSSR::GetNParticleBeams)
      .def("set_new_particle", [](srs::OSCARSSR& e) { e.SetNewParticle(); })
      .def("set_new_particle_from", [](srs::OSCARSSR& e, std::string const& b, std::string const& w) { e.SetNewParticle(b, w); })
      .def("particle_x0", [](srs::OSCARSSR const& e) { auto v = e.GetCurrentParticle().x0; return std::vector<double>{v.x, v.y, v.z}; })
      .def("particle_beta0", [](srs::OSCARSSR const& e) { auto v = e.GetCurrentParticle().beta0; return std::vector<double>{v.x, v.y, v.z}; })
      .def("particle_e0", [](srs::OSCARSSR const& e) { return e.GetCurrentParticle().e0_gev(); })
      .def("particle_gamma", [](srs::OSCARSSR const& e) { return e.GetCurrentParticle().gamma; })
      .def("particle_charge_current", [](srs::OSCARSSR const& e) {
        return std::vector<double>{e.GetCurrentParticle().q, e.GetCurrentParticle().current};
      })
      .def("particle_type", [](srs::OSCARSSR const& e) { return e.GetCurrentParticle().type; })
      .def("particle_state",
           [](srs::OSCARSSR const& e) {
             auto const& p = e.GetCurrentParticle();
             py::dict d;
             d["x0"] = std::vector<double>{p.x0.x, p.x0.y, p.x0.z};
             d["beta0"] = std::vector<double>{p.beta0.x, p.beta0.y, p.beta0.z};
             d["qmg"] = p.qmg;
             d["t0"] = p.t0;
             d["gamma"] = p.gamma;
             d["charge"] = p.q;
             d["current"] = p.current;
             return d;
           })
      .def("undulators",
           [](srs::OSCARSSR const& e) {
             // every field element, the magnetic ones first (each list in insertion order)
             py::list out;
             for (int target = 0; target < 2; ++target)
             for (auto const& u : (target == 0 ? e.Fields() : e.EFields()).fields()) {
               py::dict d;
               d["target"] = target;
               d["field"] = std::vector<double>{u.field.x, u.field.y, u.field.z};
               d["period"] = std::vector<double>{u.period.x, u.period.y, u.period.z};
               d["nperiods"] = u.nperiods;
               d["phase"] = u.phase;
               d["taper"] = u.taper;
               d["center"] = std::vector<double>{u.center.x, u.center.y, u.center.z};
               d["pole_delta"] = u.pole_delta.empty() ? py::object(py::none()) : py::object(py::cast(u.pole_delta));
               d["kind"] = (int)u.kind;
               d["width"] = std::vector<double>{u.width.x, u.width.y, u.width.z};
               d["rotations"] = std::vector<double>{u.rot.x, u.rot.y, u.rot.z};
               if (u.grid) {
                 d["grid"] = u.grid->data;
                 d["grid_n"] = std::vector<int>{u.grid->n[0], u.grid->n[1], u.grid->n[2]};
                 d["grid_start"] = std::vector<double>(u.grid->start, u.grid->start + 3);
                 d["grid_step"] = std::vector<double>(u.grid->step, u.grid->step + 3);
                 d["grid_stop"] = std::vector<double>(u.grid->stop, u.grid->stop + 3);
               }
               out.append(d);
             }
             return out;
           })
      .def("set_seed", &srs::OSCARSSR::SetSeed)
      .def("rand", [](srs::OSCARSSR const&) { return srs::Random::global().uniform(); })
      .def("norm", [](srs::OSCARSSR const&) { return srs::Random::global().normal(); })
      .def("set_ctstartstop", &srs::OSCARSSR::SetCTStartStop)
      .def("set_npoints_trajectory", &srs::OSCARSSR::SetNPointsTrajectory)
      .def("set_npoints_per_meter", &srs::OSCARSSR::SetNPointsPerMeter)
      .def("get_npoints_trajectory", &srs::OSCARSSR::GetNPointsTrajectory)
      .def("get_ctstart", &srs::OSCARSSR::GetCTStart)
      .def("get_ctstop", &srs::OSCARSSR::GetCTStop)
      .def("set_nthreads_global", &srs::OSCARSSR::SetNThreadsGlobal)
      .def("set_gpu_global", &srs::OSCARSSR::SetUseGPUGlobal)
      .def("check_gpu", &srs::OSCARSSR::CheckGPU)
      .def("calculate_trajectory",
           [](srs::OSCARSSR& e, int gpu) {
             if (gpu) e.CalculateTrajectoryGPU();
             else e.CalculateTrajectory();
             return py::make_tuple(traj_array(e.GetTrajectory()), e.GetTrajectory().dt);
           },
           py::arg("gpu") = 0)
      .def("get_trajectory", [](srs::OSCARSSR const& e) { return py::make_tuple(traj_array(e.GetTrajectory()), e.GetTrajectory().dt); })
      .def("calculate_spectrum",
           [](srs::OSCARSSR& e, std::vector<double> obs, std::vector<double> points, size_t n, double e0, double e1,
              int nparticles, int nthreads, int gpu, int precision, bool batched) {
             srs::SpectrumContainer S;
             if (points.empty()) S.init(n, e0, e1);
             else S.init(points);
             e.CalculateSpectrum(v3(obs), S, nparticles, nthreads, gpu, prec_of(precision), batched);
             return py::make_tuple(py::array_t<double>(S.e.size(), S.e.data()), py::array_t<double>(S.f.size(), S.f.data()));
           },
           py::arg("obs"), py::arg("points"), py::arg("n"), py::arg("e0"), py::arg("e1"), py::arg("nparticles"),
           py::arg("nthreads"), py::arg("gpu"), py::arg("precision"), py::arg("batched") = false)
      .def("calculate_flux",
           [](srs::OSCARSSR& e, srs::SurfacePoints const& S, double energy, int nparticles, int nthreads, int gpu,
              int dim, int precision, bool batched) {
             srs::ScalarContainer C;
             e.CalculateFlux(S, energy, C, nparticles, nthreads, gpu, dim, prec_of(precision), batched);
             return container_arrays(C);
           },
           py::arg("surface"), py::arg("energy"), py::arg("nparticles"), py::arg("nthreads"), py::arg("gpu"),
           py::arg("dim"), py::arg("precision"), py::arg("batched") = false)
      .def("calculate_power_density",
           [](srs::OSCARSSR& e, srs::SurfacePoints const& S, int dim, bool directional, int nparticles, int nthreads,
              int gpu, int precision, bool batched) {
             srs::ScalarContainer C;
             e.CalculatePowerDensity(S, C, dim, directional, nparticles, nthreads, gpu, prec_of(precision), batched);
             return container_arrays(C);
           },
           py::arg("surface"), py::arg("dim"), py::arg("directional"), py::arg("nparticles"), py::arg("nthreads"),
           py::arg("gpu"), py::arg("precision"), py::arg("batched") = false)
      .def("calculate_flux1",
           [](srs::OSCARSSR& e, srs::SurfacePoints const& S, double energy, std::string const& ofile) {
             // container points: (x1, x2, 0) of a rectangle, as the flux routes add them for
             // dim = 2; the points themselves for a point list (whose x1 = x2 = 0)
             srs::ScalarContainer C;
             for (size_t i = 0; i < S.size(); ++i)
               C.add_point(S.has_x12 ? srs::Vec3(S.x1[i], S.x2[i], 0)
                                     : srs::Vec3(S.xyz[3 * i], S.xyz[3 * i + 1], S.xyz[3 * i + 2]),
                           0);
             e.CalculateFlux1(S, energy, C, ofile);
             return container_arrays(C);
           },
           py::arg("surface"), py::arg("energy"), py::arg("ofile") = "")
      .def("calculate_total_power", &srs::OSCARSSR::CalculateTotalPower)
      .def("calculate_total_power_device", &srs::OSCARSSR::CalculateTotalPowerDevice)
      .def("calculate_efield_vs_time", [](srs::OSCARSSR& e, std::vector<double> obs) {
        std::vector<double> ef, t;
        e.CalculateElectricFieldTimeDomain(v3(obs), ef, t);
        py::array_t<double> E({(py::ssize_t)t.size(), (py::ssize_t)3});
        std::copy(ef.begin(), ef.end(), E.mutable_data());
        return py::make_tuple(E, py::array_t<double>(t.size(), t.data()));
      });
}
