// Host engine of srsoriginal_amd: an OSCARSSR-compatible C++ class for the
// radiation hot path (reference include/OSCARSSR.h:31-191, "THIS is the c++ API").
//
// Host responsibilities (as in the reference): field model, particle beams and
// sampling, the FP64 RK4 trajectory, observation surfaces and the output
// containers with their compensated accumulation.  Every radiation integral runs
// on the GPU through the C ABI of include/srs_hip.h; there is no CPU integral path
// in the product.
#pragma once
#include <cstdint>
#include <map>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "srs_hip.h"

namespace srs {

struct Vec3 {
  double x = 0, y = 0, z = 0;
  Vec3() = default;
  Vec3(double a, double b, double c) : x(a), y(b), z(c) {}
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

// ---- fields ---------------------------------------------------------------

// Ideal sine undulator with 0.25/0.75 end terminations
// (TField3D_IdealUndulator.cc:50-77, :106-139).  Optional per-pole scales
// (1 + delta_p) for the 2N interior half-period poles.
// A magnetic field element: ideal undulator (TField3D_IdealUndulator), 3-D Gaussian
// (TField3D_Gaussian) or uniform box (TField3D_UniformBox).
enum class FieldKind { Undulator = 0, Gaussian = 1, UniformBox = 2, Grid = 3 };

// A field map on a regular grid (TField3D_Grid): field vectors x-major (index
// (ix*ny + iy)*nz + iz), stored already rotated; linear interpolation along the axes
// that have more than one point, zero outside (TField3D_Grid.cc:124-250).
struct FieldGrid {
  int n[3] = {1, 1, 1};
  double start[3] = {0, 0, 0}, step[3] = {0, 0, 0}, stop[3] = {0, 0, 0};
  std::vector<double> data;  // 3 per point
  Vec3 eval(Vec3 const& X) const;  // X already rotated and translated
};

// Readers of the reference's field-map formats (TField3D_Grid.cc:318-1146) and the
// parameter interpolation between maps (:1150-1445).  rot is applied to every field
// vector as it is read (TVector3D::RotateSelfXYZ); scale as the reference applies it.
FieldGrid read_grid_file(std::string const& file, std::string const& format, Vec3 const& rot,
                         std::vector<double> const& scale, char comment = '#');
FieldGrid read_grid_interpolated(std::vector<std::pair<double, std::string>> const& mapping,
                                 std::string const& format, double parameter, Vec3 const& rot,
                                 std::vector<double> const& scale);

struct IdealUndulator {
  FieldKind kind = FieldKind::Undulator;
  Vec3 field, period, center;
  int nperiods = 0;
  double phase = 0, taper = 0;
  std::vector<double> pole_delta;  // empty = ideal
  Vec3 width;                      // Gaussian sigma / box full width (<= 0: unbounded axis)
  Vec3 rot;                        // Gaussian / box / grid: rotation of the position, X then Y then Z
  std::shared_ptr<const FieldGrid> grid;  // kind Grid (center = translation)
  Vec3 eval(Vec3 const& X) const;
};

class FieldContainer {
 public:
  void add(IdealUndulator const& u) { fields_.push_back(u); }
  void clear() { fields_.clear(); }
  size_t size() const { return fields_.size(); }
  Vec3 eval(Vec3 const& X) const;  // TFieldContainer::GetF, TFieldContainer.cc:107-117
  std::vector<IdealUndulator> const& fields() const { return fields_; }

 private:
  std::vector<IdealUndulator> fields_;
};

// ---- particles ------------------------------------------------------------

struct Trajectory {
  double dt = 0;
  std::vector<double> x, y, z, bx, by, bz, ax, ay, az;
  size_t size() const { return x.size(); }
  void clear();
};

struct Particle {
  std::string type;
  double q = 0, m = 0, current = 0;
  Vec3 x0, beta0;
  double t0 = 0, gamma = 0, qmg = 0;
  Trajectory traj;
  void set_type(std::string const& t);      // TParticleA::SetParticleType, TParticleA.cc:59-103
  void set_current(double c);               // TParticleA.cc:254-266
  void set_initial(Vec3 const& X, Vec3 const& B, double T);  // TParticleA.cc:330-382
  double e0_gev() const;
};

struct ParticleBeam {
  Particle proto;                 // type, q, m, current
  Vec3 x0, u0;
  double e0 = 0, t0 = 0, weight = 1;
  Vec3 hdir, vdir, sigma_at;
  double sigma_u[2] = {0, 0}, sigma_up[2] = {0, 0}, sigma_e = 0;
  Particle ideal() const;          // TParticleBeam::GetNewParticle("ideal"), :200-222
  Particle random() const;         // TParticleBeam::GetNewParticle(), :228-265
  void set_sigma(Vec3 const& h, const double su[2], const double sup[2], Vec3 const& at, double se);
};

// Process-global generator of the reference (TRandomA.cc:12-64).
class Random {
 public:
  static Random& global();
  void seed(int s);
  double normal();
  double uniform();

 private:
  Random();
  std::unique_ptr<std::mt19937> mt_;
  std::normal_distribution<double> normal_{0, 1};
  std::uniform_real_distribution<double> uniform_{0, 1};
};

// ---- surfaces and containers ---------------------------------------------

struct SurfacePoints {
  // points and normals (AoS xyz), plus the in-plane coordinates used for dim=2 output
  std::vector<double> xyz, normals, x1, x2;
  size_t size() const { return xyz.size() / 3; }
  bool has_x12 = false;
  static SurfacePoints rectangle(std::string const& plane, int nx1, int nx2, double w1, double w2, Vec3 const& rot,
                                 Vec3 const& trans, int normal);
  static SurfacePoints rectangle3(int nx1, int nx2, Vec3 const& X0, Vec3 const& X1, Vec3 const& X2, int normal);
  void add(Vec3 const& X, Vec3 const& N);
};

struct SpectrumContainer {
  std::vector<double> e, f, comp;
  void init(size_t n, double e0, double e1);       // TSpectrumContainer.cc:55-86
  void init(std::vector<double> const& pts);       // :91-106
  void add_to(size_t i, double v);                 // :159-175
};

struct ScalarContainer {
  std::vector<double> xyz, v, comp;
  size_t size() const { return v.size(); }
  void add_point(Vec3 const& X, double val);       // T3DScalarContainer.cc:21-26
  void add_to(size_t i, double val);               // :31-47
};

// ---- the engine -------------------------------------------------------------

enum class Precision { FP64 = 0, Mixed = 1 };

class OSCARSSR {
 public:
  OSCARSSR();

  // fields (OSCARSSR.cc:68-299): magnetic [T] and electric [V/m] elements of the same
  // kinds, each summed in insertion order (TFieldContainer::GetF)
  void AddMagneticField(IdealUndulator const& u) { bfield_.add(u); }
  void ClearMagneticFields() { bfield_.clear(); }
  Vec3 GetB(Vec3 const& X) const { return bfield_.eval(X); }
  FieldContainer const& Fields() const { return bfield_; }
  void AddElectricField(IdealUndulator const& u) { efield_.add(u); }
  void ClearElectricFields() { efield_.clear(); }
  Vec3 GetE(Vec3 const& X) const { return efield_.eval(X); }
  FieldContainer const& EFields() const { return efield_; }

  // beams and particles (OSCARSSR.cc:324-404)
  void AddParticleBeam(std::string const& type, std::string const& name, Vec3 const& x0, Vec3 const& d0, double e_gev,
                       double t0, double current, double weight);
  ParticleBeam& GetParticleBeam(std::string const& name);
  size_t GetNParticleBeams() const { return beams_.size(); }
  std::vector<std::string> BeamNames() const {  // in the order the beams were added
    std::vector<std::string> out(beams_.size());
    for (auto const& kv : beam_index_) out[kv.second] = kv.first;
    return out;
  }
  void ClearParticleBeams();
  void SetNewParticle();                                               // random, by weight
  void SetNewParticle(std::string const& beam, std::string const& ideal_or_random);
  Particle const& GetCurrentParticle() const { return particle_; }
  void SetSeed(int s) const { Random::global().seed(s); }

  // trajectory settings (OSCARSSR.cc:423-483)
  void SetNPointsTrajectory(size_t n) { npoints_ = n; }
  void SetNPointsPerMeter(size_t n);
  void SetCTStartStop(double start, double stop);
  size_t GetNPointsTrajectory() const { return npoints_; }
  double GetCTStart() const { return ct_start_; }
  double GetCTStop() const { return ct_stop_; }
  void SetNThreadsGlobal(int n) { nthreads_global_ = n; }
  void SetUseGPUGlobal(int g) { gpu_global_ = g; }
  int CheckGPU() const;

  // host FP64 RK4 trajectory (OSCARSSR.cc:544-673, :741-874)
  void CalculateTrajectory();
  void CalculateTrajectory(Particle& p) const;
  // The same trajectory integrated on the GPU (device RK4, srs_hip_trajectories):
  // agrees with the host integration to ~1e-14 relative (device sine).
  void CalculateTrajectoryGPU();
  void CalculateTrajectoryGPU(Particle& p) const;
  Trajectory const& GetTrajectory() const { return particle_.traj; }

  // radiation (OSCARSSR.cc:933-984, :1603-1677, :2418-2489): the dispatch keeps the
  // reference's container semantics per route (see the .cc); the integrals run on the GPU.
  // batched: with gpu=1, nparticles > 1 and FP64, all particles in one device pipeline
  // (srs_hip_batch_spectrum) instead of one host trajectory + one GPU call each.
  void CalculateSpectrum(Vec3 const& obs, SpectrumContainer& S, int nparticles, int nthreads, int gpu,
                         Precision prec = Precision::FP64, bool batched = false);
  // batched: with gpu=1 and nparticles > 1, draw every particle first (reference order)
  // and integrate them all in one device pipeline (srs_hip_batch_flux) instead of one
  // host trajectory + one GPU call per particle.
  void CalculateFlux(SurfacePoints const& surf, double energy_ev, ScalarContainer& C, int nparticles, int nthreads,
                     int gpu, int dim, Precision prec = Precision::FP64, bool batched = false);
  // batched: the same for the power density (srs_hip_batch_power_density, FP64 or
  // mixed).
  void CalculatePowerDensity(SurfacePoints const& surf, ScalarContainer& C, int dim, bool directional, int nparticles,
                             int nthreads, int gpu, Precision prec = Precision::FP64, bool batched = false);
  double CalculateTotalPower();
  double CalculateTotalPowerDevice();  // the sum on the device (srs_hip_total_power)
  // OSCARSSR::CalculateElectricFieldTimeDomain (OSCARSSR.cc:2627-2690): host RK4, the
  // field on the device; e [3 n], t [n].
  void CalculateElectricFieldTimeDomain(Vec3 const& obs, std::vector<double>& e, std::vector<double>& t);
  // OSCARSSR::CalculateFlux1 (OSCARSSR.cc:2209-2346), the Lienard-Wiechert form of the
  // flux with the stored A/c: P's trajectory recomputed (host RK4), the sum on the
  // device (srs_hip_flux1, FP64), each point's value added to point i of C (the
  // caller adds the points, as the reference's callers of the flux sums do), or, with
  // an ofile, "X1 X2 flux" lines written in std::scientific instead.
  void CalculateFlux1(Particle& P, SurfacePoints const& surf, double energy_ev, ScalarContainer& C,
                      std::string const& ofile = "");
  void CalculateFlux1(SurfacePoints const& surf, double energy_ev, ScalarContainer& C, std::string const& ofile = "");

 private:
  void ensure_particle();
  void device_fields(std::vector<srs_field>& out) const;
  srs_particle device_particle(Particle const& p) const;
  void flux_one(Particle& p, SurfacePoints const& surf, double energy_ev, ScalarContainer& C, double weight,
                Precision prec);
  void pd_one(Particle& p, SurfacePoints const& surf, ScalarContainer& C, bool directional, double weight,
              Precision prec);
  void spectrum_one(Particle& p, Vec3 const& obs, SpectrumContainer& S, double weight, Precision prec);

  bool no_fields() const { return bfield_.fields().empty() && efield_.fields().empty(); }

  FieldContainer bfield_, efield_;
  std::vector<ParticleBeam> beams_;
  std::vector<double> beam_weight_sums_;
  std::map<std::string, size_t> beam_index_;
  Particle particle_;
  double ct_start_ = 0, ct_stop_ = 0;
  size_t npoints_ = 0, npoints_per_meter_ = 10000;
  int nthreads_global_ = 2, gpu_global_ = 0;
};

}  // namespace srs
