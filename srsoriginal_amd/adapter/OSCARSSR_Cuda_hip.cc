// Reference-side binding of the MI355X radiation path: the four functions the
// reference engine calls when it is compiled with its GPU switch (-DCUDA), declared
// in the reference's include/OSCARSSR_Cuda.h:25-28 and called from OSCARSSR.cc:501
// (CheckGPU), :900 (spectrum), :1882 (power density) and :2588 (flux), implemented
// over the plain C ABI of include/srs_hip.h (libsrs_hip.so) instead of
// OSCARSSR_Cuda.cu.  A maintainer compiles this file into the reference and links
// -lsrs_hip; nothing else in the reference changes.
//
// Contract kept from the reference backend (OSCARSSR_Cuda.cu): the caller has computed
// the trajectory and pre-populated the containers; each call adds value * Weight to the
// first N entries through the containers' own Kahan accumulation; no GPU -> the
// reference's "No GPU found" invalid_argument.  Differences, all towards the CPU path:
// the spectrum uses omega = (E Qe)/Hbar (the CPU path's, TOSCARSSR.h:66) and the power
// density honours Directional (the CPU path's skip of N.n <= 0 terms and |.| otherwise,
// OSCARSSR.cc:1549-1573), which OSCARSSR_Cuda.cu ignored; Dimension and OutFileName are
// ignored here as they were there.  The trajectory's point at T0 carries the reference's
// uninitialised dxdt as its A/c (OSCARSSR.cc:604-630, undefined behaviour, read by the
// power density): it is staged as 0, the value of a field-free start and the one the
// MI355X engine stores there.
#include <stdexcept>
#include <string>
#include <vector>

#include "OSCARSSR_Cuda.h"
#include "srs_hip.h"

namespace {

// The trajectory of P as nine SoA arrays (x, beta, A/c) in `soa`, viewed as an srs_traj;
// A/c of the point at T0 (stored with position exactly X0, before the first step) is 0.
srs_traj View(TParticleA& P, std::vector<double>& soa) {
  TParticleTrajectoryPoints& T = P.GetTrajectory();
  TVector3D const& X0 = P.GetX0();
  size_t const n = T.GetNPoints();
  soa.resize(9 * n);
  for (size_t i = 0; i < n; ++i) {
    TVector3D const& X = T.GetX(i);
    TVector3D const& B = T.GetB(i);
    TVector3D const& A = T.GetAoverC(i);
    bool const at_t0 = X.GetX() == X0.GetX() && X.GetY() == X0.GetY() && X.GetZ() == X0.GetZ();
    for (int k = 0; k < 3; ++k) {
      soa[k * n + i] = X[k];
      soa[(3 + k) * n + i] = B[k];
      soa[(6 + k) * n + i] = at_t0 ? 0.0 : A[k];
    }
  }
  double const* d = soa.data();
  srs_traj t;
  t.n = (int64_t)n;
  t.dt = T.GetDeltaT();
  t.x = d;
  t.y = d + n;
  t.z = d + 2 * n;
  t.bx = d + 3 * n;
  t.by = d + 4 * n;
  t.bz = d + 5 * n;
  t.ax = d + 6 * n;
  t.ay = d + 7 * n;
  t.az = d + 8 * n;
  return t;
}

void Check(int rc) {
  if (rc == SRS_ERR_NO_DEVICE) throw std::invalid_argument("No GPU found");
  if (rc != SRS_OK) throw std::invalid_argument(srs_hip_last_error());
}

// Surface points (and normals) as AoS xyz.
void Points(TSurfacePoints const& S, std::vector<double>& xyz, std::vector<double>* nrm) {
  size_t const ns = S.GetNPoints();
  xyz.resize(3 * ns);
  if (nrm) nrm->resize(3 * ns);
  for (size_t i = 0; i < ns; ++i) {
    TSurfacePoint const P = S.GetPoint(i);
    for (int k = 0; k < 3; ++k) {
      xyz[3 * i + k] = P.GetPoint()[k];
      if (nrm) (*nrm)[3 * i + k] = P.GetNormal()[k];
    }
  }
}

}  // namespace

extern "C" int OSCARSSR_Cuda_GetDeviceCount() { return srs_hip_device_count(); }

extern "C" void OSCARSSR_Cuda_CalculateFluxGPU(TParticleA& Particle, TSurfacePoints const& Surface,
                                               double const Energy_eV, T3DScalarContainer& FluxContainer,
                                               int const, double const Weight, std::string const&) {
  std::vector<double> soa, xyz;
  srs_traj const t = View(Particle, soa);
  Points(Surface, xyz, nullptr);
  std::vector<double> out(Surface.GetNPoints());
  double const e = Energy_eV;
  Check(srs_hip_flux(&t, xyz.data(), (int64_t)Surface.GetNPoints(), &e, 1, Particle.GetQ(), Particle.GetCurrent(),
                     SRS_PRECISION_FP64, out.data()));
  for (size_t i = 0; i < out.size(); ++i) FluxContainer.AddToPoint(i, out[i] * Weight);
}

extern "C" void OSCARSSR_Cuda_CalculateSpectrumGPU(TParticleA& Particle, TVector3D const& ObservationPoint,
                                                   TSpectrumContainer& Spectrum, double const Weight) {
  std::vector<double> soa;
  srs_traj const t = View(Particle, soa);
  size_t const ne = Spectrum.GetNPoints();
  std::vector<double> energies(ne), out(ne);
  for (size_t i = 0; i < ne; ++i) energies[i] = Spectrum.GetEnergy(i);
  double const obs[3] = {ObservationPoint.GetX(), ObservationPoint.GetY(), ObservationPoint.GetZ()};
  Check(srs_hip_spectrum(&t, obs, energies.data(), (int64_t)ne, Particle.GetQ(), Particle.GetCurrent(),
                         SRS_PRECISION_FP64, out.data()));
  for (size_t i = 0; i < ne; ++i) Spectrum.AddToFlux(i, out[i] * Weight);
}

extern "C" void OSCARSSR_Cuda_CalculatePowerDensityGPU(TParticleA& Particle, TSurfacePoints const& Surface,
                                                       T3DScalarContainer& PowerDensityContainer, int const,
                                                       bool const Directional, double const Weight,
                                                       std::string const&) {
  std::vector<double> soa, xyz, nrm;
  srs_traj const t = View(Particle, soa);
  Points(Surface, xyz, &nrm);
  std::vector<double> out(Surface.GetNPoints());
  Check(srs_hip_power_density(&t, xyz.data(), nrm.data(), (int64_t)Surface.GetNPoints(), Directional ? 1 : 0,
                              Particle.GetQ(), Particle.GetCurrent(), SRS_PRECISION_FP64, out.data()));
  for (size_t i = 0; i < out.size(); ++i) PowerDensityContainer.AddToPoint(i, out[i] * Weight);
}
