// Host-side launch interface of the radiation kernels (internal to libsrs_hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "srs_device.h"

namespace srs {

constexpr int kMaxFluxEnergies = 4;   // energies handled per flux launch
constexpr int kBlock = 256;           // threads per block (4 wave64)

// Per-energy constants of the spectrum/flux integrand, computed on the host in the
// reference's operation order: omega = (E*Qe)/Hbar, cw = C/omega, kw = (C0*omega)*dt;
// fr = omega/(2 pi) [rev/s] and fw = fr/C [rev/m] for the paraxial mixed form.
struct EnergySet {
  int ne;
  double omega[kMaxFluxEnergies];
  double cw[kMaxFluxEnergies];
  double kw[kMaxFluxEnergies];
  double fr[kMaxFluxEnergies];
  double fw[kMaxFluxEnergies];
};

// Total power of one trajectory (with accelerations): out[0] = scale * sum of the terms.
hipError_t launch_total_power(const TrajView& t, double scale, double* out, hipStream_t s);

// ---- resident blocks per CU of the integral kernels (for the chunk choice) ----
int flux_fp64_blocks_per_cu(int ne);
int spectrum_fp64_blocks_per_cu();
int pd_fp64_blocks_per_cu();
int flux1_fp64_blocks_per_cu();
int flux_mixed_blocks_per_cu(int ne);
int pd_mixed_blocks_per_cu(int directional);

// ---- plain FP64 form ------------------------------------------------------
// Flux partial sums: part[((c*ne + e)*6 + k)*ns + j], c = trajectory chunk.
hipError_t launch_flux_fp64(const TrajView& t, const double* obs_soa, int ns, const EnergySet& es, int chunk,
                            double* part, hipStream_t s);
// Spectrum partial sums: part[(c*12 + k)*ne + e] (12 = A and G sums, re/im, xyz).
hipError_t launch_spectrum_fp64(const TrajView& t, const double obs[3], const double* omega, const double* cw, int ne,
                                int chunk, double* part, hipStream_t s);
// Near- plus far-field flux (CalculateFlux1) partial sums: part[(c*6 + k)*ns + j], the
// layout of launch_flux_fp64 with one energy (flux_reduce_kernel applies C0 dt and C2).
hipError_t launch_flux1_fp64(const TrajView& t, const double* obs_soa, int ns, double omega, int chunk, double* part,
                             hipStream_t s);
// Power density partial sums: part[c*ns + j].
hipError_t launch_pd_fp64(const TrajView& t, const double* obs_soa, const double* nrm_soa, int ns, int directional,
                          int chunk, double* part, hipStream_t s);

// ---- mixed form (FP64 phase, FP32 trig / amplitudes / tile sums) -----------
hipError_t launch_flux_mixed(const TrajView& t, const float* trajf, const double* obs_soa, int ns, const EnergySet& es,
                             int chunk, double* part, hipStream_t s);
// Mixed spectrum partials: complex field sums part[(c*6 + q)*ne + e] of trajectory
// chunk c; the energies form spectrum_mixed_groups(ne) groups of 64 lanes x 8
// consecutive energies, one block per (group, chunk).
int spectrum_mixed_groups(int ne);
hipError_t launch_spectrum_mixed(const TrajView& t, const double obs[3], const double* omega, const double* cw, int ne,
                                 int chunk, int anchor, double* part, hipStream_t s);
hipError_t launch_pd_mixed(const TrajView& t, const float* trajf, const double* obs_soa, const double* nrm_soa, int ns,
                           int directional, int chunk, double* part, hipStream_t s);
// Single-precision copies used by the mixed kernels: trajf[k*n + i] for
// k = 0..5: bx, by, (1 - bz), ax, ay, az (the last three only when t.ax..az are set).
// FP32 staging of the mixed kernels: 3 n floats (bx, by, 1 - bz), or 7 n when with_a
// (then t.ax..az must be set), then one int per kFlagSpan points, nonzero when a point
// of the span does not travel along z; traj_f32_floats(n, with_a) floats in all.
constexpr int kFlagSpan = 256;
size_t traj_f32_floats(int64_t n, int with_a);
hipError_t launch_traj_to_f32(const TrajView& t, int with_a, float* trajf, hipStream_t s);

// ---- reductions over trajectory chunks (fixed order, deterministic) --------
hipError_t launch_flux_reduce(const double* part, int nchunk, int ns, const EnergySet& es, double c2, double* out,
                              hipStream_t s);
// Time-domain electric field: e[3*i + k], arrival time tm[i] for every trajectory point.
hipError_t launch_efield_time(const TrajView& t, const double obs[3], double c0, double* e, double* tm, hipStream_t s);
hipError_t launch_spectrum_reduce(const double* part, int nchunk, int ne, const double* cw, const double* kw, double c2,
                                  double* out, hipStream_t s);
hipError_t launch_field_reduce(const double* part, int nchunk, int npts, const double* kw, double c2, double* out,
                               hipStream_t s);
hipError_t launch_pd_reduce(const double* part, int nchunk, int ns, double scale, int directional, double* out,
                            hipStream_t s);

}  // namespace srs
