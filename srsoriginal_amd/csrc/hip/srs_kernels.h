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

// Trajectory chunks of a flux launch: chunk c < n1 covers [c c1, (c + 1) c1), chunk c >= n1
// covers c2 points from n1 c1 + (c - n1) c2, the last one clipped at the trajectory's end;
// nch chunks in all.  Blocks are dispatched in chunk order, so tail chunks shorter than
// the rest (c2 < c1) shorten the drain at the end of a launch.
struct ChunkMap {
  int c1, n1, c2, nch;
  __host__ __device__ int64_t begin(int c) const {
    return c < n1 ? (int64_t)c * c1 : (int64_t)n1 * c1 + (int64_t)(c - n1) * c2;
  }
  __host__ __device__ int64_t end(int c, int64_t n) const {
    int64_t const e = begin(c) + (c < n1 ? c1 : c2);
    return e < n ? e : n;
  }
};

// Total power of one trajectory (with accelerations): out[0] = scale * sum of the terms.
hipError_t launch_total_power(const TrajView& t, double scale, double* out, hipStream_t s);

// ---- resident blocks per CU of the integral kernels (for the chunk choice) ----
int flux_fp64_blocks_per_cu(int ne);
int pd_fp64_blocks_per_cu();
int flux1_fp64_blocks_per_cu();
int flux_mixed_blocks_per_cu(int ne);
// One energy: blocks of kFlux2Threads lanes, two observation points per lane.
constexpr int kFlux2Threads = 256;
// Observation points per block of the mixed flux launch for ne energies.
constexpr int flux_mixed_points_per_block(int ne) { return ne == 1 ? 2 * kFlux2Threads : kBlock; }
int pd_mixed_blocks_per_cu(int directional);

// ---- plain FP64 form ------------------------------------------------------
// Flux partial sums: part[((c*ne + e)*6 + k)*ns + j], c = trajectory chunk.
hipError_t launch_flux_fp64(const TrajView& t, const double* obs_soa, int ns, const EnergySet& es,
                            ChunkMap const& cm, double* part, hipStream_t s);
// Near- plus far-field flux (CalculateFlux1) partial sums: part[(c*6 + k)*ns + j], the
// layout of launch_flux_fp64 with one energy (flux_reduce_kernel applies C0 dt and C2).
hipError_t launch_flux1_fp64(const TrajView& t, const double* obs_soa, int ns, double omega, int chunk, double* part,
                             hipStream_t s);
// Power density partial sums: part[c*ns + j].
hipError_t launch_pd_fp64(const TrajView& t, const double* obs_soa, const double* nrm_soa, int ns, int directional,
                          int chunk, double* part, hipStream_t s);

// ---- mixed form (FP64 phase, FP32 trig / amplitudes / tile sums) -----------
// Partial sums as launch_flux_fp64's, rounded to FP32 (each is a Kahan-compensated FP32
// sum, or an FP64 one for chunks the FP64 terms integrate): half the bytes to write and
// to fold.
hipError_t launch_flux_mixed(const TrajView& t, const float* trajf, const double* obs_soa, int ns, const EnergySet& es,
                             ChunkMap const& cm, float* part, hipStream_t s);
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

// ---- spectrum at one observer (srs_spectrum.hip), FP64 or mixed -----------
constexpr int kSpecRun = 8;  // energies per lane on a uniform grid
// Whether every run of kSpecRun consecutive energies is a uniform grid up to rounding (then
// a lane steps across its run by recurrence; otherwise one energy per lane).
bool spectrum_runs_uniform(const double* omega, int64_t ne);
struct SpectrumPlan {
  int run;     // energies per lane (8 or 1)
  int groups;  // blocks along the energies (64 lanes x run each)
  int nchunk;  // blocks along the trajectory
  int sub;     // trajectory points per wave (8 waves per block)
};
SpectrumPlan plan_spectrum(int64_t n, int64_t ne, bool uniform, int mixed, int cus);
// Device tables of a spectrum launch: per energy omega and c/omega; per run of
// kSpecRun energies the step rdw (the mean spacing of the run's block of 64 runs, the
// same for all of them) and the |tau| bound rtl of the phase correction, from
// spectrum_run_tables (host) for the form (mixed or FP64) that is launched.
struct SpecTables {
  const double* omega;
  const double* cw;
  const double* rdw;
  const double* rtl;
};
void spectrum_run_tables(const double* omega, int64_t ne, int mixed, double* rdw, double* rtl);
// Complex field sums of chunk c: part[(c*6 + q)*ne + e], nchunk x 6 x ne doubles.  With
// runs of 8 the kernel writes 1 to *range (host-visible memory, cleared by the caller)
// when the trajectory leaves the range of its phase correction; the caller then plans
// again with uniform = false.
hipError_t launch_spectrum(const TrajView& t, const double obs[3], const SpecTables& tab, int ne, const SpectrumPlan& p,
                           int mixed, double* part, int* range, hipStream_t s);
// out[e] = c2 sum |i kw[e] S|^2 over the chunks' sums S, added in chunk order.
hipError_t launch_spectrum_fold(const double* part, int nchunk, int ne, const double* kw, double c2, double* out,
                                hipStream_t s);

// ---- reductions over trajectory chunks (fixed order, deterministic) --------
hipError_t launch_flux_reduce(const double* part, int nchunk, int ns, const EnergySet& es, double c2, double* out,
                              hipStream_t s);
// The same over FP32 partials (launch_flux_mixed), added in FP64.
hipError_t launch_flux_reduce(const float* part, int nchunk, int ns, const EnergySet& es, double c2, double* out,
                              hipStream_t s);
// Time-domain electric field: e[3*i + k], arrival time tm[i] for every trajectory point.
hipError_t launch_efield_time(const TrajView& t, const double obs[3], double c0, double* e, double* tm, hipStream_t s);
hipError_t launch_pd_reduce(const double* part, int nchunk, int ns, double scale, int directional, double* out,
                            hipStream_t s);

}  // namespace srs
