// Batched multi-particle pipeline (internal to libsrs_hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "srs_kernels.h"

namespace srs {

// Ideal undulator on the device (TField3D_IdealUndulator); pole_off >= 0 selects
// the per-pole scaled form with deltas poles[pole_off .. pole_off + 2N).
struct DevUndulator {
  double field[3], period[3], center[3];
  double phase, taper;
  int nperiods;
  int pole_off;
};

// Per-particle state and trajectory grid (dt, nb, nf computed on the host with the
// reference's formulas, OSCARSSR.cc:592-598); off = first point in the packed arrays.
struct BatchParticle {
  double x0[3], beta0[3];
  double qmg, dt;
  int64_t nb, nf, off;
};

// Packed trajectories of all particles (FP64 for the plain form and A/c, FP32
// beta_x, beta_y, 1 - beta_z for the mixed form).
struct BatchTraj {
  double *x, *y, *z, *bx, *by, *bz, *ax, *ay, *az;  // ax..az may be NULL
  float *fbx, *fby, *fomb;
};

struct BatchEnergies {
  const double* omega;  // device arrays of length ne
  const double* cw;     // c / omega
  const double* c0w;    // C0 * omega (times dt per particle gives kw)
  int ne;
  double c2;
};

// Tile shape of the persistent kernel: energies per work item and trajectory points
// per LDS tile (256 observation points per block).  Chosen by the autotuner.
struct BatchConfig {
  int ne_per_item;
  int tile;
};

bool batch_config_valid(BatchConfig const& c);
int batch_max_blocks(BatchConfig const& c, bool mixed);
hipError_t launch_rk4_batch(const DevUndulator* U, int nu, const double* poles, const BatchParticle* P, int np,
                            BatchTraj const& bt, hipStream_t s);
hipError_t launch_zero_counter(unsigned* c, hipStream_t s);
hipError_t launch_batch_flux(BatchConfig const& cfg, bool mixed, BatchTraj const& bt, const BatchParticle* P,
                             const double* obs, int ns, BatchEnergies const& be, unsigned nitems, unsigned* counter,
                             double* flux_pp, int nblocks, hipStream_t s);
hipError_t launch_particle_reduce(const double* flux_pp, int np, int nout, double weight, double* out, hipStream_t s);

}  // namespace srs
