// Batched multi-particle pipeline (internal to libsrs_hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "srs_kernels.h"

namespace srs {

// Ideal undulator on the device (TField3D_IdealUndulator).  The derived constants
// are formed on the host with the reference's expressions (TField3D_IdealUndulator.cc
// :50-139): u = period/|period|, lam = |period|, len = lam*(N+2), PS = phase*lam/2pi
// and the region bounds, e.g. lo1 = (-len/2 + PS) + lam.  rlam = RN(1/lam) turns the
// reference's divisions by lam into correctly rounded Markstein quotients.
// pole_off >= 0 selects the per-pole scaled form (center, taper and phase unused:
// center = taper = PS = 0 there) with deltas poles[pole_off .. pole_off + 2N).
// One field element, constants formed on the host.  kind 0: ideal undulator (the
// members up to pole_off); kind 1: Gaussian, kind 2: uniform box, which use field
// (already rotated), center, w (sigma, or half width |w/2|; <= 0: unbounded axis) and
// the cosines / sines of the rotation angles applied to the position.
struct DevUndulator {
  double field[3], u[3], center[3];
  double lam, rlam, taper, ps;
  double lo0, hi0;  // outside: D < lo0 or D > hi0
  double lo1, hi1;  // termination region: D < lo1 or D > hi1
  double lo2, hi2;  // outer half period (x0.25): D < lo2 or D > hi2, else x0.75
  int nperiods;
  int pole_off;
  int kind;
  int target;  // 0 magnetic, 1 electric
  double w[3];
  double rc[3], rs[3];
  // kind 3, field map: 3 doubles per grid point, ((ix*gn[1] + iy)*gn[2] + iz)
  const double* gdata;
  int gn[3];
  double gstart[3], gstep[3], gstop[3];
};

// Per-particle state and trajectory grid (dt, nb, nf computed on the host with the
// reference's formulas, OSCARSSR.cc:592-598); off = first point in the packed arrays;
// pole_base = offset of this particle's pole deltas in the pole array (0 unless each
// particle carries its own field errors).
struct BatchParticle {
  double x0[3], beta0[3];
  double qmg, dt;
  int64_t nb, nf, off;
  int64_t pole_base;
  // the particle's prefactors, from its charge q and current I with the reference's
  // expressions: c0 = q/(4 pi c eps0 sqrt(2 pi)) (OSCARSSR.cc:1191), c2 = 4 pi I/(h |q| mu0 c)
  // 1e-6 1e-3 (:1194), pdq = |q I|/(16 pi^2 eps0 c) (:1563)
  double c0, c2, pdq;
};

// RK4 checkpoints: state [x, vx, y, vy, z, vz] every `every` steps of each leg;
// particle p's forward-leg slot s is state[(p*stride + s)*6], its backward-leg slot s
// is state[(p*stride + nf_slots + s)*6] with nf_slots = ceil(nf / every).
struct Checkpoints {
  double* state;  // NULL: none
  int every;
  int64_t stride;
};

// Packed trajectories of all particles (FP64 for the plain form and A/c, FP32
// beta_x, beta_y, 1 - beta_z for the mixed form) and, per particle, a flag the device
// RK4 sets when a point integrated so far does not travel along z (transverse_beam):
// the mixed kernels take the FP64 terms for such a particle.
struct BatchTraj {
  double *x, *y, *z, *bx, *by, *bz, *ax, *ay, *az;  // ax..az may be NULL
  float *fbx, *fby, *fomb;
  int* tflag;  // [np]; may be NULL (then no particle is flagged)
};

struct BatchEnergies {
  const double* omega;  // device arrays of length ne
  const double* cw;     // c / omega
  const double* c0w;    // C0 * omega (times dt per particle gives kw)
  const double* fr;     // omega / (2 pi)
  const double* fw;     // omega / (2 pi C)
  int ne;
  double c2;
};

// Tile shape of the persistent kernel: energies per work item, trajectory points per
// LDS tile and observation points per lane (1: 256 per item; 2: 512 per item, the mixed
// form at one energy per item).  Chosen by the autotuner.
struct BatchConfig {
  int ne_per_item;
  int tile;
  int ppl = 1;
};

bool batch_config_valid(BatchConfig const& c);
int batch_max_blocks(BatchConfig const& c, bool mixed);
// Steps [lo, hi) of both RK4 legs of every particle.  A segment with lo > 0 resumes
// from state[(p*2 + leg)*12 ..] (x[6] then the lagged derivative d[6]) and every
// segment leaves its end state there (state may be NULL for a single segment).
struct RK4Seg {
  int64_t lo, hi;
  double* state;
};
// Beam of the device sampler (mgev = particle mass in GeV) and its output rows, laid
// out as srs_particle.
struct DevBeam {
  double e0, sigma_e, mgev, q, m, cur;
  double x0[3], u0[3], h[3], v[3], at[3];
  double su[2], sp[2];
  double t0, ct_start, ct_stop;
  int64_t npoints;
  // the batch records of its particles (launch_sample_batch): prefactors c0, c2, pdq
  // of (q, cur) and the time grid dt, gnb, gnf of (t0, ct_start, ct_stop, npoints)
  double c0, c2, pdq, dt;
  int64_t gnb, gnf;
};
struct SampledParticle {
  double x0[3], beta0[3], qmg, t0, ct_start, ct_stop;
  int64_t npoints;
  double charge, current;
};
// beams: nb device records; wsum: their cumulative weights (device)
hipError_t launch_sample_beam(const DevBeam* beams, const double* wsum, int nb, uint64_t seed, uint64_t first,
                              int64_t np, SampledParticle* out, hipStream_t s);
// The same draws as batch records (out[p]: trajectory offset p * stride).
hipError_t launch_sample_batch(const DevBeam* beams, const double* wsum, int nb, uint64_t seed, uint64_t first,
                               int64_t np, int64_t stride, BatchParticle* out, hipStream_t s);

// local: the list holds elements other than magnetic undulators (Gaussian, box, field
// map, or any electric element); force: kForceB (magnetic elements only), kForceE
// (electric only) or kForceEB (both), OSCARSSR::SetDerivativesFunction's choice;
// npoles: length of each particle's pole-delta table at poles + pole_base (0: none),
// staged in LDS by the kernel when at most kPoleLdsMax
constexpr int kForceB = 0, kForceE = 1, kForceEB = 2;
hipError_t launch_rk4_batch(const DevUndulator* U, int nu, bool local, int force, const double* poles, int npoles,
                            const BatchParticle* P,
                            int np, BatchTraj const& bt, Checkpoints const& ck, RK4Seg const& sg, hipStream_t s);
// Pole tables of up to kPoleLdsMax deltas (64 KiB for a block of 64 particles) are read
// from LDS by the RK4 kernels; longer ones from global memory.
constexpr int kPoleLdsMax = 128;
inline int pole_lds_count(int npoles) { return npoles > 0 && npoles <= kPoleLdsMax ? npoles : 0; }
hipError_t launch_zero_counter(unsigned* c, hipStream_t s);
// One time segment of the flux: steps [lo, hi) of both RK4 legs of every particle.
// The six field sums of each (particle, energy, point) accumulate across segments in
// field_acc[((p*ne + e)*6 + q)*ns + j] (q = re/im of x, y, z), in segment order; the
// first segment starts from zero, the final one writes flux_pp.  With a single
// segment (first and final) field_acc may be NULL, or is written when keep_field.
struct FluxSeg {
  int64_t lo, hi;
  int first, final;
  int keep_field;
  double* field_acc;
};
hipError_t launch_batch_flux(BatchConfig const& cfg, bool mixed, BatchTraj const& bt, const BatchParticle* P,
                             const double* obs, int ns, BatchEnergies const& be, unsigned nitems, unsigned* counter,
                             double* flux_pp, FluxSeg const& sg, int nblocks, hipStream_t s);
hipError_t launch_particle_reduce(const double* flux_pp, int np, int nout, double weight, double* out, hipStream_t s);

// ---- adjoint of the flux loss (srs_adjoint.hip) ----------------------------
hipError_t launch_zero(double* a, int64_t n, hipStream_t s);
// loss[r] = sum_j (flux_pp[r*ns + j] - target[j])^2
hipError_t launch_loss(const double* flux_pp, const double* target, int ns, int np, double* loss, hipStream_t s);
// The observation screen of a flux gradient: flat when every point has z == z, with the
// points' x and y ranges (the mixed adjoint's paraxial bounds).
struct AdjScreen {
  int flat;
  double z, xlo, xhi, ylo, yhi;
};
// gy[c*total + point] = dL/dy_c of every trajectory point (y = x, vx, y, vy, z, vz), one energy;
// mixed: FP64 phase (paraxial on a flat screen sc within the bounds), FP32 amplitudes
// and products (bt.fbx, fby, fomb staged by the RK4)
hipError_t launch_flux_adjoint(BatchTraj const& bt, const BatchParticle* P, int np, int64_t nmax, const double* obs,
                               int ns, double omega, double cw, double c0w, double c2, const double* field_pp,
                               const double* flux_pp, const double* target, const double* dloss, double* gy,
                               int64_t total, bool mixed, AdjScreen const& sc, hipStream_t s);
// doubles of record scratch for `steps` recomputed RK4 steps of np particles
size_t adjoint_record_doubles(int64_t steps, int np);
// doubles of per-segment sweep results for `segments` segments
size_t adjoint_sweep_doubles(int64_t segments, int np, int npoles);
// grad[r*npoles + p] += dL/d delta_p (grad zeroed by the caller): both RK4 legs,
// seg_per_pass checkpoint segments recomputed per pass into rec (adjoint_record_doubles(
// min(seg_per_pass * ck.every, longest leg), np)), swept into sweep
// (adjoint_sweep_doubles(seg_per_pass, np, npoles)); lam_buf holds np * 6 doubles.
// All particles share nf, nb.
// ga (may be NULL: none) = dL/dA of every stored acceleration [3][total].
hipError_t launch_rk4_adjoint(const DevUndulator* U, int nu, int poled, const double* poles, const BatchParticle* P,
                              int np, int64_t nf, int64_t nb, Checkpoints const& ck, int64_t seg_per_pass,
                              const double* gy, const double* ga, int64_t total, double* rec, double* sweep,
                              double* lam_buf, double* grad, int npoles, hipStream_t s);
// Power density of np realisations (trajectories with accelerations in bt) on ns
// points.  launch_pd_batch sums the points of RK4 steps [lo, hi) of both legs in nch
// chunks into part (np * nch * ns doubles), writing (add = 0) or adding (add != 0:
// later time segments, queued in order on one stream); launch_pd_batch_reduce adds
// the nch chunks in order: raw[r*ns + j] = summed integrand, pd[r*ns + j] =
// pdq_r dt_r raw / 1e6 (|.| unless directional), pdq_r = |q I|/(16 pi^2 eps0 c) of particle r.
// pd_batch_chunks gives nch for npoints points.
int pd_batch_chunks(int64_t npoints, int ns, int np);
hipError_t launch_pd_batch(BatchTraj const& bt, const BatchParticle* P, int np, const double* obs, const double* nrm,
                           int ns, int directional, int64_t lo, int64_t hi, int nch, int add, double* part,
                           bool mixed, hipStream_t s);
hipError_t launch_pd_batch_reduce(const double* part, int nslot, int ns, int np, const BatchParticle* P,
                                  int directional, double* raw, double* pd, hipStream_t s);
// Spectrum of np particles at one observation point for ne energies: launch_spectrum_batch
// sums the points of RK4 steps [lo, hi) of both legs in nch chunks into part
// (np * nch * 6 * ne doubles; write or add as launch_pd_batch), with runs of `run`
// energies per lane (kSpecRun on a uniform grid with the tables of spectrum_run_tables,
// else 1; *range as launch_spectrum); launch_spectrum_batch_reduce adds the chunks in
// order and applies i (C0_r omega) dt_r and C2_r (particle r's prefactors): out[r*ne + e].
int spectrum_batch_groups(int ne, int run);
hipError_t launch_spectrum_batch(BatchTraj const& bt, const BatchParticle* P, int np, const double obs[3],
                                 SpecTables const& tab, int ne, int run, int64_t lo, int64_t hi, int nch, int add,
                                 double* part, int* range, bool mixed, hipStream_t s);
hipError_t launch_spectrum_batch_reduce(const double* part, int nch, int ne, int np, const BatchParticle* P,
                                        const double* omega, double* out, hipStream_t s);
// dL/dy (gy [6][total]) and dL/dA (ga [3][total]) of L_r = sum_j (pd_rj - target_j)^2
// scaled by dloss[r].
hipError_t launch_pd_adjoint(BatchTraj const& bt, const BatchParticle* P, int np, int64_t nmax, const double* obs,
                             const double* nrm, int ns, int directional, double scale, const double* raw,
                             const double* pd, const double* target, const double* dloss, double* gy, double* ga,
                             int64_t total, hipStream_t s);

}  // namespace srs
