/* CPU oracle: a clean-room FP64 restatement of the reference OSCARS SR CPU path
 * (dhidas/SRSOriginal, OSCARS SR 1.31.05) for the radiation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load liboracle.so, and only as the checker / the timed CPU
 * baseline.  The product (srsoriginal_amd) never links or calls it.
 *
 * Every function follows the reference operation order so that, compiled with the
 * reference's own flags (setup.py:33: -std=c++11 -O3 -pedantic -fPIC -pthread, no
 * FMA contraction on x86-64), results are bit-identical to the reference engine.
 * Each is pinned in tests/test_oracle_golden.py against vectors written by the
 * compiled reference (tests/golden/) and in tests/test_oracle_closed_form.py
 * against closed forms; see the per-function comments in oracle.cc.
 *
 * Layouts: trajectories are SoA float64 [9][n] = x,y,z, bx,by,bz, ax,ay,az;
 * point lists are AoS float64 [ns][3].
 */
#ifndef SRS_ORACLE_H
#define SRS_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* One ideal undulator (TField3D_IdealUndulator). */
/* A field element: kind 0 ideal undulator (field, period, nperiods, phase, taper,
 * center); kind 1 Gaussian, kind 2 uniform box (field already rotated, width = sigma or
 * full widths, center, rotations of the position about X, Y, Z). */
typedef struct oracle_undulator {
  double field[3];
  double period[3];
  int nperiods;
  double phase;
  double taper;
  double center[3];
  int kind;
  double width[3];
  double rotations[3];
} oracle_undulator;

/* A field set: magnetic elements, optionally one pole-scaled undulator
 * (pole_und != NULL, pole_scales[2*nperiods] = delta_p), and electric elements (eund,
 * the same kinds, in V/m). */
typedef struct oracle_fields {
  const oracle_undulator* und;
  int nund;
  const oracle_undulator* pole_und;
  const double* pole_scales;
  const oracle_undulator* eund;
  int neund;
} oracle_fields;

/* Particle state needed by the trajectory (TParticleA). */
typedef struct oracle_particle {
  double q, m, current;
  double x0[3], beta0[3];
  double t0;
  double gamma, qmg;
} oracle_particle;

/* Electron of energy E [GeV] along d0 from x0, ideal (TParticleBeam::GetNewParticle("ideal")). */
void oracle_ideal_electron(double energy_gev, const double x0[3], const double d0[3], double current, oracle_particle* out);

/* Random particles from a Gaussian beam (TParticleBeam::GetNewParticle()) using a
 * std::mt19937 seeded as TRandomA::SetSeed and a std::normal_distribution whose state
 * persists across calls exactly like the reference's process-global generator. */
/* reset_normal != 0 also clears the normal distribution's cached value (a fresh process). */
void oracle_rng_seed(int seed, int reset_normal);
double oracle_rng_normal(void);
double oracle_rng_uniform(void);
/* A Gaussian beam (the device sampler's srs_beam): e0 [GeV], sigma_e, particle mass
 * [kg] and charge [C], current, start, unit direction, horizontal / vertical unit
 * directions, lattice reference point, rms positions and angles, t0. */
typedef struct oracle_beam {
  double e0, sigma_e, mass, charge, current;
  double x0[3], u0[3], hdir[3], vdir[3], sigma_at[3];
  double sigma_u[2], sigma_up[2];
  double t0;
} oracle_beam;
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void oracle_counter_normals(uint64_t seed, uint64_t idx, double out[6]);
void oracle_sample_beam(const oracle_beam* bm, uint64_t seed, uint64_t idx, oracle_particle* out);
int oracle_beam_index(const double* weights, int nb, uint64_t seed, uint64_t idx);
void oracle_random_electron(double energy_gev, double sigma_e_gev, const double x0[3], const double d0[3],
                            const double hdir[3], const double beta[2], const double emit[2],
                            const double lattice_ref[3], double current, oracle_particle* out);

/* Number of trajectory points CalculateTrajectory produces. */
int64_t oracle_trajectory_npoints(double ct_start, double ct_stop, int64_t npoints, double t0);
/* RK4 trajectory; writes traj[9][n] (n from oracle_trajectory_npoints) and dt. */
int64_t oracle_trajectory(const oracle_fields* f, const oracle_particle* p, double ct_start, double ct_stop,
                          int64_t npoints, double* traj, double* dt);
/* Summed magnetic field at a point. */
void oracle_bfield(const oracle_fields* f, const double x[3], double out[3]);
/* Summed electric field at a point. */
void oracle_efield(const oracle_fields* f, const double x[3], double out[3]);

/* Energy grid of TSpectrumContainer::Init(N, E0, E1). */
void oracle_energy_grid(int64_t n, double e0, double e1, double* out);

/* Rectangle surface (TSurfacePoints_Rectangle::Init(plane, ...)).  Returns 0 or -1 on bad plane. */
int oracle_rect_surface(const char* plane, int nx1, int nx2, double w1, double w2, const double rot[3],
                        const double trans[3], int normal, double* xyz, double* nrm, double* x1x2);
/* Rectangle from three corners (TSurfacePoints_Rectangle::Init(NX1, NX2, X0, X1, X2, Normal)).
 * Returns 0, or -2 for a normal other than -1, 0, 1. */
int oracle_rect3_surface(int nx1, int nx2, const double x0[3], const double x1[3], const double x2[3], int normal,
                         double* xyz, double* nrm, double* x1x2);

/* Single-particle integrals; out[i] = value * weight (no container semantics). */
void oracle_spectrum(const double* traj, int64_t n, double dt, const double obs[3], const double* energies_ev,
                     int64_t ne, double q, double current, double weight, int nthreads, double* out);
void oracle_flux(const double* traj, int64_t n, double dt, const double* obs, int64_t ns, double energy_ev,
                 double q, double current, double weight, int nthreads, double* out);
void oracle_power_density(const double* traj, int64_t n, double dt, const double* obs, const double* normals,
                          int64_t ns, int directional, double q, double current, int nthreads, double* out);
double oracle_total_power(const double* traj, int64_t n, double dt, double q, double current, double gamma);

/* Loss L = sum_j (F_j - target_j)^2 of the flux at one energy for the field f, whose
 * pole-scaled undulator (pole_und, pole_scales = deltas) is differentiated, and
 * grad_out[p] = dL/d delta_p for its 2*nperiods poles, by forward-mode dual numbers
 * through the trajectory and the flux integral (one thread per pole, nthreads
 * concurrent).  Returns 0, or -1 without a pole-scaled undulator. */
int oracle_pole_loss_gradient(const oracle_fields* f, const oracle_particle* p, double ct_start, double ct_stop,
                              int64_t npoints, const double* obs, int64_t ns, double energy_ev, const double* target,
                              double* loss_out, double* grad_out, int nthreads);

/* The same for L = sum_j (P_j - target_j)^2 of the power density (W/mm^2) on ns points
 * with normals; pd_out[j] = P_j (may be NULL). */
int oracle_pole_pd_loss_gradient(const oracle_fields* f, const oracle_particle* p, double ct_start, double ct_stop,
                                 int64_t npoints, const double* obs, const double* normals, int64_t ns,
                                 int directional, const double* target, double* pd_out, double* loss_out,
                                 double* grad_out, int nthreads);

/* Electric field in the time domain at obs (OSCARSSR::CalculateElectricFieldTimeDomain,
 * OSCARSSR.cc:2646-2690): e_out[3*i + k] = E_k of trajectory point i, t_out[i] = its
 * arrival time i*dt + D/c. */
void oracle_efield_time(const double* traj, int64_t n, double dt, const double obs[3], double q, double* e_out,
                        double* t_out);

/* Near- plus far-field flux of OSCARSSR::CalculateFlux1 (OSCARSSR.cc:2209-2346) on ns
 * points (AoS xyz) at one photon energy; out[i] = its ThisFlux (no container). */
void oracle_flux1(const double* traj, int64_t n, double dt, const double* obs, int64_t ns, double energy_ev, double q,
                  double current, double* out);

/* Kahan add of TSpectrumContainer::AddToFlux / T3DScalarContainer::AddToPoint. */
void oracle_kahan_add(double* sum, double* comp, const double* v, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
