/* srs_hip.h - C ABI of libsrs_hip.so, the MI355X (gfx950) radiation kernels.
 *
 * This is the drop-in boundary that replaces the reference's CUDA backend
 * (include/OSCARSSR_Cuda.h:25-28, src/OSCARSSR_Cuda.cu).  The reference's
 * functions are declared extern "C" but take C++ objects; these take plain
 * pointers and sizes only, so any host language can bind them (ctypes, cgo, JNI).
 *
 * Contract (same as the reference backend's, OSCARSSR.cc:892-900, :1874-1882,
 * :2580-2588):
 *   - the caller computes the trajectory on the host and passes it in;
 *   - the callee returns one value per output point; the caller applies the
 *     particle weight and accumulates into its containers;
 *   - results follow the reference CPU integrands (OSCARSSR.cc:1202-1246,
 *     :2054-2092, :1522-1573); SRS_PRECISION_FP64 matches them to < 1e-10
 *     relative, SRS_PRECISION_MIXED to < 1e-5 (FP64 phase, FP32 trig/sums).
 * Every function returns SRS_OK or an error code and sets a message readable
 * with srs_hip_last_error(); nothing throws across the boundary.
 * Host-pointer entry points are synchronous.  Calls from several host threads
 * are serialised internally (one device, one stream).
 */
#ifndef SRS_HIP_H
#define SRS_HIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define SRS_OK 0
#define SRS_ERR_NO_DEVICE 1
#define SRS_ERR_INVALID 2
#define SRS_ERR_HIP 3

#define SRS_PRECISION_FP64 0
#define SRS_PRECISION_MIXED 1

/* Trajectory, structure of arrays, n points spaced dt [s] (TParticleTrajectoryPoints,
 * TParticleTrajectoryPoints.h:48-55): position x,y,z [m], beta = v/c, and A/c [1/s]
 * (ax, ay, az are read by the power density only and may be NULL otherwise). */
typedef struct srs_traj {
  int64_t n;
  double dt;
  const double* x;
  const double* y;
  const double* z;
  const double* bx;
  const double* by;
  const double* bz;
  const double* ax;
  const double* ay;
  const double* az;
} srs_traj;

/* Number of visible gfx950 devices; replaces OSCARSSR_Cuda_GetDeviceCount
 * (OSCARSSR_Cuda.cu:30-36, declared OSCARSSR_Cuda.h:25). */
int srs_hip_device_count(void);

/* Message of the last failed call on this thread ("" if none). */
const char* srs_hip_last_error(void);

/* Spectrum at one observation point for ne photon energies [eV]; flux_out[e] in
 * photons/s/0.1%bw/mm^2, unweighted.  Replaces OSCARSSR_Cuda_CalculateSpectrumGPU
 * (OSCARSSR_Cuda.h:26, OSCARSSR_Cuda.cu:1015-1204).  charge [C] and current [A]
 * are the particle's (TParticleA::GetQ / GetCurrent). */
int srs_hip_spectrum(const srs_traj* traj, const double obs[3], const double* energies_ev, int64_t ne, double charge,
                     double current, int precision, double* flux_out);

/* Flux on ns observation points (AoS xyz) for ne energies; flux_out[e*ns + j].
 * Replaces OSCARSSR_Cuda_CalculateFluxGPU (OSCARSSR_Cuda.h:27, OSCARSSR_Cuda.cu:696-890). */
int srs_hip_flux(const srs_traj* traj, const double* obs_xyz, int64_t ns, const double* energies_ev, int64_t ne,
                 double charge, double current, int precision, double* flux_out);

/* Near- plus far-field flux on ns observation points (AoS xyz) at one photon energy,
 * from the Lienard-Wiechert field with the stored A/c (the trajectory's ax..az are
 * required): the sum of OSCARSSR::CalculateFlux1 (OSCARSSR.cc:2209-2346, its CPU
 * loop :2266-2300; the reference has no GPU form of it), Omega = E 2 pi / 4.1357e-15
 * as written there.  FP64; flux_out[j] in photons/s/0.1%bw/mm^2, unweighted. */
int srs_hip_flux1(const srs_traj* traj, const double* obs_xyz, int64_t ns, double energy_ev, double charge,
                  double current, double* flux_out);

/* Power density [W/mm^2] on ns points with surface normals (AoS xyz).
 * directional != 0 skips trajectory points seen from behind the surface and keeps
 * the sign; otherwise |sum| (the reference CPU semantics, OSCARSSR.cc:1549-1573).
 * Replaces OSCARSSR_Cuda_CalculatePowerDensityGPU (OSCARSSR_Cuda.h:28,
 * OSCARSSR_Cuda.cu:1342-1547). */
int srs_hip_power_density(const srs_traj* traj, const double* obs_xyz, const double* normals_xyz, int64_t ns,
                          int directional, double charge, double current, int precision, double* pd_out);

/* Electric field in the time domain at obs: efield_out[3*i + k] = E_k [V/m] radiated
 * from trajectory point i, time_out[i] = its arrival time i*dt + |R|/c [s] (unevenly
 * spaced).  Computes OSCARSSR::CalculateElectricFieldTimeDomain (OSCARSSR.cc:2646-2690),
 * which the reference runs on the CPU only; charge [C] is the particle's. */
int srs_hip_efield_time(const srs_traj* traj, const double obs[3], double charge, double* efield_out,
                        double* time_out);

/* Total radiated power [W] of one trajectory: sum_i (|A_i|^2 - |beta_i x A_i|^2) dt times
 * |charge*current| gamma^6 / (6 pi eps0 c).  Computes OSCARSSR::CalculateTotalPower
 * (OSCARSSR.cc:1939-1978), a sequential CPU sum in the reference; here a fixed-order
 * device reduction (the same value to rounding, ~1e-15 relative). */
int srs_hip_total_power(const srs_traj* traj, double charge, double current, double gamma, double* power_out);

/* Device-resident variants: every array pointer is device memory, observation
 * points and normals are SoA ([3][ns]); work is queued on `stream` (a hipStream_t,
 * NULL = the library's stream) and not synchronised.  No counterpart in the
 * reference (which copies host data on every call); they serve pipelines that
 * keep data in HBM and timing of the kernels alone. */
int srs_hip_flux_device(const srs_traj* traj, const double* obs_soa, int64_t ns, const double* energies_ev,
                        int64_t ne, double charge, double current, int precision, double* flux_out, void* stream);

/* srs_hip_flux_device without allocation: the same flux (the same kernels, chunking and
 * results) with its scratch -- the FP32 trajectory staging of the mixed form and the
 * per-chunk partial sums -- in SRS_NOALLOC_SCRATCH_BYTES of the library's own device
 * memory (a static device array, one copy per GPU, held until the process ends)
 * instead of buffers the library allocates.  No counterpart in the reference: it is for
 * callers that cannot allocate (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  A call needing more scratch -- (trajectory chunks
 * x energies per launch (<= 4) x 48 B in FP64, 24 B in mixed precision) per observation
 * point, + 12 B per trajectory point in mixed precision -- is rejected with
 * SRS_ERR_INVALID.  Every call on a GPU
 * shares that GPU's scratch: calls that overlap on one GPU (on different streams) race.
 * Nothing in it has to start at zero (each call writes all it reads).  The library's
 * first call on a device, whichever entry point, also sets it up (streams, the scratch
 * address): make one call before capturing a graph.  128 MiB holds, for example, a
 * 201 x 201 map at one energy over 20001 trajectory points (50 chunks). */
#define SRS_NOALLOC_SCRATCH_BYTES (128u << 20)
int srs_hip_flux_device_noalloc(const srs_traj* traj, const double* obs_soa, int64_t ns, const double* energies_ev,
                                int64_t ne, double charge, double current, int precision, double* flux_out,
                                void* stream);
int srs_hip_spectrum_device(const srs_traj* traj, const double obs[3], const double* energies_ev, int64_t ne,
                            double charge, double current, int precision, double* flux_out, void* stream);
int srs_hip_power_density_device(const srs_traj* traj, const double* obs_soa, const double* normals_soa, int64_t ns,
                                 int directional, double charge, double current, int precision, double* pd_out,
                                 void* stream);
/* ---- batched multi-particle path (device RK4 + persistent work queue) ----------
 * No single counterpart in the reference: it replaces the nparticles loop of
 * OSCARSSR::CalculateFlux (OSCARSSR.cc:2465-2479: new particle, host RK4, one GPU
 * call per particle) with one device pipeline.  Particles are sampled by the caller
 * (host RNG, reference order) and passed as initial conditions. */

/* One field element, magnetic [T] (target SRS_FIELD_MAGNETIC, 0) or electric [V/m]
 * (SRS_FIELD_ELECTRIC, OSCARSSR::AddElectricField); the elements of each kind of field
 * are summed in order (TFieldContainer::GetF, TFieldContainer.cc:107-117), and the RK4
 * force is (q/m gamma)(E + v x B) in the operation order of OSCARSSR::DerivativesEB
 * (OSCARSSR.cc:769-794), which equals DerivativesB / DerivativesE (:741-765, :713-737)
 * bit for bit when the other field is absent.
 *   kind SRS_FIELD_UNDULATOR (0, so zero-initialised records are undulators): ideal
 *     undulator (TField3D_IdealUndulator.cc:50-139) of field, period, nperiods, phase,
 *     taper, center; pole_delta != NULL scales its 2*nperiods interior half-period poles
 *     by (1 + delta_p) (phase, taper and center unused then);
 *   kind SRS_FIELD_GAUSSIAN: field * prod_k exp(-(r_k/width_k)^2/2) over the axes with
 *     width_k > 0 (TField3D_Gaussian.cc:88-113);
 *   kind SRS_FIELD_UNIFORM_BOX: field inside |r_k| <= |width_k|/2 for the axes with
 *     width_k > 0, 0 outside (TField3D_UniformBox.cc:94-117);
 *   kind SRS_FIELD_GRID: a field map (TField3D_Grid.cc:124-250): grid holds
 *     3*grid_n[0]*grid_n[1]*grid_n[2] doubles, the field vectors at (ix, iy, iz) in the
 *     order ((ix*n1 + iy)*n2 + iz), already rotated; on the axes with grid_n_k > 1 the
 *     field is 0 unless grid_start_k < r_k < grid_stop_k and is interpolated linearly
 *     (x, then y, then z) from cell index trunc((r_k - start_k)/step_k);
 * where r = Rot(X) - center, Rot rotating the position by rotations[0], [1], [2] about
 * X, then Y, then Z (TVector3D::RotateSelfXYZ); field is given already rotated, as the
 * reference's constructors rotate it. */
#define SRS_FIELD_UNDULATOR 0
#define SRS_FIELD_GAUSSIAN 1
#define SRS_FIELD_UNIFORM_BOX 2
#define SRS_FIELD_GRID 3
#define SRS_FIELD_MAGNETIC 0
#define SRS_FIELD_ELECTRIC 1
typedef struct srs_field {
  double field[3];
  double period[3];
  int32_t nperiods;
  double phase;
  double taper;
  double center[3];
  const double* pole_delta;
  int32_t kind;
  double width[3];
  double rotations[3];
  const double* grid; /* SRS_FIELD_GRID only */
  int32_t grid_n[3];
  double grid_start[3];
  double grid_step[3];
  double grid_stop[3];
  int32_t target; /* SRS_FIELD_MAGNETIC or SRS_FIELD_ELECTRIC */
} srs_field;

/* One particle: initial position, beta, q/(m gamma), T0 [m], its ct range and point
 * count (set_ctstartstop / set_npoints_trajectory; dt and the forward/backward point
 * counts follow OSCARSSR.cc:592-598), and its charge [C] and beam current [A] as
 * TParticleA holds them (TParticleA.cc:254-266).  The batch entry points form each
 * particle's prefactors from its own charge and current, as the reference's
 * per-particle loop does (OSCARSSR.cc:2465-2479); charge 0 means the call's charge
 * and current. */
typedef struct srs_particle {
  double x0[3];
  double beta0[3];
  double qmg;
  double t0;
  double ct_start;
  double ct_stop;
  int64_t npoints;
  double charge;
  double current;
} srs_particle;

/* A Gaussian beam for the device sampler (TParticleBeam, TParticleBeam.cc:137-265):
 * energy e0 [GeV] with rms sigma_e [GeV], the particle's mass [kg] and charge [C],
 * start x0 [m], unit direction u0, unit horizontal / vertical directions, the lattice
 * reference point sigma_at, rms positions sigma_u [m] and angles sigma_up [rad]
 * (horizontal, vertical), t0, the ct range and point count every sampled particle's
 * trajectory gets, and the beam current [A] (0: the charge, as TParticleA.cc:258-262
 * sets it).  Sampled particles carry the beam's charge and current. */
typedef struct srs_beam {
  double e0, sigma_e;
  double mass, charge;
  double x0[3], u0[3], hdir[3], vdir[3], sigma_at[3];
  double sigma_u[2], sigma_up[2];
  double t0, ct_start, ct_stop;
  int64_t npoints;
  double current;
} srs_beam;

/* np particles of the beam drawn on the device, particle i from counter first + i
 * under key seed: Philox4x32-10 (Salmon et al., SC'11) -> 53-bit uniforms -> Box-Muller
 * normals n0..n4, then TParticleBeam::GetNewParticle's transform (TParticleBeam.cc:
 * 228-265) with n0..n4 in the reference's draw order (energy, x, y, x', y').  Draws are
 * reproducible and independent of np and of the order of calls (any particle can be
 * redrawn from its counter); the reference's host stream (std::mt19937) is the other
 * sampler, kept bit for bit on the host.  No counterpart on the reference's GPU. */
int srs_hip_sample_beam(const srs_beam* beam, int64_t np, uint64_t seed, uint64_t first, srs_particle* out);

/* The same from nbeams beams with weights (TParticleBeamContainer::GetRandomBeam,
 * TParticleBeamContainer.cc:100-142): particle i takes the first beam k with
 * u * W < w_0 + ... + w_k, u a fourth Philox block's 53-bit uniform and W the weights'
 * sum (no draw for one beam), then that beam's transform.  Weights must be >= 0 with a
 * positive sum. */
int srs_hip_sample_beams(const srs_beam* beams, const double* weights, int nbeams, int64_t np, uint64_t seed,
                         uint64_t first, srs_particle* out);

/* Points the trajectories of np particles will have (total over particles). */
int srs_hip_trajectory_points(const srs_particle* p, int64_t np, int64_t* total);

/* Device RK4 (FP64) of np particles (the GPU switch of calculate_trajectory).  traj_out
 * is host memory [9][total] (x,y,z, beta, A/c) packed particle after particle;
 * offsets_out[np + 1], dt_out[np]. */
int srs_hip_trajectories(const srs_field* und, int nund, const srs_particle* p, int64_t np, double* traj_out,
                         int64_t total, int64_t* offsets_out, double* dt_out);

#define SRS_BATCH_GRAPH 1    /* run the pipeline as a captured hipGraph, replayed while the shape is unchanged */
#define SRS_BATCH_AUTOTUNE 2 /* time the candidate tile shapes on this problem once; keep the fastest */

/* Multi-particle flux: flux_out[e*ns + j] = sum over particles (in order, Kahan) of
 * weight * flux_p(e, j), flux_p with the particle's own charge and current (the call's
 * charge and current for particles with charge 0). */
int srs_hip_batch_flux(const srs_field* und, int nund, const srs_particle* p, int64_t np, double charge,
                       double current, const double* obs_xyz, int64_t ns, const double* energies_ev, int64_t ne,
                       double weight, int precision, int flags, double* flux_out);

/* The multi-particle flux of np particles drawn on the device (the draws of
 * srs_hip_sample_beams: particle i from counter first + i under key seed) straight into
 * the batch pipeline's particle records, with no host round trip: each particle takes its
 * beam's charge, current, ct range and point count; then srs_hip_batch_flux's pipeline
 * (flags as there).  The device counterpart of the reference's per-particle loop over
 * GetNewParticle (OSCARSSR.cc:2465-2479), which has no GPU form in the reference.
 * Beams need a nonzero charge. */
int srs_hip_batch_flux_beams(const srs_field* und, int nund, const srs_beam* beams, const double* weights,
                             int nbeams, int64_t np, uint64_t seed, uint64_t first, const double* obs_xyz, int64_t ns,
                             const double* energies_ev, int64_t ne, double weight, int precision, int flags,
                             double* flux_out);

/* Multi-particle power density [W/mm^2] on ns points with normals (AoS xyz):
 * pd_out[j] = sum over particles (in order, Kahan) of weight * P_p(j), each P_p as
 * srs_hip_power_density with `directional`.  Replaces the nparticles loop of
 * OSCARSSR::CalculatePowerDensity with the GPU switch (OSCARSSR.cc:1655-1669: new
 * particle, host RK4, one GPU call per particle with Weight = 1/NP).  precision
 * SRS_PRECISION_FP64 (the reference's terms) or SRS_PRECISION_MIXED (the FP32 terms of
 * srs_hip_power_density's mixed form). */
int srs_hip_batch_power_density(const srs_field* und, int nund, const srs_particle* p, int64_t np, double charge,
                                double current, const double* obs_xyz, const double* normals_xyz, int64_t ns,
                                int directional, double weight, int precision, double* pd_out);

/* Multi-particle spectrum at one observation point: flux_out[e] = sum over particles
 * (in order, Kahan) of weight * spectrum_p(e).  Replaces the nparticles loop of
 * OSCARSSR::CalculateSpectrum (OSCARSSR.cc:963-982: new particle, host RK4, one
 * spectrum per particle with Weight = 1/NP).  precision SRS_PRECISION_FP64 (the
 * reference's FP64 terms) or SRS_PRECISION_MIXED (FP64 phase, FP32 sin/cos, amplitudes
 * and sub-sums). */
int srs_hip_batch_spectrum(const srs_field* und, int nund, const srs_particle* p, int64_t np, double charge,
                           double current, const double obs[3], const double* energies_ev, int64_t ne, double weight,
                           int precision, double* flux_out);

/* Tile shape used by the last batch call: energies per work item and trajectory
 * points per LDS tile. */
int srs_hip_batch_config(int* ne_per_item, int* tile);

/* Observation points per lane of the tile shape used by the last batch call: 1 (256
 * points per work item), or 2 (512 points per item, each staged trajectory record
 * integrated for two points; the mixed form at one energy per item). */
int srs_hip_batch_points_per_lane(int* points_per_lane);

/* Fix the tile shape of later batch flux calls that do not autotune: energies per
 * item, trajectory points per LDS tile, observation points per lane.  ne_per_item = 0
 * restores the default shape.  A shape with no kernel (see srs_hip_batch_config's
 * candidates: 1-4 energies, tiles of 64-256, two points per lane only at one energy
 * and tile 128 or 256) fails; two points per lane fail later calls in FP64.  No
 * counterpart in the reference (its per-particle loop has no tiles). */
int srs_hip_batch_set_config(int ne_per_item, int tile, int points_per_lane);

/* Kernel timing with HIP events recorded on the launch stream around each main
 * integral kernel (not the staging or the chunk reduction).  enable resets the
 * record; read waits for the recorded kernels, returns their summed duration and
 * count, and resets.  No counterpart in the reference (it has no timers). */
int srs_hip_timing_enable(int on);
int srs_hip_timing_read(double* total_ms, int64_t* launches);

/* Waits for all work queued on `stream` (NULL = the library's stream). */
int srs_hip_synchronize(void* stream);

/* ---- gradient of a flux loss with respect to per-pole field errors ----------------
 * No counterpart in the reference (it has no derivatives); the forward model is the
 * reference's pole-scaled ideal undulator, RK4 trajectory (OSCARSSR.cc:562-874) and
 * flux integral (OSCARSSR.cc:2054-2092) in FP64.  Realisation r scales the
 * 2*nperiods interior half-period poles of und[poled] by (1 + deltas[r*2N + p]);
 * every realisation flies `particle` through the field of all undulators.
 *   L_r = sum_j (F_rj - target_j)^2,  F_rj = flux of realisation r at obs point j.
 * forward: loss_out[r] (and flux_out[r*ns + j] when not NULL), with RK4 checkpoints
 * every checkpoint_every steps kept on the device; *tape_out identifies that state.
 * backward: grad_out[r*2N + p] = dloss[r] * dL_r/d delta_p (dloss NULL = ones), by a
 * hand-written adjoint of the flux integral and of the RK4 (each checkpoint segment
 * recomputed).  A later forward, batch or trajectory call replaces the tape. */
int srs_hip_pole_forward(const srs_field* und, int nund, int poled, const srs_particle* particle, int64_t nreal,
                         const double* deltas, double charge, double current, const double* obs_xyz, int64_t ns,
                         double energy_ev, const double* target, int checkpoint_every, double* flux_out,
                         double* loss_out, int64_t* tape_out);
/* srs_hip_pole_forward with a precision: SRS_PRECISION_FP64 (as srs_hip_pole_forward) or
 * SRS_PRECISION_MIXED, where the forward flux takes the mixed terms (FP64 phase, FP32
 * trig, amplitudes and tile sums) and srs_hip_pole_backward the mixed adjoint (the same
 * FP64 phase; the amplitudes, their derivatives and the products in FP32, sums in
 * FP64).  No counterpart in the reference (its flux has no gradient). */
int srs_hip_pole_forward_ex(const srs_field* und, int nund, int poled, const srs_particle* particle, int64_t nreal,
                            const double* deltas, double charge, double current, const double* obs_xyz, int64_t ns,
                            double energy_ev, const double* target, int checkpoint_every, int precision,
                            double* flux_out, double* loss_out, int64_t* tape_out);
/* The same for the power density (OSCARSSR.cc:1477-1598: W/mm^2 on ns points with
 * surface normals, AoS xyz; directional as in srs_hip_power_density):
 *   L_r = sum_j (P_rj - target_j)^2,  P_rj = power density of realisation r at point j;
 * pd_out[r*ns + j] = P_rj when not NULL.  The stored acceleration A/c of each
 * trajectory point (lagged one step, as the reference stores it) is differentiated
 * through the RK4 step that produced it.  No counterpart in the reference. */
int srs_hip_pole_pd_forward(const srs_field* und, int nund, int poled, const srs_particle* particle,
                            int64_t nreal, const double* deltas, double charge, double current,
                            const double* obs_xyz, const double* normals_xyz, int64_t ns, int directional,
                            const double* target, int checkpoint_every, double* pd_out, double* loss_out,
                            int64_t* tape_out);
/* Backward of the last srs_hip_pole_forward or srs_hip_pole_pd_forward. */
int srs_hip_pole_backward(int64_t tape, const double* dloss, double* grad_out);

#ifdef __cplusplus
}
#endif
#endif
