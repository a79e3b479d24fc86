// C ABI of libsrs_hip.so (declared in include/srs_hip.h).
//
// Host side of the boundary: validates arguments, derives the per-energy
// constants in the reference's operation order, stages host arrays into
// grow-only device buffers (no allocation per call once warm), picks the
// trajectory chunking and launches the kernels on one stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/srs_hip.h"
#include "srs_batch.h"
#include "srs_kernels.h"

#include <map>

namespace {

using srs::EnergySet;
using srs::TrajView;

// Reference constants, TOSCARSSR.h:20-69.
constexpr double kPi = 3.14159265358979323846;
constexpr double kFourPi = 4.0 * kPi;
constexpr double kTwoPi = 2.0 * kPi;
constexpr double kPi2 = kPi * kPi;
constexpr double kSqrt2Pi = 2.5066282746310002416123552393401;
constexpr double kC = 2.99792458e8;
constexpr double kH = 6.62606876e-34;
constexpr double kHbar = 1.054571596e-34;
constexpr double kQe = 1.602176462e-19;
constexpr double kEps0 = 8.854187817E-12;
constexpr double kMu0 = 1.2566370614E-6;

thread_local std::string g_err;

int fail(int code, std::string const& msg) {
  g_err = msg;
  return code;
}

struct Buf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      if (e != hipSuccess) return e;
      p = nullptr;
      cap = 0;
    }
    size_t const want = std::max<size_t>(bytes, 1 << 16);
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) cap = want;
    return e;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

struct Context {
  std::mutex mu;
  bool ready = false;
  hipStream_t stream = nullptr;
  // second stream of the batched pipeline: the RK4 segments run there while the
  // flux integrates the previous segment on `stream` (same default priority)
  hipStream_t stream2 = nullptr;
  static constexpr int kMaxSegments = 8;
  hipEvent_t seg_ev[kMaxSegments + 1] = {};
  Buf traj, trajf, obs, nrm, energies, part, out;
  // optional HIP-event timing of the integral kernels (srs_hip_timing_*)
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
  size_t ev_used = 0;
  // batched path
  Buf b_part, b_und, b_poles, b_grid, b_traj, b_trajf, b_energies, b_flux_pp, b_out, b_counter, b_obs;
  Buf b_state, b_field;  // RK4 state between segments; field sums accumulated over segments
  Buf b_beams;           // beam records of the device-sampled batch
  bool und_local = false;  // the staged field list holds elements other than magnetic undulators
  int und_force = srs::kForceB;  // the derivative form: magnetic, electric or both kinds of elements
  int pole_count = 0;            // pole deltas in each particle's table in b_poles (0: none)
  hipGraph_t b_graph = nullptr;
  hipGraphExec_t b_exec = nullptr;
  std::string b_key;
  std::map<std::string, srs::BatchConfig> b_tuned;
  srs::BatchConfig b_last{1, 128};
  srs::BatchConfig b_fixed{0, 0};  // srs_hip_batch_set_config (ne_per_item 0: none)
  int cus = 0;
  // pinned host flag of the spectrum kernels: the trajectory left the range of the
  // energy-run phase correction (the call is then evaluated one energy per lane)
  int* spec_range = nullptr;
  // pole-error gradient: the forward call leaves its tape here for the backward call
  Buf g_ckpt, g_field, g_target, g_loss, g_dloss, g_gy, g_rec, g_grad, g_lam, g_sweep, g_nrm, g_part;
  struct Tape {
    int64_t id = 0;
    bool valid = false;
    int np = 0, ns = 0, nund = 0, poled = 0, npoles = 0, every = 0;
    int64_t nf = 0, nb = 0;
    int64_t seg_per_pass = 1;  // checkpoint segments recomputed per adjoint pass
    int64_t total = 0, nmax = 0;
    srs::BatchTraj bt{};
    srs::Checkpoints ck{};
    double omega = 0, cw = 0, c0w = 0, c2 = 0;  // flux
    int quantity = 0;                             // 0 flux, 1 power density
    bool mixed = false;                           // flux: mixed-precision forward and adjoint
    srs::AdjScreen screen{};                      // flux: the points' plane and extent
    int directional = 0;
    double pd_scale = 0;
  } tape;
  int64_t tape_counter = 0;
};

Context& ctx() {
  static Context c;
  return c;
}

// Scratch of srs_hip_flux_device_noalloc: device memory of the library itself, so that
// entry point allocates nothing.  Every GPU the library runs on holds its own copy until
// the process ends.  Host code never passes the array's name (a host-side stand-in) to
// the device: it uses the address hipGetSymbolAddress returns for the current device.
constexpr int kScratchDevices = 64;
static __device__ __align__(256) unsigned char scratch[SRS_NOALLOC_SCRATCH_BYTES];
void* g_scratch_dev[kScratchDevices] = {};  // per device, filled once with that device current

int scratch_address(int dev, void** p) {
  if (dev < 0 || dev >= kScratchDevices) return fail(SRS_ERR_INVALID, "device index out of range for the scratch");
  if (!g_scratch_dev[dev]) {
    void* a = nullptr;
    hipError_t const e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fail(SRS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[dev] = a;
  }
  *p = g_scratch_dev[dev];
  return SRS_OK;
}

#define HIPCHK(expr)                                                                 \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess) return fail(SRS_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(_e)); \
  } while (0)

#define TRY(expr)              \
  do {                         \
    int _r = (expr);           \
    if (_r != SRS_OK) return _r; \
  } while (0)

int device_count_raw() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int init(Context& c) {
  if (c.ready) return SRS_OK;
  if (device_count_raw() < 1) return fail(SRS_ERR_NO_DEVICE, "You are requesting the GPU, but none were found");
  HIPCHK(hipSetDevice(0));
  HIPCHK(hipStreamCreate(&c.stream));
  HIPCHK(hipStreamCreate(&c.stream2));
  void* sp = nullptr;
  TRY(scratch_address(0, &sp));  // set-up of the allocation-free entry point on the library's device
  for (auto& e : c.seg_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  HIPCHK(hipHostMalloc((void**)&c.spec_range, sizeof(int), hipHostMallocDefault));
  c.ready = true;
  return SRS_OK;
}

// Brackets the main integral kernel of a launch with HIP events on its stream.
struct KernelTimer {
  Context& c;
  hipStream_t s;
  hipEvent_t stop = nullptr;
  KernelTimer(Context& cc, hipStream_t ss) : c(cc), s(ss) {
    if (!c.timing) return;
    if (c.ev_used == c.ev_pool.size()) {
      hipEvent_t a, b;
      if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
      c.ev_pool.push_back({a, b});
    }
    auto& p = c.ev_pool[c.ev_used++];
    (void)hipEventRecord(p.first, s);
    stop = p.second;
  }
  ~KernelTimer() {
    if (stop) (void)hipEventRecord(stop, s);
  }
};

double omega_of(double e_ev) { return e_ev * kQe / kHbar; }                       // TOSCARSSR.h:66
double c0_of(double q) { return q / (kFourPi * kC * kEps0 * kSqrt2Pi); }          // OSCARSSR.cc:1191
double c2_of(double q, double cur) {                                               // OSCARSSR.cc:1194
  return kFourPi * cur / (kH * std::fabs(q) * kMu0 * kC) * 1e-6 * 0.001;
}
double pd_qfac_of(double q, double cur) {                                           // OSCARSSR.cc:1563
  return std::fabs(q * cur) / (16 * kPi2 * kEps0 * kC);
}
double pd_scale_of(double q, double cur, double dt) { return pd_qfac_of(q, cur) * dt; }

int check_traj(const srs_traj* t, bool need_a) {
  if (!t) return fail(SRS_ERR_INVALID, "trajectory is NULL");
  if (t->n < 1) return fail(SRS_ERR_INVALID, "no points in trajectory.  Is particle or beam defined?");
  if (t->n > (int64_t)1 << 31) return fail(SRS_ERR_INVALID, "trajectory longer than 2^31 points");
  if (!t->x || !t->y || !t->z || !t->bx || !t->by || !t->bz) return fail(SRS_ERR_INVALID, "trajectory array is NULL");
  if (need_a && (!t->ax || !t->ay || !t->az)) return fail(SRS_ERR_INVALID, "trajectory acceleration array is NULL");
  return SRS_OK;
}

// Chunking of the trajectory for a launch of bx x nch equal blocks (bx = point
// blocks of 256 lanes) of which the GPU holds `resident` at a time: the launch runs
// in ceil(bx*nch / resident) rounds of one chunk each, so the chunk (a multiple of
// align) that minimises rounds x chunk -- plus a small cost per chunk for its partial
// sums -- keeps every CU busy to the end.  (A C1 flux map, 158 x 12 blocks of 1792
// points with 1024 resident, runs 2 rounds where 1.85 would do; 19 chunks of 1056
// points fill 3 rounds to 97 %.)
int pick_chunk(int64_t n, int64_t lanes, int align, int64_t resident) {
  int64_t const bx = (lanes + srs::kBlock - 1) / srs::kBlock;
  int64_t const maxch = (n + align - 1) / align;
  resident = std::max<int64_t>(1, resident);
  int64_t const limit = std::max<int64_t>(1, std::min<int64_t>(maxch, std::min<int64_t>(4096, 4 * resident / bx + 2)));
  int64_t best = n, best_cost = -1;
  for (int64_t nch = 1; nch <= limit; ++nch) {
    int64_t chunk = (n + nch - 1) / nch;
    chunk = (chunk + align - 1) / align * align;
    int64_t const nc = (n + chunk - 1) / chunk;
    int64_t const rounds = (bx * nc + resident - 1) / resident;
    int64_t const cost = rounds * chunk + nc;
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = chunk;
    }
  }
  return (int)best;
}
int nchunks(int64_t n, int chunk) { return (int)((n + chunk - 1) / chunk); }

// Trajectory chunks of a flux launch: pick_chunk's chunk c1, and, when the launch takes
// more than one round of resident blocks, its last part re-split into chunks of c1 / 8
// (32-aligned), about one round of them (resident / point blocks).  Blocks run in chunk
// order, so the blocks that drain the device at the end are the short ones.
constexpr int kTailRatio = 8;
constexpr double kTailRounds = 1.0;
srs::ChunkMap flux_chunks(int64_t n, int64_t lanes, int align, int64_t resident) {
  int const c1 = pick_chunk(n, lanes, align, resident);
  int64_t const bx = (lanes + srs::kBlock - 1) / srs::kBlock;
  int const nu = nchunks(n, c1);
  srs::ChunkMap m{c1, nu, c1, nu};
  int64_t const c2 = (int64_t)(c1 / kTailRatio + align - 1) / align * align;
  if (bx * nu <= resident || c2 >= c1 || nu < 2) return m;
  int64_t const want = (int64_t)std::ceil(kTailRounds * (double)resident / (double)bx) * c2;
  int64_t const n1 = std::max<int64_t>(1, (n - std::min<int64_t>(want, n)) / c1);
  int64_t const rest = n - n1 * c1;
  if (rest <= 0) return m;
  m.n1 = (int)n1;
  m.c2 = (int)c2;
  m.nch = (int)(n1 + (rest + c2 - 1) / c2);
  return m;
}

// Copy a host trajectory into the context's device buffer (9 SoA arrays).
int stage_traj(Context& c, const srs_traj* h, bool need_a, TrajView* v) {
  size_t const n = (size_t)h->n;
  HIPCHK(c.traj.ensure(9 * n * sizeof(double)));
  double* d = c.traj.as<double>();
  const double* src[9] = {h->x, h->y, h->z, h->bx, h->by, h->bz, h->ax, h->ay, h->az};
  for (int k = 0; k < 9; ++k) {
    if (k >= 6 && !need_a) break;
    HIPCHK(hipMemcpyAsync(d + k * n, src[k], n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  }
  *v = TrajView{d, d + n, d + 2 * n, d + 3 * n, d + 4 * n, d + 5 * n, need_a ? d + 6 * n : nullptr,
                need_a ? d + 7 * n : nullptr, need_a ? d + 8 * n : nullptr, h->n, h->dt};
  return SRS_OK;
}

TrajView view_of(const srs_traj* t) {
  return TrajView{t->x, t->y, t->z, t->bx, t->by, t->bz, t->ax, t->ay, t->az, t->n, t->dt};
}

// ---- core launches on device data ------------------------------------------

// Device scratch of one flux evaluation, in one buffer: the FP32 staging of the mixed form
// (traj_f32_floats) at offset 0, then the chunk partial sums part[((c*ne0 + e)*6 + k)*ns + j],
// FP32 for the mixed form and FP64 for the FP64 form (ne0 = energies per launch, at most kMaxFluxEnergies; launches of later energy groups
// reuse them in stream order).
struct FluxPlan {
  srs::ChunkMap cm{};
  int nch = 0, ne0 = 0;
  size_t part_off = 0, bytes = 0;
};

int plan_flux(Context& c, const TrajView& t, int64_t ns, int64_t ne, int precision, FluxPlan* pl) {
  if (!c.cus) HIPCHK(hipDeviceGetAttribute(&c.cus, hipDeviceAttributeMultiprocessorCount, 0));
  pl->ne0 = (int)std::min<int64_t>(srs::kMaxFluxEnergies, ne);
  int const per_cu = precision == SRS_PRECISION_MIXED ? srs::flux_mixed_blocks_per_cu(pl->ne0)
                                                      : srs::flux_fp64_blocks_per_cu(pl->ne0);
  // point blocks of the launch, as the lanes of that many 256-lane blocks for the chunk choice
  int64_t const ppb = precision == SRS_PRECISION_MIXED ? srs::flux_mixed_points_per_block(pl->ne0) : srs::kBlock;
  pl->cm = flux_chunks(t.n, (ns + ppb - 1) / ppb * srs::kBlock, 32, (int64_t)c.cus * per_cu);
  pl->nch = pl->cm.nch;
  size_t const staged = precision == SRS_PRECISION_MIXED ? srs::traj_f32_floats(t.n, 0) * sizeof(float) : 0;
  pl->part_off = (staged + 255) / 256 * 256;
  size_t const psize = precision == SRS_PRECISION_MIXED ? sizeof(float) : sizeof(double);
  pl->bytes = pl->part_off + (size_t)pl->nch * pl->ne0 * 6 * ns * psize;
  return SRS_OK;
}

// The flux kernels on device data, with the scratch of plan_flux at `scratch` (device
// memory of at least pl.bytes).  Queues work on s only: no allocation, no synchronisation.
int run_flux(Context& c, const TrajView& t, const double* obs_soa, int64_t ns, const double* energies_ev, int64_t ne,
             double q, double cur, int precision, double* out, hipStream_t s, FluxPlan const& pl, void* scratch) {
  double const c0 = c0_of(q), c2 = c2_of(q, cur);
  int const maxe = srs::kMaxFluxEnergies;
  float* const trajf = static_cast<float*>(scratch);
  void* const part = static_cast<char*>(scratch) + pl.part_off;  // FP32 partials (mixed) or FP64
  if (precision == SRS_PRECISION_MIXED) HIPCHK(srs::launch_traj_to_f32(t, 0, trajf, s));
  for (int64_t e0 = 0; e0 < ne; e0 += maxe) {
    EnergySet es{};
    es.ne = (int)std::min<int64_t>(maxe, ne - e0);
    for (int e = 0; e < es.ne; ++e) {
      double const w = omega_of(energies_ev[e0 + e]);
      es.omega[e] = w;
      es.cw[e] = kC / w;
      es.kw[e] = (c0 * w) * t.dt;
      es.fr[e] = w / kTwoPi;
      es.fw[e] = es.fr[e] / kC;
    }
    {
      KernelTimer timer(c, s);
      if (precision == SRS_PRECISION_MIXED) {
        HIPCHK(srs::launch_flux_mixed(t, trajf, obs_soa, (int)ns, es, pl.cm, static_cast<float*>(part), s));
      } else {
        HIPCHK(srs::launch_flux_fp64(t, obs_soa, (int)ns, es, pl.cm, static_cast<double*>(part), s));
      }
    }
    if (precision == SRS_PRECISION_MIXED)
      HIPCHK(srs::launch_flux_reduce(static_cast<const float*>(part), pl.nch, (int)ns, es, c2, out + e0 * ns, s));
    else
      HIPCHK(srs::launch_flux_reduce(static_cast<const double*>(part), pl.nch, (int)ns, es, c2, out + e0 * ns, s));
  }
  return SRS_OK;
}

// run_flux with its scratch in the context's grow-only buffer.
int run_flux(Context& c, const TrajView& t, const double* obs_soa, int64_t ns, const double* energies_ev, int64_t ne,
             double q, double cur, int precision, double* out, hipStream_t s) {
  FluxPlan pl;
  TRY(plan_flux(c, t, ns, ne, precision, &pl));
  HIPCHK(c.part.ensure(pl.bytes));
  return run_flux(c, t, obs_soa, ns, energies_ev, ne, q, cur, precision, out, s, pl, c.part.p);
}

// CalculateFlux1's sum (OSCARSSR.cc:2209-2346): Omega = (E 2 pi) / 4.1357e-15 as the
// reference writes it (not omega_of), C0 with sqrt(2 pi), the C2 of the other flux routes.
int run_flux1(Context& c, const TrajView& t, const double* obs_soa, int64_t ns, double energy_ev, double q, double cur,
              double* out, hipStream_t s) {
  if (!c.cus) HIPCHK(hipDeviceGetAttribute(&c.cus, hipDeviceAttributeMultiprocessorCount, 0));
  int const chunk = pick_chunk(t.n, ns, 32, (int64_t)c.cus * srs::flux1_fp64_blocks_per_cu());
  int const nch = nchunks(t.n, chunk);
  HIPCHK(c.part.ensure((size_t)nch * 6 * ns * sizeof(double)));
  EnergySet es{};
  es.ne = 1;
  es.omega[0] = energy_ev * kTwoPi / 4.1357e-15;
  es.kw[0] = c0_of(q) * t.dt;
  {
    KernelTimer timer(c, s);
    HIPCHK(srs::launch_flux1_fp64(t, obs_soa, (int)ns, es.omega[0], chunk, c.part.as<double>(), s));
  }
  HIPCHK(srs::launch_flux_reduce(c.part.as<double>(), nch, (int)ns, es, c2_of(q, cur), out, s));
  return SRS_OK;
}

int run_spectrum(Context& c, const TrajView& t, const double obs[3], const double* energies_ev, int64_t ne, double q,
                 double cur, int precision, double* out, hipStream_t s) {
  double const c0 = c0_of(q), c2 = c2_of(q, cur);
  int const mixed = precision == SRS_PRECISION_MIXED ? 1 : 0;
  // per energy omega = (E Qe)/Hbar, c/omega and kw = (C0 omega) dt; per run the
  // step and |tau| bound of the phase correction
  int64_t const nr = (ne + srs::kSpecRun - 1) / srs::kSpecRun;
  std::vector<double> h(3 * ne + 2 * nr);
  for (int64_t e = 0; e < ne; ++e) {
    double const w = omega_of(energies_ev[e]);
    h[e] = w;
    h[ne + e] = kC / w;
    h[2 * ne + e] = (c0 * w) * t.dt;
  }
  bool uniform = srs::spectrum_runs_uniform(h.data(), ne);
  if (uniform) srs::spectrum_run_tables(h.data(), ne, mixed, h.data() + 3 * ne, h.data() + 3 * ne + nr);
  HIPCHK(c.energies.ensure(h.size() * sizeof(double)));
  double* d = c.energies.as<double>();
  HIPCHK(hipMemcpyAsync(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s));
  srs::SpecTables const tab{d, d + ne, d + 3 * ne, d + 3 * ne + nr};
  if (!c.cus) HIPCHK(hipDeviceGetAttribute(&c.cus, hipDeviceAttributeMultiprocessorCount, 0));
  for (;;) {
    srs::SpectrumPlan const pl = srs::plan_spectrum(t.n, ne, uniform, mixed, c.cus);
    HIPCHK(c.part.ensure((size_t)pl.nchunk * 6 * ne * sizeof(double)));
    *c.spec_range = 0;
    {
      KernelTimer timer(c, s);
      HIPCHK(srs::launch_spectrum(t, obs, tab, (int)ne, pl, mixed, c.part.as<double>(), c.spec_range, s));
    }
    HIPCHK(srs::launch_spectrum_fold(c.part.as<double>(), pl.nchunk, (int)ne, d + 2 * ne, c2, out, s));
    // the host-side energy table must outlive the async copy; the flag is read after
    HIPCHK(hipStreamSynchronize(s));
    if (!uniform || *(volatile int*)c.spec_range == 0) break;
    uniform = false;  // beyond the correction's range: one energy per lane
  }
  return SRS_OK;
}

int run_pd(Context& c, const TrajView& t, const double* obs_soa, const double* nrm_soa, int64_t ns, int directional,
           double q, double cur, int precision, double* out, hipStream_t s) {
  if (!c.cus) HIPCHK(hipDeviceGetAttribute(&c.cus, hipDeviceAttributeMultiprocessorCount, 0));
  int const per_cu = precision == SRS_PRECISION_MIXED ? srs::pd_mixed_blocks_per_cu(directional)
                                                      : srs::pd_fp64_blocks_per_cu();
  int const chunk = pick_chunk(t.n, ns, 32, (int64_t)c.cus * per_cu);
  int const nch = nchunks(t.n, chunk);
  HIPCHK(c.part.ensure((size_t)nch * ns * sizeof(double)));
  if (precision == SRS_PRECISION_MIXED) {
    HIPCHK(c.trajf.ensure(srs::traj_f32_floats(t.n, 1) * sizeof(float)));
    HIPCHK(srs::launch_traj_to_f32(t, 1, c.trajf.as<float>(), s));
    KernelTimer timer(c, s);
    HIPCHK(srs::launch_pd_mixed(t, c.trajf.as<float>(), obs_soa, nrm_soa, (int)ns, directional, chunk,
                                c.part.as<double>(), s));
  } else {
    KernelTimer timer(c, s);
    HIPCHK(srs::launch_pd_fp64(t, obs_soa, nrm_soa, (int)ns, directional, chunk, c.part.as<double>(), s));
  }
  HIPCHK(srs::launch_pd_reduce(c.part.as<double>(), nch, (int)ns, pd_scale_of(q, cur, t.dt), directional, out, s));
  return SRS_OK;
}

int stage_points(Context& c, Buf& b, const double* aos, int64_t ns) {
  std::vector<double> soa(3 * ns);
  for (int64_t j = 0; j < ns; ++j)
    for (int k = 0; k < 3; ++k) soa[k * ns + j] = aos[3 * j + k];
  HIPCHK(b.ensure(3 * ns * sizeof(double)));
  HIPCHK(hipMemcpy(b.p, soa.data(), 3 * ns * sizeof(double), hipMemcpyHostToDevice));
  return SRS_OK;
}

int check_common(int64_t nout, int precision) {
  if (nout < 1) return fail(SRS_ERR_INVALID, "no output points");
  if (nout > (int64_t)1 << 30) return fail(SRS_ERR_INVALID, "too many output points");
  if (precision != SRS_PRECISION_FP64 && precision != SRS_PRECISION_MIXED)
    return fail(SRS_ERR_INVALID, "precision must be SRS_PRECISION_FP64 or SRS_PRECISION_MIXED");
  return SRS_OK;
}

// ---- batched multi-particle path -------------------------------------------------

int particle_grid(const srs_particle& p, double* dt, int64_t* nb, int64_t* nf) {
  // OSCARSSR.cc:574-598
  if (p.ct_start > p.t0) return fail(SRS_ERR_INVALID, "start time is greater than T0");
  if (p.ct_start >= p.ct_stop) return fail(SRS_ERR_INVALID, "start time is greater than stop time");
  if (p.npoints < 2) return fail(SRS_ERR_INVALID, "number of trajectory points must be >= 2");
  double const d = ((p.ct_stop - p.ct_start) / kC / (size_t)(p.npoints - 1));
  size_t const f = 1 + (p.ct_stop - p.t0) / kC / d;
  size_t const b = (p.t0 - p.ct_start) / kC / d;
  *dt = d;
  *nb = (int64_t)b;
  *nf = (int64_t)f;
  return SRS_OK;
}

// A particle's charge and current: its own, or the call's for a particle with charge 0.
void particle_charge(const srs_particle& p, double charge, double current, double* q, double* cur) {
  *q = p.charge != 0 ? p.charge : charge;
  *cur = p.charge != 0 ? p.current : current;
}

int build_particles(const srs_particle* p, int64_t np, double charge, double current,
                    std::vector<srs::BatchParticle>& out, int64_t* total) {
  out.resize(np);
  int64_t off = 0;
  for (int64_t i = 0; i < np; ++i) {
    srs::BatchParticle& q = out[i];
    double qq, ii;
    particle_charge(p[i], charge, current, &qq, &ii);
    q.c0 = c0_of(qq);
    q.c2 = c2_of(qq, ii);
    q.pdq = pd_qfac_of(qq, ii);
    for (int k = 0; k < 3; ++k) {
      q.x0[k] = p[i].x0[k];
      q.beta0[k] = p[i].beta0[k];
    }
    q.qmg = p[i].qmg;
    TRY(particle_grid(p[i], &q.dt, &q.nb, &q.nf));
    q.off = off;
    q.pole_base = 0;
    off += q.nb + q.nf;
  }
  *total = off;
  if (off > ((int64_t)1 << 40)) return fail(SRS_ERR_INVALID, "too many trajectory points");
  return SRS_OK;
}

int stage_undulators(Context& c, const srs_field* und, int nund, hipStream_t s) {
  if (nund < 1 || !und) return fail(SRS_ERR_INVALID, "no magnetic or electric field given");
  std::vector<srs::DevUndulator> du(nund);
  std::vector<double> poles;
  // field maps: their values go to one device buffer, one after the other
  size_t grid_doubles = 0;
  std::vector<size_t> grid_off(nund, 0);
  for (int u = 0; u < nund; ++u) {
    srs_field const& h = und[u];
    if (h.kind != SRS_FIELD_GRID) continue;
    if (!h.grid) return fail(SRS_ERR_INVALID, "field map values are NULL");
    int64_t np = 1;
    int ndim = 0;
    for (int k = 0; k < 3; ++k) {
      if (h.grid_n[k] < 1) return fail(SRS_ERR_INVALID, "invalid number of points in at least one dimension");
      np *= h.grid_n[k];
      ndim += h.grid_n[k] > 1;
    }
    if (ndim == 0) return fail(SRS_ERR_INVALID, "invalid dimensions");
    if (np > ((int64_t)1 << 31)) return fail(SRS_ERR_INVALID, "field map larger than 2^31 points");
    grid_off[u] = grid_doubles;
    grid_doubles += 3 * (size_t)np;
  }
  if (grid_doubles) HIPCHK(c.b_grid.ensure(grid_doubles * sizeof(double)));
  for (int u = 0; u < nund; ++u) {
    srs::DevUndulator& d = du[u];
    srs_field const& h = und[u];
    d = srs::DevUndulator{};
    d.kind = h.kind;
    if (h.target != SRS_FIELD_MAGNETIC && h.target != SRS_FIELD_ELECTRIC)
      return fail(SRS_ERR_INVALID, "field target must be SRS_FIELD_MAGNETIC or SRS_FIELD_ELECTRIC");
    if (h.target == SRS_FIELD_ELECTRIC && h.pole_delta)
      return fail(SRS_ERR_INVALID, "pole errors apply to magnetic undulators only");
    d.target = h.target;
    d.pole_off = -1;
    if (h.kind == SRS_FIELD_GRID) {
      if (h.pole_delta) return fail(SRS_ERR_INVALID, "pole errors apply to undulators only");
      size_t const np = (size_t)h.grid_n[0] * h.grid_n[1] * h.grid_n[2];
      double* const dst = c.b_grid.as<double>() + grid_off[u];
      HIPCHK(hipMemcpyAsync(dst, h.grid, 3 * np * sizeof(double), hipMemcpyHostToDevice, s));
      d.gdata = dst;
      for (int k = 0; k < 3; ++k) {
        d.center[k] = h.center[k];
        d.rc[k] = std::cos(h.rotations[k]);
        d.rs[k] = std::sin(h.rotations[k]);
        d.gn[k] = h.grid_n[k];
        d.gstart[k] = h.grid_start[k];
        d.gstep[k] = h.grid_step[k];
        d.gstop[k] = h.grid_stop[k];
      }
      continue;
    }
    if (h.kind == SRS_FIELD_GAUSSIAN || h.kind == SRS_FIELD_UNIFORM_BOX) {
      if (h.pole_delta) return fail(SRS_ERR_INVALID, "pole errors apply to undulators only");
      for (int k = 0; k < 3; ++k) {
        d.field[k] = h.field[k];
        d.center[k] = h.center[k];
        // Gaussian: sigma; box: half width, -1 for an unbounded axis
        d.w[k] = h.kind == SRS_FIELD_GAUSSIAN ? h.width[k] : (h.width[k] > 0 ? std::fabs(h.width[k] / 2.) : -1.0);
        d.rc[k] = std::cos(h.rotations[k]);  // the values TVector3D::RotateSelf* computes per call
        d.rs[k] = std::sin(h.rotations[k]);
      }
      continue;
    }
    if (h.kind != SRS_FIELD_UNDULATOR) return fail(SRS_ERR_INVALID, "unknown field kind");
    double const pm = std::sqrt(h.period[0] * h.period[0] + h.period[1] * h.period[1] + h.period[2] * h.period[2]);
    if (!(pm > 0)) return fail(SRS_ERR_INVALID, "undulator period must be a nonzero vector");
    bool const poled = h.pole_delta != nullptr;
    for (int k = 0; k < 3; ++k) {
      d.field[k] = h.field[k];
      d.u[k] = h.period[k] / pm;
      d.center[k] = poled ? 0.0 : h.center[k];
    }
    double const lam = pm, len = pm * (h.nperiods + 2);
    double const ps = poled ? 0.0 : h.phase * pm / kTwoPi;
    d.lam = lam;
    d.rlam = 1.0 / lam;
    d.taper = poled ? 0.0 : h.taper;
    d.ps = ps;
    d.lo0 = -len / 2. + ps;
    d.hi0 = len / 2. + ps;
    d.lo1 = -len / 2. + ps + lam;
    d.hi1 = len / 2. + ps - lam;
    d.lo2 = -len / 2. + ps + lam / 2.;
    d.hi2 = len / 2. + ps - lam / 2.;
    d.nperiods = h.nperiods;
    d.pole_off = -1;
    if (poled) {
      d.pole_off = (int)poles.size();
      poles.insert(poles.end(), h.pole_delta, h.pole_delta + 2 * h.nperiods);
    }
  }
  c.pole_count = (int)poles.size();
  if (poles.empty()) poles.push_back(0.0);
  c.und_local = false;
  bool has_b = false, has_e = false;
  for (int u = 0; u < nund; ++u) {
    has_e = has_e || und[u].target == SRS_FIELD_ELECTRIC;
    has_b = has_b || und[u].target != SRS_FIELD_ELECTRIC;
    c.und_local = c.und_local || und[u].kind != SRS_FIELD_UNDULATOR || und[u].target == SRS_FIELD_ELECTRIC;
  }
  c.und_force = has_e ? (has_b ? srs::kForceEB : srs::kForceE) : srs::kForceB;  // OSCARSSR.cc:690-703
  HIPCHK(c.b_und.ensure(du.size() * sizeof(srs::DevUndulator)));
  HIPCHK(c.b_poles.ensure(poles.size() * sizeof(double)));
  HIPCHK(hipMemcpyAsync(c.b_und.p, du.data(), du.size() * sizeof(srs::DevUndulator), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c.b_poles.p, poles.data(), poles.size() * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));  // host staging vectors go out of scope
  return SRS_OK;
}

// Packed trajectory buffers of np particles with `total` points: FP64 rows (A/c when
// with_a), the FP32 rows, then np ints of per-particle "not along z" flags (the device
// RK4 writes them).
int stage_batch_traj(Context& c, int64_t total, int64_t np, bool with_a, srs::BatchTraj* bt) {
  HIPCHK(c.b_traj.ensure((size_t)(with_a ? 9 : 6) * total * sizeof(double)));
  HIPCHK(c.b_trajf.ensure(((size_t)3 * total + (size_t)np) * sizeof(float)));
  double* d = c.b_traj.as<double>();
  float* f = c.b_trajf.as<float>();
  size_t const n = (size_t)total;
  *bt = srs::BatchTraj{d,          d + n,      d + 2 * n, d + 3 * n, d + 4 * n, d + 5 * n, with_a ? d + 6 * n : nullptr,
                       with_a ? d + 7 * n : nullptr, with_a ? d + 8 * n : nullptr, f, f + n, f + 2 * n,
                       reinterpret_cast<int*>(f + 3 * n)};
  return SRS_OK;
}

struct BatchRun {
  int nund;
  srs::BatchTraj bt;
  srs::BatchEnergies be;
  int np, ns, ne;
  bool mixed;
  double weight;
  int64_t max_leg = 0;                         // longest RK4 leg over the particles
  srs::Checkpoints ck{nullptr, 1, 0};          // RK4 checkpoints (gradient forward)
  double* field_out = nullptr;                 // keep the field sums here (gradient forward)
};

// Time segments of the pipeline: one for short trajectories, else up to kMaxSegments
// of >= 2048 steps, so the device RK4 (latency-bound, a few waves) of segment s + 1
// runs on stream2 while the integral of segment s runs on the main stream.  The first
// segment is a quarter of the others (at least 512 steps): the integral waits for it
// alone, while every later RK4 segment runs beside the integral of the one before.
struct Segments {
  int n = 1;
  int64_t first = 0, len = 0;  // steps of segment 0 and of each later one
  int64_t lo(int k) const { return k == 0 ? 0 : first + (int64_t)(k - 1) * len; }
  int64_t hi(int k) const { return first + (int64_t)k * len; }
  int64_t longest() const { return std::max(first, len); }
};
Segments segments_of(int64_t max_leg) {
  Segments g;
  int const n = max_leg <= 4096 ? 1 : (int)std::min<int64_t>(Context::kMaxSegments, (max_leg + 2047) / 2048);
  if (n == 1) {
    g.first = g.len = max_leg;
    return g;
  }
  int64_t const L = (max_leg + n - 1) / n;
  g.n = n;
  g.first = std::max<int64_t>(512, L / 4);
  g.len = (max_leg - g.first + n - 2) / (n - 1);  // hi(n - 1) >= max_leg
  return g;
}

// Queue the whole pipeline: RK4 segments, persistent flux per segment, particle reduction.
int enqueue_batch(Context& c, BatchRun const& r, srs::BatchConfig const& cfg, int np_run, hipStream_t s,
                  bool reduce = true) {
  int const tile_pts = srs::kBlock * cfg.ppl;
  int const ntiles = (r.ns + tile_pts - 1) / tile_pts;
  int const ngroups = (r.ne + cfg.ne_per_item - 1) / cfg.ne_per_item;
  unsigned const nitems = (unsigned)((int64_t)np_run * ntiles * ngroups);
  Segments const g = segments_of(r.max_leg);
  int const nseg = g.n;
  // leave room for the RK4 blocks next to the persistent grid
  int const reserve = nseg > 1 ? (np_run + 63) / 64 + 1 : 0;
  int nblocks = std::max(1, c.cus * srs::batch_max_blocks(cfg, r.mixed) - reserve);
  if ((unsigned)nblocks > nitems) nblocks = (int)nitems;
  auto* P = c.b_part.as<srs::BatchParticle>();
  double* field = r.field_out;
  if (nseg > 1) {
    HIPCHK(c.b_state.ensure((size_t)np_run * 24 * sizeof(double)));
    if (!field) {
      HIPCHK(c.b_field.ensure((size_t)np_run * r.ne * 6 * r.ns * sizeof(double)));
      field = c.b_field.as<double>();
    }
  }
  srs::RK4Seg rs{0, g.hi(0), nseg > 1 ? c.b_state.as<double>() : nullptr};
  if (nseg == 1) {
    HIPCHK(srs::launch_rk4_batch(c.b_und.as<srs::DevUndulator>(), r.nund, c.und_local, c.und_force, c.b_poles.as<double>(), c.pole_count, P, np_run, r.bt,
                                 r.ck, rs, s));
    HIPCHK(srs::launch_zero_counter(c.b_counter.as<unsigned>(), s));
    srs::FluxSeg const fs{0, g.hi(0), 1, 1, field != nullptr, field};
    HIPCHK(srs::launch_batch_flux(cfg, r.mixed, r.bt, P, c.b_obs.as<double>(), r.ns, r.be, nitems,
                                  c.b_counter.as<unsigned>(), c.b_flux_pp.as<double>(), fs, nblocks, s));
  } else {
    HIPCHK(hipEventRecord(c.seg_ev[Context::kMaxSegments], s));  // fork: the RK4 follows the staging on s
    HIPCHK(hipStreamWaitEvent(c.stream2, c.seg_ev[Context::kMaxSegments], 0));
    for (int k = 0; k < nseg; ++k) {
      rs.lo = g.lo(k);
      rs.hi = g.hi(k);
      HIPCHK(srs::launch_rk4_batch(c.b_und.as<srs::DevUndulator>(), r.nund, c.und_local, c.und_force, c.b_poles.as<double>(), c.pole_count, P, np_run, r.bt,
                                   r.ck, rs, c.stream2));
      HIPCHK(hipEventRecord(c.seg_ev[k], c.stream2));
    }
    for (int k = 0; k < nseg; ++k) {
      HIPCHK(hipStreamWaitEvent(s, c.seg_ev[k], 0));  // the last wait also joins stream2
      HIPCHK(srs::launch_zero_counter(c.b_counter.as<unsigned>(), s));
      srs::FluxSeg const fs{g.lo(k), g.hi(k), k == 0, k == nseg - 1, 1, field};
      HIPCHK(srs::launch_batch_flux(cfg, r.mixed, r.bt, P, c.b_obs.as<double>(), r.ns, r.be, nitems,
                                    c.b_counter.as<unsigned>(), c.b_flux_pp.as<double>(), fs, nblocks, s));
    }
  }
  if (reduce)
    HIPCHK(srs::launch_particle_reduce(c.b_flux_pp.as<double>(), np_run, r.ne * r.ns, r.weight, c.b_out.as<double>(),
                                       s));
  return SRS_OK;
}

std::string batch_key(BatchRun const& r, srs::BatchConfig const& cfg, Context const& c) {
  char buf[768];
  // every value a captured launch holds by value: shapes, tile shape, the field list's
  // length and whether it needs the general field evaluation (the RK4 instantiation
  // depends on it), the pole table length (the RK4's LDS copy of it), the CU count
  // behind the persistent grid size, the flux prefactor c2 and the weight, and the
  // address of every buffer a launch reads or writes
  snprintf(buf, sizeof buf,
           "%d/%d/%d/%d/%d/%d/%d/%d/%d/%d/%d/%lld/%.17g/%.17g/%p/%p/%p/%p/%p/%p/%p/%p/%p/%p/%p/%p/%p/%p", r.np, r.ns,
           r.ne, (int)r.mixed, cfg.ne_per_item, cfg.tile, cfg.ppl, r.nund, (int)c.und_local + 2 * c.und_force,
           c.pole_count, c.cus, (long long)r.max_leg,
           r.weight, r.be.c2, c.b_traj.p, c.b_trajf.p, c.b_flux_pp.p, c.b_out.p, c.b_obs.p, c.b_energies.p,
           c.b_state.p, c.b_field.p, c.b_poles.p, c.b_part.p, c.b_counter.p, c.b_und.p,
           (void*)r.bt.bz, (void*)r.bt.tflag);  // row offsets within the packed buffers follow the total point count
  return buf;
}

// Power density of np particles (or realisations) already staged in b_part / b_und /
// b_poles with trajectory buffers bt (accelerations included), observation points in
// b_obs and normals in g_nrm: the device RK4 in time segments on stream2 overlapped
// with pd_batch_kernel per segment on s (adding in segment order), then the ordered reduction into
// raw (np * ns) and b_flux_pp (np * ns, the per-particle power density).
int enqueue_pd_batch(Context& c, srs::BatchTraj const& bt, int np, int nund, int64_t max_leg, int64_t max_pts,
                     srs::Checkpoints const& ck, int ns, int dir, double* raw, hipStream_t s,
                     bool mixed = false) {
  Segments const g = segments_of(max_leg);
  int const nseg = g.n;
  int const nch = srs::pd_batch_chunks(nseg > 1 ? 2 * g.longest() : max_pts, ns, np);
  HIPCHK(c.g_part.ensure((size_t)np * nch * ns * sizeof(double)));
  if (nseg > 1) HIPCHK(c.b_state.ensure((size_t)np * 24 * sizeof(double)));
  auto* Pd = c.b_part.as<srs::BatchParticle>();
  auto* U = c.b_und.as<srs::DevUndulator>();
  double* const poles = c.b_poles.as<double>();
  if (nseg == 1) {
    HIPCHK(srs::launch_rk4_batch(U, nund, c.und_local, c.und_force, poles, c.pole_count, Pd, np, bt, ck, srs::RK4Seg{0, INT64_MAX, nullptr}, s));
    HIPCHK(srs::launch_pd_batch(bt, Pd, np, c.b_obs.as<double>(), c.g_nrm.as<double>(), ns, dir, 0, INT64_MAX, nch, 0,
                                c.g_part.as<double>(), mixed, s));
  } else {
    HIPCHK(hipEventRecord(c.seg_ev[Context::kMaxSegments], s));  // fork: the RK4 follows the staging on s
    HIPCHK(hipStreamWaitEvent(c.stream2, c.seg_ev[Context::kMaxSegments], 0));
    for (int k = 0; k < nseg; ++k) {
      HIPCHK(srs::launch_rk4_batch(U, nund, c.und_local, c.und_force, poles, c.pole_count, Pd, np, bt, ck,
                                   srs::RK4Seg{g.lo(k), g.hi(k), c.b_state.as<double>()}, c.stream2));
      HIPCHK(hipEventRecord(c.seg_ev[k], c.stream2));
    }
    for (int k = 0; k < nseg; ++k) {
      HIPCHK(hipStreamWaitEvent(s, c.seg_ev[k], 0));  // the last wait also joins stream2
      HIPCHK(srs::launch_pd_batch(bt, Pd, np, c.b_obs.as<double>(), c.g_nrm.as<double>(), ns, dir, g.lo(k), g.hi(k),
                                  nch, k > 0, c.g_part.as<double>(), mixed, s));
    }
  }
  HIPCHK(srs::launch_pd_batch_reduce(c.g_part.as<double>(), nch, ns, np, Pd, dir, raw,
                                     c.b_flux_pp.as<double>(), s));
  return SRS_OK;
}

// ---- pole-error gradient: shared parts of the forward entry points ----------------
int pole_validate(const srs_field* und, int nund, int poled, const srs_particle* particle, int64_t nreal,
                  const double* deltas, const double* obs_xyz, int64_t ns, const double* target, int checkpoint_every,
                  double* loss_out, int64_t* tape_out) {
  TRY(check_common(ns, SRS_PRECISION_FP64));
  if (!und || nund < 1) return fail(SRS_ERR_INVALID, "no magnetic field (undulator) given");
  if (poled < 0 || poled >= nund) return fail(SRS_ERR_INVALID, "poled undulator index out of range");
  if (und[poled].nperiods < 1) return fail(SRS_ERR_INVALID, "the pole-scaled undulator needs nperiods >= 1");
  for (int u = 0; u < nund; ++u)
    if (u != poled && und[u].pole_delta)
      return fail(SRS_ERR_INVALID, "only the differentiated undulator may carry pole errors");
  for (int u = 0; u < nund; ++u)
    if (und[u].kind != SRS_FIELD_UNDULATOR || und[u].target != SRS_FIELD_MAGNETIC)
      return fail(SRS_ERR_INVALID, "the pole-error gradient supports magnetic undulator fields only");
  if (!particle) return fail(SRS_ERR_INVALID, "no particle");
  if (nreal < 1 || nreal > 65536) return fail(SRS_ERR_INVALID, "number of realisations must be in 1..65536");
  if (!deltas || !target || !loss_out || !tape_out || !obs_xyz) return fail(SRS_ERR_INVALID, "NULL argument");
  if (checkpoint_every < 1) return fail(SRS_ERR_INVALID, "checkpoint interval must be >= 1");
  // all undulators must share one period direction (the adjoint projects dB/dX on it)
  double u0[3];
  for (int u = 0; u < nund; ++u) {
    double const* pr = und[u].period;
    double const pm = std::sqrt(pr[0] * pr[0] + pr[1] * pr[1] + pr[2] * pr[2]);
    if (!(pm > 0)) return fail(SRS_ERR_INVALID, "undulator period must be a nonzero vector");
    for (int k = 0; k < 3; ++k) {
      double const uk = pr[k] / pm;
      if (u == 0) u0[k] = uk;
      else if (uk != u0[k]) return fail(SRS_ERR_INVALID, "the gradient needs undulators with one period direction");
    }
  }
  return SRS_OK;
}

struct PoleRun {
  std::vector<srs::BatchParticle> P;
  int64_t total = 0, nmax = 0, stride = 0;
  int npoles = 0, every = 1;
  srs::BatchTraj bt{};
  srs::Checkpoints ck{};
};

// Realisations (copies of the particle, pole_base = r * npoles), fields, deltas,
// observation points, target, trajectory buffers (with accelerations when with_a)
// and checkpoints, all queued on s.  Invalidates the tape.
int pole_stage(Context& c, const srs_field* und, int nund, int poled, const srs_particle* particle, int64_t nreal,
               const double* deltas, double charge, double current, const double* obs_xyz, int64_t ns,
               const double* target, int every, bool with_a, PoleRun& pr, hipStream_t s) {
  pr.npoles = 2 * und[poled].nperiods;
  pr.every = every;
  {
    std::vector<srs_particle> rows((size_t)nreal, *particle);
    TRY(build_particles(rows.data(), nreal, charge, current, pr.P, &pr.total));
    for (int64_t r = 0; r < nreal; ++r) pr.P[r].pole_base = r * pr.npoles;
  }
  pr.nmax = pr.P[0].nb + pr.P[0].nf;
  pr.stride = (pr.P[0].nf + every - 1) / every + (pr.P[0].nb + every - 1) / every;
  c.tape.valid = false;
  if (!c.cus) HIPCHK(hipDeviceGetAttribute(&c.cus, hipDeviceAttributeMultiprocessorCount, 0));
  {
    std::vector<srs_field> uu(und, und + nund);
    uu[poled].pole_delta = deltas;  // pole_off 0 into the realisations' delta array
    TRY(stage_undulators(c, uu.data(), nund, s));
  }
  size_t const nd = (size_t)nreal * pr.npoles;
  c.pole_count = pr.npoles;  // each realisation's own table (pole_base = r * npoles)
  HIPCHK(c.b_poles.ensure(nd * sizeof(double)));
  HIPCHK(hipMemcpyAsync(c.b_poles.p, deltas, nd * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(c.b_part.ensure(nreal * sizeof(srs::BatchParticle)));
  HIPCHK(hipMemcpyAsync(c.b_part.p, pr.P.data(), nreal * sizeof(srs::BatchParticle), hipMemcpyHostToDevice, s));
  TRY(stage_points(c, c.b_obs, obs_xyz, ns));
  HIPCHK(c.g_target.ensure(ns * sizeof(double)));
  HIPCHK(hipMemcpyAsync(c.g_target.p, target, ns * sizeof(double), hipMemcpyHostToDevice, s));
  TRY(stage_batch_traj(c, pr.total, nreal, with_a, &pr.bt));
  HIPCHK(c.g_ckpt.ensure((size_t)nreal * pr.stride * 6 * sizeof(double)));
  pr.ck = srs::Checkpoints{c.g_ckpt.as<double>(), every, pr.stride};
  HIPCHK(c.b_flux_pp.ensure((size_t)nreal * ns * sizeof(double)));
  HIPCHK(c.g_loss.ensure((size_t)nreal * sizeof(double)));
  return SRS_OK;
}

// Loss of every realisation from the per-realisation maps in b_flux_pp; copies out
// the losses (and the maps when map_out != NULL) and waits.
int pole_finish(Context& c, PoleRun const&, int nreal, int ns, double* map_out, double* loss_out, hipStream_t s) {
  HIPCHK(srs::launch_loss(c.b_flux_pp.as<double>(), c.g_target.as<double>(), ns, nreal, c.g_loss.as<double>(), s));
  HIPCHK(hipMemcpyAsync(loss_out, c.g_loss.p, nreal * sizeof(double), hipMemcpyDeviceToHost, s));
  if (map_out)
    HIPCHK(hipMemcpyAsync(map_out, c.b_flux_pp.p, (size_t)nreal * ns * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return SRS_OK;
}

Context::Tape& pole_tape(Context& c, PoleRun const& pr, int nreal, int ns, int nund, int poled) {
  Context::Tape& t = c.tape;
  t.id = ++c.tape_counter;
  t.np = nreal;
  t.ns = ns;
  t.nund = nund;
  t.poled = poled;
  t.npoles = pr.npoles;
  t.every = pr.every;
  t.nf = pr.P[0].nf;
  t.nb = pr.P[0].nb;
  {
    // as many segments per pass as a 4 GiB record budget holds (all of them at C4 sizes)
    int64_t const nseg = (std::max(t.nf, t.nb) + t.every - 1) / t.every;
    double const per_seg = (double)srs::adjoint_record_doubles(t.every, nreal) * sizeof(double);
    t.seg_per_pass = std::max<int64_t>(1, std::min<int64_t>(nseg, (int64_t)(4.0 * (1 << 30) / per_seg)));
  }
  t.total = pr.total;
  t.nmax = pr.nmax;
  t.bt = pr.bt;
  t.ck = pr.ck;
  t.valid = true;
  return t;
}

// Device records of nbeams beams and their cumulative weights (the running sums of
// TParticleBeamContainer::AddNewParticleBeam); with grid, also each beam's batch
// prefactors (its charge and current) and time grid (OSCARSSR.cc:574-598).
int beam_records(const srs_beam* beams, const double* weights, int nbeams, bool grid, std::vector<srs::DevBeam>& db,
                 std::vector<double>& wsum) {
  if (!beams || (nbeams > 1 && !weights)) return fail(SRS_ERR_INVALID, "NULL argument");
  if (nbeams < 1) return fail(SRS_ERR_INVALID, "no beam defined");
  db.assign(nbeams, srs::DevBeam{});
  wsum.assign(nbeams, 0.0);
  double acc = 0;
  for (int k = 0; k < nbeams; ++k) {
    srs_beam const& beam = beams[k];
    if (!(beam.mass > 0)) return fail(SRS_ERR_INVALID, "beam particle mass must be > 0");
    double const mgev = 1e-9 * beam.mass * kC * kC / kQe;  // TSRS::kgToGeV, TOSCARSSR.h:68
    if (!(beam.e0 > mgev)) return fail(SRS_ERR_INVALID, "beam energy must exceed the particle's rest energy");
    if (!(beam.sigma_e >= 0)) return fail(SRS_ERR_INVALID, "'sigma_energy_GeV' cannot be less than zero");
    if (beam.npoints < 2) return fail(SRS_ERR_INVALID, "npoints must be >= 2");
    double const w = nbeams > 1 ? weights[k] : 1.0;
    if (!(w >= 0)) return fail(SRS_ERR_INVALID, "beam weights must be >= 0");
    acc += w;
    wsum[k] = acc;
    srs::DevBeam& b = db[k];
    b.e0 = beam.e0;
    b.sigma_e = beam.sigma_e;
    b.mgev = mgev;
    b.q = beam.charge;
    b.m = beam.mass;
    b.cur = beam.current != 0 ? beam.current : beam.charge;  // TParticleA.cc:258-262
    for (int i = 0; i < 3; ++i) {
      b.x0[i] = beam.x0[i];
      b.u0[i] = beam.u0[i];
      b.h[i] = beam.hdir[i];
      b.v[i] = beam.vdir[i];
      b.at[i] = beam.sigma_at[i];
    }
    for (int i = 0; i < 2; ++i) {
      b.su[i] = beam.sigma_u[i];
      b.sp[i] = beam.sigma_up[i];
    }
    b.t0 = beam.t0;
    b.ct_start = beam.ct_start;
    b.ct_stop = beam.ct_stop;
    b.npoints = beam.npoints;
    if (grid) {
      if (!(beam.charge != 0)) return fail(SRS_ERR_INVALID, "beam charge must be nonzero");
      b.c0 = c0_of(b.q);
      b.c2 = c2_of(b.q, b.cur);
      b.pdq = pd_qfac_of(b.q, b.cur);
      srs_particle g{};
      g.t0 = beam.t0;
      g.ct_start = beam.ct_start;
      g.ct_stop = beam.ct_stop;
      g.npoints = beam.npoints;
      TRY(particle_grid(g, &b.dt, &b.gnb, &b.gnf));
    }
  }
  if (!(acc > 0)) return fail(SRS_ERR_INVALID, "beam weights must have a positive sum");
  return SRS_OK;
}

// The batch flux pipeline for np particle records already in b_part (total trajectory
// points, longest leg max_leg) and the field list in b_und: observation points and
// energies staged, the tile shape (autotuned on request), then the pipeline eagerly or
// as a replayed hipGraph, and the result copied to flux_out.  Called with c.mu held.
int batch_flux_staged(Context& c, hipStream_t s, int nund, int64_t np, int64_t total, int64_t max_leg, double charge,
                      double current, const double* obs_xyz, int64_t ns, const double* energies_ev, int64_t ne,
                      double weight, int precision, int flags, double* flux_out) {
  TRY(stage_points(c, c.b_obs, obs_xyz, ns));
  double const c0 = c0_of(charge);
  std::vector<double> eh(5 * ne);
  for (int64_t e = 0; e < ne; ++e) {
    double const w = omega_of(energies_ev[e]);
    eh[e] = w;
    eh[ne + e] = kC / w;
    eh[2 * ne + e] = c0 * w;
    eh[3 * ne + e] = w / kTwoPi;
    eh[4 * ne + e] = eh[3 * ne + e] / kC;
  }
  HIPCHK(c.b_energies.ensure(5 * ne * sizeof(double)));
  HIPCHK(hipMemcpyAsync(c.b_energies.p, eh.data(), 5 * ne * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(c.b_flux_pp.ensure((size_t)np * ne * ns * sizeof(double)));
  HIPCHK(c.b_out.ensure((size_t)ne * ns * sizeof(double)));
  HIPCHK(c.b_counter.ensure(sizeof(unsigned)));
  BatchRun r;
  r.nund = nund;
  TRY(stage_batch_traj(c, total, np, false, &r.bt));
  double* ed = c.b_energies.as<double>();
  r.be = srs::BatchEnergies{ed, ed + ne, ed + 2 * ne, ed + 3 * ne, ed + 4 * ne, (int)ne, c2_of(charge, current)};
  r.np = (int)np;
  r.ns = (int)ns;
  r.ne = (int)ne;
  r.mixed = precision == SRS_PRECISION_MIXED;
  r.weight = weight;
  r.max_leg = max_leg;
  // the segmented pipeline's buffers are sized here: nothing may allocate while a
  // graph is being captured
  HIPCHK(c.b_state.ensure((size_t)np * 24 * sizeof(double)));
  HIPCHK(c.b_field.ensure((size_t)np * ne * 6 * ns * sizeof(double)));
  HIPCHK(hipStreamSynchronize(s));  // host staging vectors

  // tile shape: default, or the autotuned one for this problem signature
  srs::BatchConfig cfg{(int)std::min<int64_t>(ne, 4) == 3 ? 3 : (ne >= 4 ? 4 : (int)ne), 128};
  if (!srs::batch_config_valid(cfg)) cfg = srs::BatchConfig{1, 128};
  if (c.b_fixed.ne_per_item > 0) {
    if (c.b_fixed.ppl == 2 && !r.mixed)
      return fail(SRS_ERR_INVALID, "two observation points per lane need the mixed precision form");
    cfg = c.b_fixed;
  }
  char sig[128];
  snprintf(sig, sizeof sig, "%d/%lld/%lld/%lld", (int)r.mixed, (long long)((ns + 255) / 256), (long long)ne,
           (long long)(total / np));
  if (flags & SRS_BATCH_AUTOTUNE) {
    auto it = c.b_tuned.find(sig);
    if (it != c.b_tuned.end()) {
      cfg = it->second;
    } else {
      // time each candidate on a subset of the particles (same RK4 and inputs), large
      // enough that every candidate runs many items per persistent block
      static const srs::BatchConfig cands[] = {{1, 64},  {1, 128}, {1, 256}, {2, 128},    {2, 256},   {3, 64},
                                               {3, 128}, {3, 256}, {4, 128}, {4, 256},    {1, 128, 2}, {1, 256, 2}};
      int const nsub = (int)std::min<int64_t>(np, 256);
      hipEvent_t e0, e1;
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
      float best = 1e30f;
      bool const tsave = c.timing;
      c.timing = false;
      for (auto const& k : cands) {
        if (k.ne_per_item > ne || (k.ne_per_item > 1 && ne % k.ne_per_item != 0 && k.ne_per_item != ne)) continue;
        if (!srs::batch_config_valid(k) || (k.ppl == 2 && !r.mixed)) continue;
        TRY(enqueue_batch(c, r, k, nsub, s));  // warm-up
        HIPCHK(hipEventRecord(e0, s));
        TRY(enqueue_batch(c, r, k, nsub, s));
        HIPCHK(hipEventRecord(e1, s));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
          best = ms;
          cfg = k;
        }
      }
      c.timing = tsave;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      c.b_tuned[sig] = cfg;
    }
  }
  c.b_last = cfg;

  if (flags & SRS_BATCH_GRAPH) {
    std::string const key = batch_key(r, cfg, c);
    if (key != c.b_key || !c.b_exec) {
      if (c.b_exec) (void)hipGraphExecDestroy(c.b_exec);
      if (c.b_graph) (void)hipGraphDestroy(c.b_graph);
      c.b_exec = nullptr;
      c.b_graph = nullptr;
      bool const tsave = c.timing;
      c.timing = false;  // no timing events inside the captured graph
      HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      int const rc = enqueue_batch(c, r, cfg, r.np, s);
      hipGraph_t g = nullptr;
      hipError_t const ec = hipStreamEndCapture(s, &g);
      c.timing = tsave;
      if (rc != SRS_OK) return rc;
      if (ec != hipSuccess) return fail(SRS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
      c.b_graph = g;
      HIPCHK(hipGraphInstantiate(&c.b_exec, c.b_graph, nullptr, nullptr, 0));
      c.b_key = key;
    }
    KernelTimer timer(c, s);
    HIPCHK(hipGraphLaunch(c.b_exec, s));
  } else {
    KernelTimer timer(c, s);
    TRY(enqueue_batch(c, r, cfg, r.np, s));
  }
  HIPCHK(hipMemcpyAsync(flux_out, c.b_out.p, (size_t)ne * ns * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return SRS_OK;
}


}  // namespace

extern "C" {

int srs_hip_device_count(void) {
  g_err.clear();
  return device_count_raw();
}

const char* srs_hip_last_error(void) { return g_err.c_str(); }

int srs_hip_spectrum(const srs_traj* traj, const double obs[3], const double* energies_ev, int64_t ne, double charge,
                     double current, int precision, double* flux_out) {
  g_err.clear();
  TRY(check_traj(traj, false));
  TRY(check_common(ne, precision));
  if (!obs || !energies_ev || !flux_out) return fail(SRS_ERR_INVALID, "NULL argument");
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  TRY(init(c));
  TrajView t;
  TRY(stage_traj(c, traj, false, &t));
  HIPCHK(c.out.ensure(ne * sizeof(double)));
  TRY(run_spectrum(c, t, obs, energies_ev, ne, charge, current, precision, c.out.as<double>(), c.stream));
  HIPCHK(hipMemcpyAsync(flux_out, c.out.p, ne * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return SRS_OK;
}

int srs_hip_flux(const srs_traj* traj, const double* obs_xyz, int64_t ns, const double* energies_ev, int64_t ne,
                 double charge, double current, int precision, double* flux_out) {
  g_err.clear();
  TRY(check_traj(traj, false));
  TRY(check_common(ns, precision));
  if (ne < 1) return fail(SRS_ERR_INVALID, "no photon energies");
  if (!obs_xyz || !energies_ev || !flux_out) return fail(SRS_ERR_INVALID, "NULL argument");
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  TRY(init(c));
  TrajView t;
  TRY(stage_traj(c, traj, false, &t));
  TRY(stage_points(c, c.obs, obs_xyz, ns));
  HIPCHK(c.out.ensure(ne * ns * sizeof(double)));
  TRY(run_flux(c, t, c.obs.as<double>(), ns, energies_ev, ne, charge, current, precision, c.out.as<double>(), c.stream));
  HIPCHK(hipMemcpyAsync(flux_out, c.out.p, ne * ns * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return SRS_OK;
}

int srs_hip_flux1(const srs_traj* traj, const double* obs_xyz, int64_t ns, double energy_ev, double charge,
                  double current, double* flux_out) {
  g_err.clear();
  TRY(check_traj(traj, true));
  TRY(check_common(ns, SRS_PRECISION_FP64));
  if (!obs_xyz || !flux_out) return fail(SRS_ERR_INVALID, "NULL argument");
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  TRY(init(c));
  TrajView t;
  TRY(stage_traj(c, traj, true, &t));
  TRY(stage_points(c, c.obs, obs_xyz, ns));
  HIPCHK(c.out.ensure(ns * sizeof(double)));
  TRY(run_flux1(c, t, c.obs.as<double>(), ns, energy_ev, charge, current, c.out.as<double>(), c.stream));
  HIPCHK(hipMemcpyAsync(flux_out, c.out.p, ns * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return SRS_OK;
}

int srs_hip_power_density(const srs_traj* traj, const double* obs_xyz, const double* normals_xyz, int64_t ns,
                          int directional, double charge, double current, int precision, double* pd_out) {
  g_err.clear();
  TRY(check_traj(traj, true));
  TRY(check_common(ns, precision));
  if (!obs_xyz || !normals_xyz || !pd_out) return fail(SRS_ERR_INVALID, "NULL argument");
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  TRY(init(c));
  TrajView t;
  TRY(stage_traj(c, traj, true, &t));
  TRY(stage_points(c, c.obs, obs_xyz, ns));
  TRY(stage_points(c, c.nrm, normals_xyz, ns));
  HIPCHK(c.out.ensure(ns * sizeof(double)));
  TRY(run_pd(c, t, c.obs.as<double>(), c.nrm.as<double>(), ns, directional ? 1 : 0, charge, current, precision,
             c.out.as<double>(), c.stream));
  HIPCHK(hipMemcpyAsync(pd_out, c.out.p, ns * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return SRS_OK;
}

int srs_hip_efield_time(const srs_traj* traj, const double obs[3], double charge, double* efield_out,
                        double* time_out) {
  g_err.clear();
  TRY(check_traj(traj, true));
  if (!obs || !efield_out || !time_out) return fail(SRS_ERR_INVALID, "NULL argument");
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  TRY(init(c));
  TrajView t;
  TRY(stage_traj(c, traj, true, &t));
  size_t const n = (size_t)traj->n;
  HIPCHK(c.out.ensure(4 * n * sizeof(double)));
  double* const d = c.out.as<double>();
  HIPCHK(srs::launch_efield_time(t, obs, charge / (kFourPi * kEps0), d, d + 3 * n, c.stream));  // OSCARSSR.cc:2661
  HIPCHK(hipMemcpyAsync(efield_out, d, 3 * n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipMemcpyAsync(time_out, d + 3 * n, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return SRS_OK;
}

static_assert(sizeof(srs::SampledParticle) == sizeof(srs_particle) &&
                  offsetof(srs::SampledParticle, qmg) == offsetof(srs_particle, qmg) &&
                  offsetof(srs::SampledParticle, npoints) == offsetof(srs_particle, npoints) &&
                  offsetof(srs::SampledParticle, charge) == offsetof(srs_particle, charge) &&
                  offsetof(srs::SampledParticle, current) == offsetof(srs_particle, current),
              "sampled rows are srs_particle rows");

int srs_hip_sample_beams(const srs_beam* beams, const double* weights, int nbeams, int64_t np, uint64_t seed,
                         uint64_t first, srs_particle* out) {
  g_err.clear();
  if (!beams || !out || (nbeams > 1 && !weights)) return fail(SRS_ERR_INVALID, "NULL argument");
  if (np < 1) return fail(SRS_ERR_INVALID, "no particles requested");
  std::vector<srs::DevBeam> db;
  std::vector<double> wsum;
  TRY(beam_records(beams, weights, nbeams, false, db, wsum));
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  TRY(init(c));
  // one buffer: beam records | cumulative weights | rows, each part 256-byte aligned
  size_t const beam_bytes = ((nbeams * sizeof(srs::DevBeam) + 255) / 256) * 256;
  size_t const w_bytes = ((nbeams * sizeof(double) + 255) / 256) * 256;
  HIPCHK(c.out.ensure(beam_bytes + w_bytes + (size_t)np * sizeof(srs::SampledParticle)));
  char* const base = c.out.as<char>();
  auto* const dbeams = reinterpret_cast<srs::DevBeam*>(base);
  auto* const dw = reinterpret_cast<double*>(base + beam_bytes);
  auto* const drows = reinterpret_cast<srs::SampledParticle*>(base + beam_bytes + w_bytes);
  HIPCHK(hipMemcpyAsync(dbeams, db.data(), nbeams * sizeof(srs::DevBeam), hipMemcpyHostToDevice, c.stream));
  HIPCHK(hipMemcpyAsync(dw, wsum.data(), nbeams * sizeof(double), hipMemcpyHostToDevice, c.stream));
  HIPCHK(srs::launch_sample_beam(dbeams, dw, nbeams, seed, first, np, drows, c.stream));
  HIPCHK(hipMemcpyAsync(out, drows, (size_t)np * sizeof(srs_particle), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return SRS_OK;
}

int srs_hip_sample_beam(const srs_beam* beam, int64_t np, uint64_t seed, uint64_t first, srs_particle* out) {
  return srs_hip_sample_beams(beam, nullptr, 1, np, seed, first, out);
}

int srs_hip_total_power(const srs_traj* traj, double charge, double current, double gamma, double* power_out) {
  g_err.clear();
  TRY(check_traj(traj, true));
  if (!power_out) return fail(SRS_ERR_INVALID, "NULL argument");
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  TRY(init(c));
  TrajView t;
  TRY(stage_traj(c, traj, true, &t));
  HIPCHK(c.out.ensure(sizeof(double)));
  // OSCARSSR.cc:1975: |q I| gamma^6 / (6 pi eps0 c)
  double const scale = std::fabs(charge * current) * std::pow(gamma, 6) / (6 * kPi * kEps0 * kC);
  {
    KernelTimer timer(c, c.stream);
    HIPCHK(srs::launch_total_power(t, scale, c.out.as<double>(), c.stream));
  }
  HIPCHK(hipMemcpyAsync(power_out, c.out.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return SRS_OK;
}

int srs_hip_flux_device(const srs_traj* traj, const double* obs_soa, int64_t ns, const double* energies_ev,
                        int64_t ne, double charge, double current, int precision, double* flux_out, void* stream) {
  g_err.clear();
  TRY(check_traj(traj, false));
  TRY(check_common(ns, precision));
  if (ne < 1) return fail(SRS_ERR_INVALID, "no photon energies");
  if (!obs_soa || !energies_ev || !flux_out) return fail(SRS_ERR_INVALID, "NULL argument");
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  TRY(init(c));
  hipStream_t s = stream ? (hipStream_t)stream : c.stream;
  return run_flux(c, view_of(traj), obs_soa, ns, energies_ev, ne, charge, current, precision, flux_out, s);
}

int srs_hip_flux_device_noalloc(const srs_traj* traj, const double* obs_soa, int64_t ns, const double* energies_ev,
                                int64_t ne, double charge, double current, int precision, double* flux_out,
                                void* stream) {
  g_err.clear();
  TRY(check_traj(traj, false));
  TRY(check_common(ns, precision));
  if (ne < 1) return fail(SRS_ERR_INVALID, "no photon energies");
  if (!obs_soa || !energies_ev || !flux_out) return fail(SRS_ERR_INVALID, "NULL argument");
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  TRY(init(c));
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  void* sp = nullptr;
  TRY(scratch_address(dev, &sp));
  TrajView const t = view_of(traj);
  FluxPlan pl;
  TRY(plan_flux(c, t, ns, ne, precision, &pl));
  if (pl.bytes > (size_t)SRS_NOALLOC_SCRATCH_BYTES) {
    char msg[256];
    snprintf(msg, sizeof msg,
             "this flux needs %zu bytes of scratch, more than the %zu of srs_hip_flux_device_noalloc "
             "(use srs_hip_flux_device)",
             pl.bytes, (size_t)SRS_NOALLOC_SCRATCH_BYTES);
    return fail(SRS_ERR_INVALID, msg);
  }
  hipStream_t s = stream ? (hipStream_t)stream : c.stream;
  bool const tsave = c.timing;
  c.timing = false;  // the timer creates its events on demand: not here
  int const rc = run_flux(c, t, obs_soa, ns, energies_ev, ne, charge, current, precision, flux_out, s, pl, sp);
  c.timing = tsave;
  return rc;
}

int srs_hip_spectrum_device(const srs_traj* traj, const double obs[3], const double* energies_ev, int64_t ne,
                            double charge, double current, int precision, double* flux_out, void* stream) {
  g_err.clear();
  TRY(check_traj(traj, false));
  TRY(check_common(ne, precision));
  if (!obs || !energies_ev || !flux_out) return fail(SRS_ERR_INVALID, "NULL argument");
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  TRY(init(c));
  hipStream_t s = stream ? (hipStream_t)stream : c.stream;
  return run_spectrum(c, view_of(traj), obs, energies_ev, ne, charge, current, precision, flux_out, s);
}

int srs_hip_power_density_device(const srs_traj* traj, const double* obs_soa, const double* normals_soa, int64_t ns,
                                 int directional, double charge, double current, int precision, double* pd_out,
                                 void* stream) {
  g_err.clear();
  TRY(check_traj(traj, true));
  TRY(check_common(ns, precision));
  if (!obs_soa || !normals_soa || !pd_out) return fail(SRS_ERR_INVALID, "NULL argument");
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  TRY(init(c));
  hipStream_t s = stream ? (hipStream_t)stream : c.stream;
  return run_pd(c, view_of(traj), obs_soa, normals_soa, ns, directional ? 1 : 0, charge, current, precision, pd_out, s);
}

int srs_hip_timing_enable(int on) {
  g_err.clear();
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  TRY(init(c));
  c.timing = on != 0;
  c.ev_used = 0;
  return SRS_OK;
}

int srs_hip_timing_read(double* total_ms, int64_t* launches) {
  g_err.clear();
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  double sum = 0;
  for (size_t i = 0; i < c.ev_used; ++i) {
    float ms = 0;
    HIPCHK(hipEventSynchronize(c.ev_pool[i].second));
    HIPCHK(hipEventElapsedTime(&ms, c.ev_pool[i].first, c.ev_pool[i].second));
    sum += ms;
  }
  if (total_ms) *total_ms = sum;
  if (launches) *launches = (int64_t)c.ev_used;
  c.ev_used = 0;
  return SRS_OK;
}

int srs_hip_trajectory_points(const srs_particle* p, int64_t np, int64_t* total) {
  g_err.clear();
  if (!p || np < 1 || !total) return fail(SRS_ERR_INVALID, "no particles");
  std::vector<srs::BatchParticle> P;
  return build_particles(p, np, 0.0, 0.0, P, total);
}

int srs_hip_trajectories(const srs_field* und, int nund, const srs_particle* p, int64_t np, double* traj_out,
                         int64_t total, int64_t* offsets_out, double* dt_out) {
  g_err.clear();
  if (!p || np < 1 || !traj_out || !offsets_out || !dt_out) return fail(SRS_ERR_INVALID, "NULL argument");
  std::vector<srs::BatchParticle> P;
  int64_t tot = 0;
  TRY(build_particles(p, np, 0.0, 0.0, P, &tot));
  if (tot != total) return fail(SRS_ERR_INVALID, "traj_out size does not match the trajectory points");
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  c.tape.valid = false;  // shares the batch buffers with the gradient tape
  TRY(init(c));
  hipStream_t s = c.stream;
  TRY(stage_undulators(c, und, nund, s));
  HIPCHK(c.b_part.ensure(np * sizeof(srs::BatchParticle)));
  HIPCHK(hipMemcpyAsync(c.b_part.p, P.data(), np * sizeof(srs::BatchParticle), hipMemcpyHostToDevice, s));
  srs::BatchTraj bt;
  TRY(stage_batch_traj(c, tot, np, true, &bt));
  HIPCHK(srs::launch_rk4_batch(c.b_und.as<srs::DevUndulator>(), nund, c.und_local, c.und_force, c.b_poles.as<double>(),
                               c.pole_count, c.b_part.as<srs::BatchParticle>(), (int)np, bt, srs::Checkpoints{nullptr, 1, 0},
                               srs::RK4Seg{0, INT64_MAX, nullptr}, s));
  double* src[9] = {bt.x, bt.y, bt.z, bt.bx, bt.by, bt.bz, bt.ax, bt.ay, bt.az};
  for (int k = 0; k < 9; ++k)
    HIPCHK(hipMemcpyAsync(traj_out + (size_t)k * tot, src[k], tot * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int64_t i = 0; i < np; ++i) {
    offsets_out[i] = P[i].off;
    dt_out[i] = P[i].dt;
  }
  offsets_out[np] = tot;
  return SRS_OK;
}

int srs_hip_batch_flux(const srs_field* und, int nund, const srs_particle* p, int64_t np, double charge,
                       double current, const double* obs_xyz, int64_t ns, const double* energies_ev, int64_t ne,
                       double weight, int precision, int flags, double* flux_out) {
  g_err.clear();
  TRY(check_common(ns, precision));
  if (ne < 1) return fail(SRS_ERR_INVALID, "no photon energies");
  if (!p || np < 1) return fail(SRS_ERR_INVALID, "no particles");
  if (np > (1 << 24)) return fail(SRS_ERR_INVALID, "too many particles");
  if (!obs_xyz || !energies_ev || !flux_out) return fail(SRS_ERR_INVALID, "NULL argument");
  std::vector<srs::BatchParticle> P;
  int64_t total = 0;
  TRY(build_particles(p, np, charge, current, P, &total));
  int64_t const ntiles = (ns + srs::kBlock - 1) / srs::kBlock;
  if ((double)np * ntiles * ne > 4.0e9) return fail(SRS_ERR_INVALID, "too many work items");
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  TRY(init(c));
  c.tape.valid = false;  // the batch buffers below are shared with the gradient tape
  hipStream_t s = c.stream;
  if (!c.cus) HIPCHK(hipDeviceGetAttribute(&c.cus, hipDeviceAttributeMultiprocessorCount, 0));
  TRY(stage_undulators(c, und, nund, s));
  HIPCHK(c.b_part.ensure(np * sizeof(srs::BatchParticle)));
  HIPCHK(hipMemcpyAsync(c.b_part.p, P.data(), np * sizeof(srs::BatchParticle), hipMemcpyHostToDevice, s));
  int64_t max_leg = 0;
  for (auto const& q : P) max_leg = std::max(max_leg, std::max(q.nf, q.nb));
  return batch_flux_staged(c, s, nund, np, total, max_leg, charge, current, obs_xyz, ns, energies_ev, ne, weight,
                           precision, flags, flux_out);
}

int srs_hip_batch_flux_beams(const srs_field* und, int nund, const srs_beam* beams, const double* weights,
                             int nbeams, int64_t np, uint64_t seed, uint64_t first, const double* obs_xyz, int64_t ns,
                             const double* energies_ev, int64_t ne, double weight, int precision, int flags,
                             double* flux_out) {
  g_err.clear();
  TRY(check_common(ns, precision));
  if (ne < 1) return fail(SRS_ERR_INVALID, "no photon energies");
  if (np < 1) return fail(SRS_ERR_INVALID, "no particles");
  if (np > (1 << 24)) return fail(SRS_ERR_INVALID, "too many particles");
  if (!obs_xyz || !energies_ev || !flux_out) return fail(SRS_ERR_INVALID, "NULL argument");
  std::vector<srs::DevBeam> db;
  std::vector<double> wsum;
  TRY(beam_records(beams, weights, nbeams, true, db, wsum));
  // every particle gets its beam's time grid; trajectories are laid out at a fixed
  // stride (the longest beam's point count)
  int64_t stride = 0, max_leg = 0;
  for (auto const& b : db) {
    stride = std::max(stride, b.gnb + b.gnf);
    max_leg = std::max(max_leg, std::max(b.gnb, b.gnf));
  }
  if ((double)stride * np > (double)((int64_t)1 << 40)) return fail(SRS_ERR_INVALID, "too many trajectory points");
  int64_t const ntiles = (ns + srs::kBlock - 1) / srs::kBlock;
  if ((double)np * ntiles * ne > 4.0e9) return fail(SRS_ERR_INVALID, "too many work items");
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  TRY(init(c));
  c.tape.valid = false;  // the batch buffers below are shared with the gradient tape
  hipStream_t s = c.stream;
  if (!c.cus) HIPCHK(hipDeviceGetAttribute(&c.cus, hipDeviceAttributeMultiprocessorCount, 0));
  TRY(stage_undulators(c, und, nund, s));
  // beam records | cumulative weights, 256-byte aligned; the records of the particles
  // are drawn into b_part on the device
  size_t const beam_bytes = ((nbeams * sizeof(srs::DevBeam) + 255) / 256) * 256;
  HIPCHK(c.b_beams.ensure(beam_bytes + nbeams * sizeof(double)));
  auto* const dbeams = c.b_beams.as<srs::DevBeam>();
  auto* const dw = reinterpret_cast<double*>(c.b_beams.as<char>() + beam_bytes);
  HIPCHK(hipMemcpyAsync(dbeams, db.data(), nbeams * sizeof(srs::DevBeam), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dw, wsum.data(), nbeams * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(c.b_part.ensure(np * sizeof(srs::BatchParticle)));
  HIPCHK(srs::launch_sample_batch(dbeams, dw, nbeams, seed, first, np, stride, c.b_part.as<srs::BatchParticle>(), s));
  return batch_flux_staged(c, s, nund, np, np * stride, max_leg, db[0].q, db[0].cur, obs_xyz, ns, energies_ev, ne,
                           weight, precision, flags, flux_out);
}

int srs_hip_batch_power_density(const srs_field* und, int nund, const srs_particle* p, int64_t np, double charge,
                                double current, const double* obs_xyz, const double* normals_xyz, int64_t ns,
                                int directional, double weight, int precision, double* pd_out) {
  g_err.clear();
  TRY(check_common(ns, precision));
  if (!p || np < 1) return fail(SRS_ERR_INVALID, "no particles");
  if (np > (1 << 24)) return fail(SRS_ERR_INVALID, "too many particles");
  if (!obs_xyz || !normals_xyz || !pd_out) return fail(SRS_ERR_INVALID, "NULL argument");
  std::vector<srs::BatchParticle> P;
  int64_t total = 0;
  TRY(build_particles(p, np, charge, current, P, &total));
  int64_t max_leg = 0, max_pts = 0;
  for (auto const& q : P) {
    max_leg = std::max(max_leg, std::max(q.nf, q.nb));
    max_pts = std::max(max_pts, q.nf + q.nb);
  }
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  TRY(init(c));
  c.tape.valid = false;  // the batch buffers below are shared with the gradient tape
  hipStream_t s = c.stream;
  TRY(stage_undulators(c, und, nund, s));
  HIPCHK(c.b_part.ensure(np * sizeof(srs::BatchParticle)));
  HIPCHK(hipMemcpyAsync(c.b_part.p, P.data(), np * sizeof(srs::BatchParticle), hipMemcpyHostToDevice, s));
  TRY(stage_points(c, c.b_obs, obs_xyz, ns));
  TRY(stage_points(c, c.g_nrm, normals_xyz, ns));
  srs::BatchTraj bt;
  TRY(stage_batch_traj(c, total, np, true, &bt));
  HIPCHK(c.b_flux_pp.ensure((size_t)np * ns * sizeof(double)));
  HIPCHK(c.g_field.ensure((size_t)np * ns * sizeof(double)));
  HIPCHK(c.b_out.ensure((size_t)ns * sizeof(double)));
  {
    KernelTimer timer(c, s);
    TRY(enqueue_pd_batch(c, bt, (int)np, nund, max_leg, max_pts, srs::Checkpoints{nullptr, 1, 0}, (int)ns,
                         directional ? 1 : 0, c.g_field.as<double>(), s,
                         precision == SRS_PRECISION_MIXED));
    HIPCHK(srs::launch_particle_reduce(c.b_flux_pp.as<double>(), (int)np, (int)ns, weight, c.b_out.as<double>(), s));
  }
  HIPCHK(hipMemcpyAsync(pd_out, c.b_out.p, (size_t)ns * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return SRS_OK;
}

int srs_hip_batch_spectrum(const srs_field* und, int nund, const srs_particle* p, int64_t np, double charge,
                           double current, const double obs[3], const double* energies_ev, int64_t ne, double weight,
                           int precision, double* flux_out) {
  g_err.clear();
  TRY(check_common(ne, precision));
  if (!p || np < 1) return fail(SRS_ERR_INVALID, "no particles");
  if (np > (1 << 24)) return fail(SRS_ERR_INVALID, "too many particles");
  bool const mixed = precision == SRS_PRECISION_MIXED;
  if (!obs || !energies_ev || !flux_out) return fail(SRS_ERR_INVALID, "NULL argument");
  std::vector<srs::BatchParticle> P;
  int64_t total = 0;
  TRY(build_particles(p, np, charge, current, P, &total));
  int64_t max_leg = 0, max_pts = 0;
  for (auto const& q : P) {
    max_leg = std::max(max_leg, std::max(q.nf, q.nb));
    max_pts = std::max(max_pts, q.nf + q.nb);
  }
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  TRY(init(c));
  c.tape.valid = false;  // the batch buffers below are shared with the gradient tape
  hipStream_t s = c.stream;
  TRY(stage_undulators(c, und, nund, s));
  HIPCHK(c.b_part.ensure(np * sizeof(srs::BatchParticle)));
  HIPCHK(hipMemcpyAsync(c.b_part.p, P.data(), np * sizeof(srs::BatchParticle), hipMemcpyHostToDevice, s));
  // per energy omega and c/omega; per run the step and |tau| bound of the phase
  // correction (srs_spectrum.h)
  int64_t const nr = (ne + srs::kSpecRun - 1) / srs::kSpecRun;
  std::vector<double> eh(2 * ne + 2 * nr);
  for (int64_t e = 0; e < ne; ++e) {
    double const w = omega_of(energies_ev[e]);
    eh[e] = w;
    eh[ne + e] = kC / w;
  }
  bool const uniform = srs::spectrum_runs_uniform(eh.data(), ne);
  if (uniform) srs::spectrum_run_tables(eh.data(), ne, mixed ? 1 : 0, eh.data() + 2 * ne, eh.data() + 2 * ne + nr);
  HIPCHK(c.b_energies.ensure(eh.size() * sizeof(double)));
  HIPCHK(hipMemcpyAsync(c.b_energies.p, eh.data(), eh.size() * sizeof(double), hipMemcpyHostToDevice, s));
  double* const ed = c.b_energies.as<double>();
  srs::SpecTables const tab{ed, ed + ne, ed + 2 * ne, ed + 2 * ne + nr};
  int const run = uniform ? srs::kSpecRun : 1;
  srs::BatchTraj bt;
  TRY(stage_batch_traj(c, total, np, false, &bt));
  Segments const g = segments_of(max_leg);
  int const nseg = g.n;
  // chunks per segment: at least ~4 blocks (energy group x particle x chunk) per CU
  if (!c.cus) HIPCHK(hipDeviceGetAttribute(&c.cus, hipDeviceAttributeMultiprocessorCount, 0));
  int64_t const tiles = (int64_t)srs::spectrum_batch_groups((int)ne, run) * np;
  int64_t const pts = nseg > 1 ? 2 * g.longest() : max_pts;
  int const nch = (int)std::max<int64_t>(1, std::min<int64_t>((4 * c.cus + tiles - 1) / tiles, (pts + 255) / 256));
  HIPCHK(c.g_part.ensure((size_t)np * nch * 6 * ne * sizeof(double)));
  HIPCHK(c.b_flux_pp.ensure((size_t)np * ne * sizeof(double)));
  HIPCHK(c.b_out.ensure((size_t)ne * sizeof(double)));
  if (nseg > 1) HIPCHK(c.b_state.ensure((size_t)np * 24 * sizeof(double)));
  HIPCHK(hipStreamSynchronize(s));  // host staging vectors
  auto* Pd = c.b_part.as<srs::BatchParticle>();
  auto* U = c.b_und.as<srs::DevUndulator>();
  srs::Checkpoints const ck{nullptr, 1, 0};
  *c.spec_range = 0;
  {
    KernelTimer timer(c, s);
    if (nseg == 1) {
      HIPCHK(srs::launch_rk4_batch(U, nund, c.und_local, c.und_force, c.b_poles.as<double>(), c.pole_count, Pd, (int)np, bt, ck,
                                   srs::RK4Seg{0, INT64_MAX, nullptr}, s));
      HIPCHK(srs::launch_spectrum_batch(bt, Pd, (int)np, obs, tab, (int)ne, run, 0, INT64_MAX, nch, 0,
                                        c.g_part.as<double>(), c.spec_range, mixed, s));
    } else {
      HIPCHK(hipEventRecord(c.seg_ev[Context::kMaxSegments], s));  // fork: the RK4 follows the staging on s
      HIPCHK(hipStreamWaitEvent(c.stream2, c.seg_ev[Context::kMaxSegments], 0));
      for (int k = 0; k < nseg; ++k) {
        HIPCHK(srs::launch_rk4_batch(U, nund, c.und_local, c.und_force, c.b_poles.as<double>(), c.pole_count, Pd, (int)np, bt, ck,
                                     srs::RK4Seg{g.lo(k), g.hi(k), c.b_state.as<double>()}, c.stream2));
        HIPCHK(hipEventRecord(c.seg_ev[k], c.stream2));
      }
      for (int k = 0; k < nseg; ++k) {
        HIPCHK(hipStreamWaitEvent(s, c.seg_ev[k], 0));  // the last wait also joins stream2
        HIPCHK(srs::launch_spectrum_batch(bt, Pd, (int)np, obs, tab, (int)ne, run, g.lo(k), g.hi(k), nch, k > 0,
                                          c.g_part.as<double>(), c.spec_range, mixed, s));
      }
    }
    HIPCHK(srs::launch_spectrum_batch_reduce(c.g_part.as<double>(), nch, (int)ne, (int)np, Pd, ed,
                                             c.b_flux_pp.as<double>(), s));
    HIPCHK(srs::launch_particle_reduce(c.b_flux_pp.as<double>(), (int)np, (int)ne, weight, c.b_out.as<double>(), s));
  }
  HIPCHK(hipStreamSynchronize(s));
  if (run > 1 && *(volatile int*)c.spec_range != 0) {
    // a trajectory left the range of the runs' phase correction: the whole spectrum
    // again, one energy per lane, over the trajectories the RK4 stored
    HIPCHK(srs::launch_spectrum_batch(bt, Pd, (int)np, obs, tab, (int)ne, 1, 0, INT64_MAX, nch, 0,
                                      c.g_part.as<double>(), c.spec_range, mixed, s));
    HIPCHK(srs::launch_spectrum_batch_reduce(c.g_part.as<double>(), nch, (int)ne, (int)np, Pd, ed,
                                             c.b_flux_pp.as<double>(), s));
    HIPCHK(srs::launch_particle_reduce(c.b_flux_pp.as<double>(), (int)np, (int)ne, weight, c.b_out.as<double>(), s));
  }
  HIPCHK(hipMemcpyAsync(flux_out, c.b_out.p, (size_t)ne * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return SRS_OK;
}

int srs_hip_batch_config(int* ne_per_item, int* tile) {
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  if (ne_per_item) *ne_per_item = c.b_last.ne_per_item;
  if (tile) *tile = c.b_last.tile;
  return SRS_OK;
}

int srs_hip_batch_points_per_lane(int* points_per_lane) {
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  if (points_per_lane) *points_per_lane = c.b_last.ppl;
  return SRS_OK;
}

int srs_hip_batch_set_config(int ne_per_item, int tile, int points_per_lane) {
  g_err.clear();
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  if (ne_per_item == 0) {
    c.b_fixed = srs::BatchConfig{0, 0};
    return SRS_OK;
  }
  srs::BatchConfig const k{ne_per_item, tile, points_per_lane};
  if (!srs::batch_config_valid(k)) return fail(SRS_ERR_INVALID, "no batch kernel for this tile shape");
  c.b_fixed = k;
  return SRS_OK;
}

int srs_hip_synchronize(void* stream) {
  g_err.clear();
  Context& c = ctx();
  if (!stream && !c.ready) return SRS_OK;
  hipStream_t s = stream ? (hipStream_t)stream : c.stream;
  HIPCHK(hipStreamSynchronize(s));
  return SRS_OK;
}


// ---- gradient of the flux loss with respect to per-pole field errors ---------------

int srs_hip_pole_forward(const srs_field* und, int nund, int poled, const srs_particle* particle, int64_t nreal,
                         const double* deltas, double charge, double current, const double* obs_xyz, int64_t ns,
                         double energy_ev, const double* target, int checkpoint_every, double* flux_out,
                         double* loss_out, int64_t* tape_out) {
  return srs_hip_pole_forward_ex(und, nund, poled, particle, nreal, deltas, charge, current, obs_xyz, ns, energy_ev,
                                 target, checkpoint_every, SRS_PRECISION_FP64, flux_out, loss_out, tape_out);
}

int srs_hip_pole_forward_ex(const srs_field* und, int nund, int poled, const srs_particle* particle, int64_t nreal,
                            const double* deltas, double charge, double current, const double* obs_xyz, int64_t ns,
                            double energy_ev, const double* target, int checkpoint_every, int precision,
                            double* flux_out, double* loss_out, int64_t* tape_out) {
  g_err.clear();
  TRY(pole_validate(und, nund, poled, particle, nreal, deltas, obs_xyz, ns, target, checkpoint_every, loss_out,
                    tape_out));
  if (!(energy_ev > 0)) return fail(SRS_ERR_INVALID, "photon energy must be > 0");
  if (precision != SRS_PRECISION_FP64 && precision != SRS_PRECISION_MIXED)
    return fail(SRS_ERR_INVALID, "precision must be SRS_PRECISION_FP64 or SRS_PRECISION_MIXED");
  particle_charge(*particle, charge, current, &charge, &current);
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  TRY(init(c));
  hipStream_t s = c.stream;
  PoleRun pr;
  TRY(pole_stage(c, und, nund, poled, particle, nreal, deltas, charge, current, obs_xyz, ns, target, checkpoint_every,
                 false, pr, s));
  double const w = omega_of(energy_ev), c0 = c0_of(charge);
  double const eh[5] = {w, kC / w, c0 * w, w / kTwoPi, (w / kTwoPi) / kC};
  HIPCHK(c.b_energies.ensure(5 * sizeof(double)));
  HIPCHK(hipMemcpyAsync(c.b_energies.p, eh, 5 * sizeof(double), hipMemcpyHostToDevice, s));
  double* ed = c.b_energies.as<double>();
  srs::BatchEnergies const be{ed, ed + 1, ed + 2, ed + 3, ed + 4, 1, c2_of(charge, current)};
  HIPCHK(c.g_field.ensure((size_t)nreal * 6 * ns * sizeof(double)));
  HIPCHK(c.b_counter.ensure(sizeof(unsigned)));
  {
    BatchRun r;
    r.nund = nund;
    r.bt = pr.bt;
    r.be = be;
    r.np = (int)nreal;
    r.ns = (int)ns;
    r.ne = 1;
    r.mixed = precision == SRS_PRECISION_MIXED;
    r.weight = 1.0;
    r.max_leg = std::max(pr.P[0].nf, pr.P[0].nb);
    r.ck = pr.ck;
    r.field_out = c.g_field.as<double>();
    KernelTimer timer(c, s);
    TRY(enqueue_batch(c, r, srs::BatchConfig{1, 256}, (int)nreal, s, false));
  }
  TRY(pole_finish(c, pr, (int)nreal, (int)ns, flux_out, loss_out, s));
  Context::Tape& t = pole_tape(c, pr, (int)nreal, (int)ns, nund, poled);
  t.quantity = 0;
  t.mixed = precision == SRS_PRECISION_MIXED;
  {
    srs::AdjScreen sc{1, obs_xyz[2], obs_xyz[0], obs_xyz[0], obs_xyz[1], obs_xyz[1]};
    for (int64_t j = 0; j < ns; ++j) {
      double const* o = obs_xyz + 3 * j;
      sc.flat = sc.flat && o[2] == sc.z;
      sc.xlo = std::min(sc.xlo, o[0]);
      sc.xhi = std::max(sc.xhi, o[0]);
      sc.ylo = std::min(sc.ylo, o[1]);
      sc.yhi = std::max(sc.yhi, o[1]);
    }
    t.screen = sc;
  }
  t.omega = w;
  t.cw = kC / w;
  t.c0w = c0 * w;
  t.c2 = be.c2;
  *tape_out = t.id;
  return SRS_OK;
}

int srs_hip_pole_pd_forward(const srs_field* und, int nund, int poled, const srs_particle* particle,
                            int64_t nreal, const double* deltas, double charge, double current,
                            const double* obs_xyz, const double* normals_xyz, int64_t ns, int directional,
                            const double* target, int checkpoint_every, double* pd_out, double* loss_out,
                            int64_t* tape_out) {
  g_err.clear();
  TRY(pole_validate(und, nund, poled, particle, nreal, deltas, obs_xyz, ns, target, checkpoint_every, loss_out,
                    tape_out));
  if (!normals_xyz) return fail(SRS_ERR_INVALID, "NULL argument");
  particle_charge(*particle, charge, current, &charge, &current);
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  TRY(init(c));
  hipStream_t s = c.stream;
  PoleRun pr;
  TRY(pole_stage(c, und, nund, poled, particle, nreal, deltas, charge, current, obs_xyz, ns, target, checkpoint_every,
                 true, pr, s));
  TRY(stage_points(c, c.g_nrm, normals_xyz, ns));
  HIPCHK(c.g_field.ensure((size_t)nreal * ns * sizeof(double)));
  double const scale = pd_scale_of(charge, current, pr.P[0].dt);
  {
    KernelTimer timer(c, s);
    TRY(enqueue_pd_batch(c, pr.bt, (int)nreal, nund, std::max(pr.P[0].nf, pr.P[0].nb), pr.P[0].nb + pr.P[0].nf, pr.ck,
                         (int)ns, directional ? 1 : 0, c.g_field.as<double>(), s));
  }
  TRY(pole_finish(c, pr, (int)nreal, (int)ns, pd_out, loss_out, s));
  Context::Tape& t = pole_tape(c, pr, (int)nreal, (int)ns, nund, poled);
  t.quantity = 1;
  t.mixed = false;
  t.directional = directional ? 1 : 0;
  t.pd_scale = scale;
  *tape_out = t.id;
  return SRS_OK;
}

int srs_hip_pole_backward(int64_t tape, const double* dloss, double* grad_out) {
  g_err.clear();
  if (!grad_out) return fail(SRS_ERR_INVALID, "NULL argument");
  Context& c = ctx();
  std::lock_guard<std::mutex> lock(c.mu);
  TRY(init(c));
  Context::Tape const& t = c.tape;
  if (!t.valid || t.id != tape)
    return fail(SRS_ERR_INVALID, "no forward tape for this call (another forward or batch call replaced it)");
  hipStream_t s = c.stream;
  HIPCHK(c.g_dloss.ensure(t.np * sizeof(double)));
  {
    std::vector<double> dl((size_t)t.np, 1.0);
    if (dloss) std::copy(dloss, dloss + t.np, dl.begin());
    HIPCHK(hipMemcpyAsync(c.g_dloss.p, dl.data(), t.np * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  bool const pd = t.quantity == 1;
  HIPCHK(c.g_gy.ensure((size_t)(pd ? 9 : 6) * t.total * sizeof(double)));
  double* const gy = c.g_gy.as<double>();
  double* const ga = pd ? gy + 6 * t.total : nullptr;
  int64_t const rec_steps = std::min<int64_t>(t.seg_per_pass * t.every, std::max(t.nf, t.nb));
  HIPCHK(c.g_rec.ensure(srs::adjoint_record_doubles(rec_steps, t.np) * sizeof(double)));
  HIPCHK(c.g_lam.ensure((size_t)t.np * 6 * sizeof(double)));
  HIPCHK(c.g_sweep.ensure(srs::adjoint_sweep_doubles(t.seg_per_pass, t.np, t.npoles) * sizeof(double)));
  HIPCHK(c.g_grad.ensure((size_t)t.np * t.npoles * sizeof(double)));
  auto* Pd = c.b_part.as<srs::BatchParticle>();
  {
    KernelTimer timer(c, s);
    if (pd)
      HIPCHK(srs::launch_pd_adjoint(t.bt, Pd, t.np, t.nmax, c.b_obs.as<double>(), c.g_nrm.as<double>(), t.ns,
                                    t.directional, t.pd_scale, c.g_field.as<double>(), c.b_flux_pp.as<double>(),
                                    c.g_target.as<double>(), c.g_dloss.as<double>(), gy, ga, t.total, s));
    else
      HIPCHK(srs::launch_flux_adjoint(t.bt, Pd, t.np, t.nmax, c.b_obs.as<double>(), t.ns, t.omega, t.cw, t.c0w, t.c2,
                                      c.g_field.as<double>(), c.b_flux_pp.as<double>(), c.g_target.as<double>(),
                                      c.g_dloss.as<double>(), gy, t.total, t.mixed, t.screen, s));
  }
  HIPCHK(srs::launch_zero(c.g_grad.as<double>(), (int64_t)t.np * t.npoles, s));
  {
    KernelTimer timer(c, s);
    HIPCHK(srs::launch_rk4_adjoint(c.b_und.as<srs::DevUndulator>(), t.nund, t.poled, c.b_poles.as<double>(), Pd, t.np,
                                   t.nf, t.nb, t.ck, t.seg_per_pass, gy, ga, t.total, c.g_rec.as<double>(),
                                   c.g_sweep.as<double>(), c.g_lam.as<double>(), c.g_grad.as<double>(), t.npoles, s));
  }
  HIPCHK(hipMemcpyAsync(grad_out, c.g_grad.p, (size_t)t.np * t.npoles * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return SRS_OK;
}

}  // extern "C"
