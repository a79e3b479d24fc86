// C ABI of libsrs_hip.so (declared in include/srs_hip.h).
//
// Host side of the boundary: validates arguments, derives the per-energy
// constants in the reference's operation order, stages host arrays into
// grow-only device buffers (no allocation per call once warm), picks the
// trajectory chunking and launches the kernels on one stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/srs_hip.h"
#include "srs_kernels.h"

namespace {

using srs::EnergySet;
using srs::TrajView;

// Reference constants, TOSCARSSR.h:20-69.
constexpr double kPi = 3.14159265358979323846;
constexpr double kFourPi = 4.0 * kPi;
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
  Buf traj, trajf, obs, nrm, energies, part, out;
  // optional HIP-event timing of the integral kernels (srs_hip_timing_*)
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
  size_t ev_used = 0;
};

Context& ctx() {
  static Context c;
  return c;
}

#define HIPCHK(expr)                                                                 \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess) return fail(SRS_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(_e)); \
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
    hipEventRecord(p.first, s);
    stop = p.second;
  }
  ~KernelTimer() {
    if (stop) hipEventRecord(stop, s);
  }
};

double omega_of(double e_ev) { return e_ev * kQe / kHbar; }                       // TOSCARSSR.h:66
double c0_of(double q) { return q / (kFourPi * kC * kEps0 * kSqrt2Pi); }          // OSCARSSR.cc:1191
double c2_of(double q, double cur) {                                               // OSCARSSR.cc:1194
  return kFourPi * cur / (kH * std::fabs(q) * kMu0 * kC) * 1e-6 * 0.001;
}
double pd_scale_of(double q, double cur, double dt) {                              // OSCARSSR.cc:1563
  return std::fabs(q * cur) / (16 * kPi2 * kEps0 * kC) * dt;
}

int check_traj(const srs_traj* t, bool need_a) {
  if (!t) return fail(SRS_ERR_INVALID, "trajectory is NULL");
  if (t->n < 1) return fail(SRS_ERR_INVALID, "no points in trajectory.  Is particle or beam defined?");
  if (t->n > (int64_t)1 << 31) return fail(SRS_ERR_INVALID, "trajectory longer than 2^31 points");
  if (!t->x || !t->y || !t->z || !t->bx || !t->by || !t->bz) return fail(SRS_ERR_INVALID, "trajectory array is NULL");
  if (need_a && (!t->ax || !t->ay || !t->az)) return fail(SRS_ERR_INVALID, "trajectory acceleration array is NULL");
  return SRS_OK;
}

// Chunking of the trajectory so that a launch has ~2048 blocks (8 per CU on 256 CUs).
int pick_chunk(int64_t n, int64_t lanes, int align) {
  int64_t const bx = (lanes + srs::kBlock - 1) / srs::kBlock;
  int64_t nch = (2048 + bx - 1) / bx;
  int64_t const maxch = (n + align - 1) / align;
  nch = std::max<int64_t>(1, std::min(nch, maxch));
  int64_t chunk = (n + nch - 1) / nch;
  chunk = ((chunk + align - 1) / align) * align;
  return (int)chunk;
}
int nchunks(int64_t n, int chunk) { return (int)((n + chunk - 1) / chunk); }

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

int run_flux(Context& c, const TrajView& t, const double* obs_soa, int64_t ns, const double* energies_ev, int64_t ne,
             double q, double cur, int precision, double* out, hipStream_t s) {
  double const c0 = c0_of(q), c2 = c2_of(q, cur);
  int const chunk = pick_chunk(t.n, ns, 256);
  int const nch = nchunks(t.n, chunk);
  int const maxe = srs::kMaxFluxEnergies;
  HIPCHK(c.part.ensure((size_t)nch * maxe * 6 * ns * sizeof(double)));
  if (precision == SRS_PRECISION_MIXED) {
    HIPCHK(c.trajf.ensure((size_t)6 * t.n * sizeof(float)));
    HIPCHK(srs::launch_traj_to_f32(t, c.trajf.as<float>(), s));
  }
  for (int64_t e0 = 0; e0 < ne; e0 += maxe) {
    EnergySet es{};
    es.ne = (int)std::min<int64_t>(maxe, ne - e0);
    for (int e = 0; e < es.ne; ++e) {
      double const w = omega_of(energies_ev[e0 + e]);
      es.omega[e] = w;
      es.cw[e] = kC / w;
      es.kw[e] = (c0 * w) * t.dt;
    }
    {
      KernelTimer timer(c, s);
      if (precision == SRS_PRECISION_MIXED) {
        HIPCHK(srs::launch_flux_mixed(t, c.trajf.as<float>(), obs_soa, (int)ns, es, chunk, c.part.as<double>(), s));
      } else {
        HIPCHK(srs::launch_flux_fp64(t, obs_soa, (int)ns, es, chunk, c.part.as<double>(), s));
      }
    }
    HIPCHK(srs::launch_flux_reduce(c.part.as<double>(), nch, (int)ns, es, c2, out + e0 * ns, s));
  }
  return SRS_OK;
}

int run_spectrum(Context& c, const TrajView& t, const double obs[3], const double* energies_ev, int64_t ne, double q,
                 double cur, int precision, double* out, hipStream_t s) {
  double const c0 = c0_of(q), c2 = c2_of(q, cur);
  std::vector<double> h(3 * ne);
  for (int64_t e = 0; e < ne; ++e) {
    double const w = omega_of(energies_ev[e]);
    h[e] = w;
    h[ne + e] = kC / w;
    h[2 * ne + e] = (c0 * w) * t.dt;
  }
  HIPCHK(c.energies.ensure(3 * ne * sizeof(double)));
  double* d = c.energies.as<double>();
  HIPCHK(hipMemcpyAsync(d, h.data(), 3 * ne * sizeof(double), hipMemcpyHostToDevice, s));
  if (precision == SRS_PRECISION_MIXED) {
    int64_t const lanes = (ne + srs::spectrum_mixed_run() - 1) / srs::spectrum_mixed_run();
    int const chunk = pick_chunk(t.n, lanes, 32);
    int const nch = nchunks(t.n, chunk);
    HIPCHK(c.part.ensure((size_t)nch * 6 * ne * sizeof(double)));
    {
      KernelTimer timer(c, s);
      HIPCHK(srs::launch_spectrum_mixed(t, obs, d, d + ne, (int)ne, chunk, 0, c.part.as<double>(), s));
    }
    HIPCHK(srs::launch_field_reduce(c.part.as<double>(), nch, (int)ne, d + 2 * ne, c2, out, s));
  } else {
    int const chunk = pick_chunk(t.n, ne, 256);
    int const nch = nchunks(t.n, chunk);
    HIPCHK(c.part.ensure((size_t)nch * 12 * ne * sizeof(double)));
    {
      KernelTimer timer(c, s);
      HIPCHK(srs::launch_spectrum_fp64(t, obs, d, d + ne, (int)ne, chunk, c.part.as<double>(), s));
    }
    HIPCHK(srs::launch_spectrum_reduce(c.part.as<double>(), nch, (int)ne, d + ne, d + 2 * ne, c2, out, s));
  }
  // the host-side energy table must outlive the async copy
  HIPCHK(hipStreamSynchronize(s));
  return SRS_OK;
}

int run_pd(Context& c, const TrajView& t, const double* obs_soa, const double* nrm_soa, int64_t ns, int directional,
           double q, double cur, int precision, double* out, hipStream_t s) {
  int const chunk = pick_chunk(t.n, ns, 256);
  int const nch = nchunks(t.n, chunk);
  HIPCHK(c.part.ensure((size_t)nch * ns * sizeof(double)));
  if (precision == SRS_PRECISION_MIXED) {
    HIPCHK(c.trajf.ensure((size_t)6 * t.n * sizeof(float)));
    HIPCHK(srs::launch_traj_to_f32(t, c.trajf.as<float>(), s));
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

#define TRY(expr)              \
  do {                         \
    int _r = (expr);           \
    if (_r != SRS_OK) return _r; \
  } while (0)

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

int srs_hip_synchronize(void* stream) {
  g_err.clear();
  Context& c = ctx();
  if (!stream && !c.ready) return SRS_OK;
  hipStream_t s = stream ? (hipStream_t)stream : c.stream;
  HIPCHK(hipStreamSynchronize(s));
  return SRS_OK;
}

}  // extern "C"
