// Spectrum of one trajectory at one observation point (OSCARSSR.cc:1202-1246), FP64
// and mixed forms, on the energy-run recurrence of srs_spectrum.h.
//
// Block: kSpecWaves waves on the same 64 lanes x RUN energies, each wave on its own
// part of the block's trajectory chunk (a fixed split: every sum is formed in the same
// order on every call); each wave stages the geometry of 64 points at a time in LDS (one
// point per lane, FP64, the reference's operation order) and every lane walks them.  At
// the end the waves' FP64 sums are added in LDS in wave order and the block writes one
// partial per energy: part[(c*6 + q)*ne + e], c = chunk.
// spectrum_fold_kernel adds the chunks in a fixed order (deterministic) and applies
// the prefactors.  A run kernel whose trajectory reaches a |tau| beyond a lane's series
// range writes 1 to *range (host memory); the caller then evaluates with runs of one.
#include "srs_kernels.h"
#include "srs_spectrum.h"

namespace srs {
namespace {

constexpr int kSpecWaves = 8;
constexpr int kSpecThreads = 64 * kSpecWaves;
constexpr int kFoldRun = 4;  // energies per lane folded per LDS round (8 x 4 x 6 x 64 doubles)

// Issue balance of the two waves that share a SIMD (waves w and w ^ 4 of a block of 8,
// the CU placing a block's waves on its 4 SIMDs in turn).  Both have the same number of
// points; left alone, the older takes most issue slots and the younger finishes its
// points alone, with nothing to hide its latencies.  After each point a wave publishes
// its count and takes the lower priority while it is ahead of its partner, so the pair
// advances together.  Only the order of issue changes, never a result.
struct IssueBalance {
  int* prog;
  int self, partner;
  __device__ __forceinline__ void step(int done) {
    int const other = __builtin_amdgcn_readfirstlane(
        __hip_atomic_load(prog + partner, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
    __hip_atomic_store(prog + self, done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (done > other)
      __builtin_amdgcn_s_setprio(0);
    else
      __builtin_amdgcn_s_setprio(2);
  }
};

template <bool MIXED, int RUN>
__global__ __launch_bounds__(kSpecThreads) void spectrum_kernel(TrajView t, double ox, double oy, double oz,
                                                                SpecTables tab, int ne, int sub,
                                                                double* __restrict__ part, int* __restrict__ range) {
  using T = typename SpecLane<MIXED, RUN>::T;
  constexpr int FR = RUN < kFoldRun ? RUN : kFoldRun;  // energies of the run folded per round
  struct Tile {
    double tau[64], psi[64];
    T wc[64], ws[64];
    T a[3][64], g[3][64];
  };
  __shared__ union {
    Tile tile[kSpecWaves];
    double fold[kSpecWaves][FR][6][64];
  } sm;
  __shared__ int prog[kSpecWaves];
  int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) prog[wave] = 0;
  SpecLane<MIXED, RUN> L;
  L.init(tab.omega, tab.cw, tab.rdw, tab.rtl, ne, (blockIdx.x * 64 + lane) * RUN);
  int64_t const i0 = ((int64_t)blockIdx.y * kSpecWaves + wave) * sub;
  int64_t const i1 = (i0 + sub < t.n) ? i0 + sub : t.n;
  int const ntiles = (sub + 63) / 64;
  Tile& tl = sm.tile[wave];
  IssueBalance bal{prog, wave, wave ^ 4};
  int done = 0;
  double tmax = 0.0;  // largest |tau| this lane staged
  double const cwmax = wave_max(L.cw_max());
  // mixed: tiles whose |tau| stays within a tenth of every lane's series range take the
  // first-order correction
  double const tlim1 = kFirstOrderFrac * wave_min(L.tau_lim);
  // the step of every run of this block (one value per block, spectrum_run_tables)
  double const dwb = RUN > 1 ? tab.rdw[(blockIdx.x * 64 * RUN < ne ? blockIdx.x * 64 * RUN : 0) / RUN] : 0.0;
  __syncthreads();  // prog is set
  for (int tt = 0; tt < ntiles; ++tt) {
    int64_t const base = i0 + (int64_t)tt * 64;
    int const cnt = base < i1 ? ((i1 - base) < 64 ? (int)(i1 - base) : 64) : 0;
    if (tt > 0) wave_sync();  // the previous tile is consumed
    double dmin = 1e300, tabs = 0.0;
    if (lane < cnt) {
      int64_t const i = base + lane;
      SpecPoint P;
      spec_geometry(ox, oy, oz, t.x[i], t.y[i], t.z[i], t.bx[i], t.by[i], t.bz[i], t.dt * (double)i, P);
      tl.tau[lane] = P.tau;
      double psi = 0.0;
      T wc = (T)1, ws = (T)0;
      if constexpr (RUN > 1) spec_step<MIXED>(dwb, P.tau, psi, wc, ws);
      tl.psi[lane] = psi;
      tl.wc[lane] = wc;
      tl.ws[lane] = ws;
      tabs = __builtin_fabs(P.tau);
      tmax = __builtin_fmax(tmax, tabs);
      dmin = P.d;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        tl.a[k][lane] = (T)P.a[k];
        tl.g[k][lane] = (T)P.g[k];
      }
    }
    wave_sync();
    // mixed: this tile keeps the x, y near field unless (c/omega)/D <= kSpecNear throughout
    bool const xy_near = !MIXED || cwmax > kSpecNear * wave_min(dmin);
    bool const second = !MIXED || wave_max(tabs) > tlim1;
    for (int k0 = 0; k0 < cnt; k0 += kSub) {
      int const k1 = (k0 + kSub < cnt) ? k0 + kSub : cnt;
      if constexpr (MIXED) {
        // the next point's values are read from LDS while the current one is integrated
        auto run = [&](auto near) {
          double tn = tl.tau[k0], pn = tl.psi[k0];
          T wcn = tl.wc[k0], wsn = tl.ws[k0];
          T an0 = tl.a[0][k0], an1 = tl.a[1][k0], an2 = tl.a[2][k0];
          T gn0 = tl.g[0][k0], gn1 = tl.g[1][k0], gn2 = tl.g[2][k0];
          for (int k = k0; k < k1; ++k) {
            double const tc = tn, psc = pn;
            T const wcc = wcn, wsc = wsn;
            T const a0 = an0, a1 = an1, a2 = an2, g0 = gn0, g1 = gn1, g2 = gn2;
            int const kn = k + 1 < k1 ? k + 1 : k;
            tn = tl.tau[kn];
            pn = tl.psi[kn];
            wcn = tl.wc[kn];
            wsn = tl.ws[kn];
            an0 = tl.a[0][kn];
            an1 = tl.a[1][kn];
            an2 = tl.a[2][kn];
            gn0 = tl.g[0][kn];
            gn1 = tl.g[1][kn];
            gn2 = tl.g[2][kn];
            L.template add<decltype(near)::value, decltype(near)::value>(tc, psc, wcc, wsc, a0, a1, a2, g0, g1, g2);
            bal.step(++done);
          }
        };
        // every term of the tile in full, or the fast form: no x, y near field and the
        // first-order correction
        if (xy_near || second)
          run(std::true_type{});
        else
          run(std::false_type{});
      } else {
        for (int k = k0; k < k1; ++k) {
          L.add(tl.tau[k], tl.psi[k], tl.wc[k], tl.ws[k], tl.a[0][k], tl.a[1][k], tl.a[2][k], tl.g[0][k], tl.g[1][k],
                tl.g[2][k]);
          bal.step(++done);
        }
      }
      L.flush();
    }
  }
  // every lane's series must have covered the wave's largest |tau|
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) tmax = __builtin_fmax(tmax, __shfl_xor(tmax, m, 64));
  if (tmax > L.tau_lim) *range = 1;
  __builtin_amdgcn_s_setprio(0);
  // the waves' sums in wave order, FR energies of the run per round
#pragma unroll
  for (int r0 = 0; r0 < RUN; r0 += FR) {
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < FR; ++rr)
#pragma unroll
      for (int q = 0; q < 6; ++q) sm.fold[wave][rr][q][lane] = L.acc[r0 + rr][q];
    __syncthreads();
    for (int idx = threadIdx.x; idx < FR * 6 * 64; idx += kSpecThreads) {
      int const rr = idx / 384, q = (idx / 64) % 6, l = idx % 64;
      double v = sm.fold[0][rr][q][l];
#pragma unroll
      for (int w = 1; w < kSpecWaves; ++w) v += sm.fold[w][rr][q][l];
      int const e = (blockIdx.x * 64 + l) * RUN + r0 + rr;
      if (e < ne) part[((size_t)blockIdx.y * 6 + q) * ne + e] = v;
    }
  }
}

// out[e] = C2 sum_axis |i (C0 omega_e) dt S_axis|^2 with S = the chunks' sums in
// chunk order.  Block: 32 energies x 8 chunk slices (slice s adds chunks s, s+8, ...;
// the slices are then added in slice order), loads coalesced across the energies.
constexpr int kFoldE = 32, kFoldSlices = 8;

__global__ __launch_bounds__(kFoldE* kFoldSlices) void spectrum_fold_kernel(const double* __restrict__ part,
                                                                            int nchunk, int ne,
                                                                            const double* __restrict__ kw, double c2,
                                                                            double* __restrict__ out) {
  __shared__ double v[kFoldSlices][6][kFoldE];
  int const el = threadIdx.x % kFoldE, sl = threadIdx.x / kFoldE;
  int const e = blockIdx.x * kFoldE + el;
  double s[6] = {0, 0, 0, 0, 0, 0};
  if (e < ne) {
    for (int c = sl; c < nchunk; c += kFoldSlices)
#pragma unroll
      for (int q = 0; q < 6; ++q) s[q] += part[((size_t)c * 6 + q) * ne + e];
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) v[sl][q][el] = s[q];
  __syncthreads();
  if (sl != 0 || e >= ne) return;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    double a = v[0][q][el];
#pragma unroll
    for (int k = 1; k < kFoldSlices; ++k) a += v[k][q][el];
    s[q] = a;
  }
  double const k2 = kw[e];
  double p[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    double const re = -(s[2 * q + 1] * k2);
    double const im = s[2 * q] * k2;
    p[q] = re * re + im * im;
  }
  out[e] = c2 * (p[0] + p[1] + p[2]);
}

template <class K>
int resident_blocks(K kernel, int& cache) {
  if (cache == 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(kernel), kSpecThreads, 0) !=
            hipSuccess ||
        nb < 1)
      nb = 1;
    cache = nb;
  }
  return cache;
}

int spectrum_blocks_per_cu(int mixed, int run) {
  static int c[4];
  if (mixed) return run > 1 ? resident_blocks(spectrum_kernel<true, kSpecRun>, c[0]) : resident_blocks(spectrum_kernel<true, 1>, c[1]);
  return run > 1 ? resident_blocks(spectrum_kernel<false, kSpecRun>, c[2]) : resident_blocks(spectrum_kernel<false, 1>, c[3]);
}

}  // namespace

// Energies per block of the run kernels: 64 lanes x kSpecRun, all stepping with one
// spacing (the block's mean step), so that the step's angle and phasor are formed once
// per trajectory point and block.
constexpr int64_t kSpecGroup = 64 * kSpecRun;

// The mean step of the group of energies starting at g0.
static double group_step(const double* omega, int64_t ne, int64_t g0) {
  int64_t const len = (ne - g0) < kSpecGroup ? ne - g0 : kSpecGroup;
  return len > 1 ? (omega[g0 + len - 1] - omega[g0]) / (double)(len - 1) : 0.0;
}

bool spectrum_runs_uniform(const double* omega, int64_t ne) {
  for (int64_t g0 = 0; g0 < ne; g0 += kSpecGroup) {
    double const dw = group_step(omega, ne, g0);
    int64_t const gend = (ne - g0) < kSpecGroup ? ne : g0 + kSpecGroup;
    for (int64_t e0 = g0; e0 < gend; e0 += kSpecRun) {
      int64_t const len = (gend - e0) < kSpecRun ? gend - e0 : kSpecRun;
      double const w0 = omega[e0];
      if (!(w0 > 0.0) || !(w0 < 1e300)) return false;
      for (int64_t r = 1; r < len; ++r) {
        double const wr = omega[e0 + r];
        if (!(wr >= 0.6 * w0 && wr <= 1.6 * w0)) return false;
        // a uniform grid up to rounding: the correction angle stays in its series' range
        if (!(__builtin_fabs(wr - (w0 + (double)r * dw)) <= 64.0 * 0x1p-52 * wr)) return false;
      }
    }
  }
  return true;
}

void spectrum_run_tables(const double* omega, int64_t ne, int mixed, double* rdw, double* rtl) {
  // theta_r = (omega_r - omega_0 - r dw) tau + e_r - e_0 - r e_psi with each rounding
  // error e_x at most 2^-53 |x tau|, so |theta_r| <= |tau| K, K the largest over the run
  // of |omega_r - omega_0 - r dw| (plus one ulp of omega_r for its own rounding)
  // + 2^-53 (omega_r + omega_0 + r |dw|); tau_lim = range / (1.01 K).  dw is the mean
  // step of the run's group of kSpecGroup energies (one per block).
  double const thr = mixed ? (double)kThetaF32 : kThetaF64;
  for (int64_t e0 = 0, g = 0; e0 < ne; e0 += kSpecRun, ++g) {
    int64_t const len = (ne - e0) < kSpecRun ? ne - e0 : kSpecRun;
    double const w0 = omega[e0];
    double const dw = group_step(omega, ne, e0 / kSpecGroup * kSpecGroup);
    double K = 0.0;
    for (int64_t r = 1; r < len; ++r) {
      double const wr = omega[e0 + r];
      double const dev = __builtin_fabs(wr - __builtin_fma((double)r, dw, w0));
      double const k = dev + 0x1p-52 * wr + 0x1p-53 * (wr + w0 + (double)r * __builtin_fabs(dw));
      K = k > K ? k : K;
    }
    rdw[g] = dw;
    rtl[g] = K > 0.0 ? thr / (1.01 * K) : 1e300;
  }
}

SpectrumPlan plan_spectrum(int64_t n, int64_t ne, bool uniform, int mixed, int cus) {
  SpectrumPlan p;
  p.run = uniform ? kSpecRun : 1;
  p.groups = (int)((ne + 64 * p.run - 1) / (64 * p.run));
  int64_t const target = (int64_t)cus * spectrum_blocks_per_cu(mixed, p.run);
  int64_t nchb = (target + p.groups - 1) / p.groups;
  int64_t const most = (n + kSpecWaves * 16 - 1) / (kSpecWaves * 16);  // >= 16 points per wave
  if (nchb > most) nchb = most;
  if (nchb < 1) nchb = 1;
  int64_t const chunk = (n + nchb - 1) / nchb;
  p.sub = (int)((chunk + kSpecWaves - 1) / kSpecWaves);
  p.nchunk = (int)((n + (int64_t)p.sub * kSpecWaves - 1) / ((int64_t)p.sub * kSpecWaves));
  return p;
}

hipError_t launch_spectrum(const TrajView& t, const double obs[3], const SpecTables& tab, int ne, const SpectrumPlan& p,
                           int mixed, double* part, int* range, hipStream_t s) {
  dim3 grid(p.groups, p.nchunk);
  if (mixed) {
    if (p.run > 1)
      spectrum_kernel<true, kSpecRun><<<grid, kSpecThreads, 0, s>>>(t, obs[0], obs[1], obs[2], tab, ne, p.sub, part, range);
    else
      spectrum_kernel<true, 1><<<grid, kSpecThreads, 0, s>>>(t, obs[0], obs[1], obs[2], tab, ne, p.sub, part, range);
  } else {
    if (p.run > 1)
      spectrum_kernel<false, kSpecRun><<<grid, kSpecThreads, 0, s>>>(t, obs[0], obs[1], obs[2], tab, ne, p.sub, part, range);
    else
      spectrum_kernel<false, 1><<<grid, kSpecThreads, 0, s>>>(t, obs[0], obs[1], obs[2], tab, ne, p.sub, part, range);
  }
  return hipGetLastError();
}

hipError_t launch_spectrum_fold(const double* part, int nchunk, int ne, const double* kw, double c2, double* out,
                                hipStream_t s) {
  spectrum_fold_kernel<<<(ne + kFoldE - 1) / kFoldE, kFoldE * kFoldSlices, 0, s>>>(part, nchunk, ne, kw, c2, out);
  return hipGetLastError();
}

}  // namespace srs
