// Mixed-precision radiation kernels for gfx950.
//
// FP64 is kept where stability needs it: the observer distance D and the
// retarded time tau = dt*i + D/c, the phase omega*tau (up to 5e11 rad) and its
// reduction to [-pi, pi].  Everything after the reduction runs in FP32: sin/cos
// of the reduced phase (v_sin_f32 / v_cos_f32), the amplitude terms and the
// accumulation.  Cancellation-prone differences are rewritten so that FP32 sees
// only well-conditioned quantities:
//   beta_z - n_z = (1 - n_z) - (1 - beta_z),  1 - beta_z stored per point (FP64 -> FP32),
//   1 - n_z = (n_x^2 + n_y^2) / (1 + n_z)          for n_z > 0.
// Sums run in FP32 over 32-point sub-tiles; the sub-tile sums are added to FP32
// accumulators with Kahan compensation; each trajectory chunk writes FP64 partial
// sums and the cross-chunk reduction is FP64 (shared with the plain form).
#include "srs_kernels.h"
#include "srs_pd.h"
#include "srs_terms.h"

namespace srs {
namespace {

constexpr int TILE = kBlock;
constexpr int SUB = kSub;



// trajf layout (float): [0] bx, [1] by, [2] 1 - bz, and with_a: [3] ax, [4] ay, [5] az,
// [6] 1 - |beta|^2 (from FP64: (1 - bz)(1 + bz) - bx^2 - by^2); after the rows, one int
// per 256 points (kFlagSpan), nonzero when a point of that span does not travel along z
// (transverse_beam).  The caller states with_a and sizes trajf with traj_f32_floats.
// Launched with 256-lane blocks: block b writes flag b.
__global__ __launch_bounds__(kFlagSpan) void traj_to_f32_kernel(TrajView t, int with_a, float* __restrict__ f) {
  int64_t const i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool tr = false;
  if (i < t.n) {
    f[0 * t.n + i] = (float)t.bx[i];
    f[1 * t.n + i] = (float)t.by[i];
    f[2 * t.n + i] = (float)(1.0 - t.bz[i]);
    if (with_a) {
      f[3 * t.n + i] = (float)t.ax[i];
      f[4 * t.n + i] = (float)t.ay[i];
      f[5 * t.n + i] = (float)t.az[i];
      double const omb = 1.0 - t.bz[i];
      f[6 * t.n + i] = (float)(omb * (1.0 + t.bz[i]) - t.bx[i] * t.bx[i] - t.by[i] * t.by[i]);
    }
    tr = transverse_beam(t.bx[i], t.by[i]);
  }
  int const any = __syncthreads_or(tr);
  if (threadIdx.x == 0) reinterpret_cast<int*>(f + (with_a ? 7 : 3) * t.n)[blockIdx.x] = any;
}

// Whether any point of [i0, i1) does not travel along z, from the span flags (block-uniform).
__device__ __forceinline__ bool chunk_transverse(const int* __restrict__ flags, int64_t i0, int64_t i1) {
  int f = 0;
  for (int64_t b = i0 / kFlagSpan + threadIdx.x; b <= (i1 - 1) / kFlagSpan; b += kBlock) f |= flags[b];
  return __syncthreads_or(f) != 0;
}

// A block owns 256 observation points and one trajectory chunk.  It takes the
// paraxial terms when its points lie on one plane z = Z and every (point,
// trajectory point) pair of the chunk passes their bounds (a block-uniform
// decision), and the plain mixed terms otherwise.
template <int NE, int TILE_>
__global__ __launch_bounds__(kBlock) void flux_mixed_kernel(TrajView t, const float* __restrict__ tf,
                                                             const int* __restrict__ flags,
                                                             const double* __restrict__ obs, int ns, EnergySet es,
                                                             int chunk, double* __restrict__ part) {
  __shared__ MixedTile<TILE_, NE> T;
  __shared__ double red[4 * (kBlock / 64)];
  int const tid = threadIdx.x;
  int const j0 = blockIdx.x * kBlock;
  int const j = j0 + tid;
  int const jj = j < ns ? j : j0;  // idle lanes repeat the block's first point
  double const ox = obs[jj], oy = obs[ns + jj], oz = obs[2 * ns + jj];
  float cwf[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) cwf[e] = (float)es.cw[e];
  int64_t const i0 = (int64_t)blockIdx.y * chunk;
  int64_t const i1 = (i0 + chunk < t.n) ? i0 + chunk : t.n;
  if (chunk_transverse(flags, i0, i1)) {
    // the beam of this chunk is not along z, which the mixed terms assume (they stage
    // 1 - beta_z): the block sums the FP64 terms instead
    AccFP64<NE> d;
    d.zero();
    for (int64_t base = i0; base < i1; base += TILE_) {
      int const cnt = (i1 - base) < TILE_ ? (int)(i1 - base) : TILE_;
      __syncthreads();
      for (int k = tid; k < cnt; k += kBlock) {
        int64_t const i = base + k;
        T.raw.x[k] = t.x[i];
        T.raw.y[k] = t.y[i];
        T.raw.z[k] = t.z[i];
        T.raw.bx[k] = t.bx[i];
        T.raw.by[k] = t.by[i];
        T.raw.bz[k] = t.bz[i];
      }
      __syncthreads();
      flux_terms_fp64<NE, TILE_>(T.raw, cnt, base, t.dt, ox, oy, oz, es, d);
    }
    if (j < ns) {
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        size_t const b = ((size_t)(blockIdx.y * NE + e) * 6) * ns + j;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          part[b + (2 * k) * (size_t)ns] = d.re[e][k];
          part[b + (2 * k + 1) * (size_t)ns] = d.im[e][k];
        }
      }
    }
    return;
  }
  ParaBlock const pb = para_block(ox, oy, oz, obs[2 * ns + j0], red);
  bool const para = para_check(pb, t.x, t.y, t.z, i0, i1, es);
  AccMixed<NE> a;
  a.zero();
  if (para) {
    for (int64_t base = i0; base < i1; base += TILE_) {
      int const cnt = (i1 - base) < TILE_ ? (int)(i1 - base) : TILE_;
      __syncthreads();
      int const cntp = (cnt + kSub - 1) / kSub * kSub;
      for (int k = tid; k < cntp; k += kBlock) {
        int64_t const i = base + k;
        if (k < cnt)
          para_stage<TILE_, NE>(T.para, k, t.x[i], t.y[i], t.z[i], tf[i], tf[t.n + i], tf[2 * t.n + i], i, t.dt, pb,
                                es);
        else
          para_pad<TILE_, NE>(T.para, k);
      }
      __syncthreads();
      flux_terms_para<NE, TILE_>(T.para, cntp, ox, oy, es, cwf, a);
    }
  } else {
    for (int64_t base = i0; base < i1; base += TILE_) {
      int const cnt = (i1 - base) < TILE_ ? (int)(i1 - base) : TILE_;
      __syncthreads();
      for (int k = tid; k < cnt; k += kBlock) {
        int64_t const i = base + k;
        T.raw.x[k] = t.x[i];
        T.raw.y[k] = t.y[i];
        T.raw.z[k] = t.z[i];
        T.raw.fbx[k] = tf[i];
        T.raw.fby[k] = tf[t.n + i];
        T.raw.fomb[k] = tf[2 * t.n + i];
      }
      __syncthreads();
      flux_terms_mixed<NE, TILE_>(T.raw, cnt, base, t.dt, ox, oy, oz, es, cwf, a);
    }
  }
  if (j < ns) {
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      size_t const b = ((size_t)(blockIdx.y * NE + e) * 6) * ns + j;
#pragma unroll
      for (int q = 0; q < 6; ++q) part[b + (size_t)q * ns] = a.value(e, q);
    }
  }
}

// Spectrum with the energy-grid recurrence: a lane owns RUN consecutive energies.
// For each trajectory point it anchors exp(i omega_0 tau) with an FP64-reduced
// phase and steps exp(i d_omega tau) across the run (FP64-reduced as well); the
// run restarts from a fresh FP64 anchor, so rounding drift is bounded by RUN steps.
// Uniform steps are exact to the grid's own rounding (the reference computes each
// energy separately, TSpectrumContainer.cc:82); d_omega is the run's mean spacing.
//
// The workload is small (C0: 4e7 terms) and a lane's output (6 sums x RUN energies)
// is large next to its work, so the block splits its trajectory chunk among its
// SPW waves (all on the same 64 x RUN energies) and adds the waves' sums in LDS in
// wave order before writing one partial per block: the FP64 partials, and the chunk
// reduction that reads them, shrink SPW-fold.
constexpr int RUN = 8;
constexpr int SPW = 8;  // waves per spectrum block (8 x 64 lanes: 256 VGPRs for the 48 Kahan sums)

__global__ __launch_bounds__(64 * SPW) void spectrum_mixed_kernel(TrajView t, double ox, double oy, double oz,
                                                                   const double* __restrict__ omega, int ne, int chunk,
                                                                   double* __restrict__ part) {
  __shared__ double stau[SPW][64];
  __shared__ float sa[SPW][3][64], sg[SPW][3][64];
  __shared__ double red[SPW][6][64];
  int const tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  int const e0 = (blockIdx.x * 64 + lane) * RUN;
  int const ea = e0 < ne ? e0 : ne - 1;
  int const eb = (e0 + RUN - 1 < ne) ? e0 + RUN - 1 : ne - 1;
  double const w0 = omega[ea];
  double const dw = eb > ea ? (omega[eb] - omega[ea]) / (double)(eb - ea) : 0.0;
  int64_t const i0 = (int64_t)blockIdx.y * chunk;
  int64_t const i1 = (i0 + chunk < t.n) ? i0 + chunk : t.n;
  // this wave's part of the chunk; every wave runs the same number of 64-point tiles
  int64_t const sub = ((int64_t)chunk + SPW - 1) / SPW;
  int64_t const w0i = i0 + wave * sub;
  int64_t const w1i = (w0i + sub < i1) ? w0i + sub : i1;
  int const ntiles = (int)((sub + 63) / 64);
  float cw[RUN];
#pragma unroll
  for (int r = 0; r < RUN; ++r) {
    int const e = (e0 + r < ne) ? e0 + r : ne - 1;
    cw[r] = (float)(kC / omega[e]);
  }
  KahanF ex[RUN][6];
#pragma unroll
  for (int r = 0; r < RUN; ++r)
#pragma unroll
    for (int q = 0; q < 6; ++q) ex[r][q] = KahanF{0.f, 0.f};

  for (int tt = 0; tt < ntiles; ++tt) {
    int64_t const base = w0i + (int64_t)tt * 64;
    int const cnt = base < w1i ? ((w1i - base) < 64 ? (int)(w1i - base) : 64) : 0;
    __syncthreads();
    if (lane < cnt) {
      int64_t const i = base + lane;
      double const rx = ox - t.x[i], ry = oy - t.y[i], rz = oz - t.z[i];
      double const d2 = rx * rx + ry * ry + rz * rz;
      double const D = __builtin_sqrt(d2);
      stau[wave][lane] = t.dt * (double)i + D / kC;
      double const inv = 1.0 / D;
      double const nx = rx * inv, ny = ry * inv, nz = rz * inv;
      sa[wave][0][lane] = (float)((t.bx[i] - nx) * inv);
      sa[wave][1][lane] = (float)((t.by[i] - ny) * inv);
      sa[wave][2][lane] = (float)((t.bz[i] - nz) * inv);
      double const inv2 = inv * inv;
      sg[wave][0][lane] = (float)(nx * inv2);
      sg[wave][1][lane] = (float)(ny * inv2);
      sg[wave][2][lane] = (float)(nz * inv2);
    }
    __syncthreads();
    for (int k0 = 0; k0 < cnt; k0 += SUB) {
      int const k1 = (k0 + SUB < cnt) ? k0 + SUB : cnt;
      float sub_s[RUN][6];
#pragma unroll
      for (int r = 0; r < RUN; ++r)
#pragma unroll
        for (int q = 0; q < 6; ++q) sub_s[r][q] = 0.f;
      for (int k = k0; k < k1; ++k) {
        double const tau = stau[wave][k];
        float ps, pc, ws, wc;
        sincos_reduced_f32(reduce_two_pi(w0 * tau), &ps, &pc);
        sincos_reduced_f32(reduce_two_pi(dw * tau), &ws, &wc);
        float const a0 = sa[wave][0][k], a1 = sa[wave][1][k], a2 = sa[wave][2][k];
        float const g0 = sg[wave][0][k], g1 = sg[wave][1][k], g2 = sg[wave][2][k];
#pragma unroll
        for (int r = 0; r < RUN; ++r) {
          float const cs = cw[r] * ps, cc = cw[r] * pc;
          sub_s[r][0] = __builtin_fmaf(a0, pc, __builtin_fmaf(g0, cs, sub_s[r][0]));
          sub_s[r][1] = __builtin_fmaf(a0, ps, __builtin_fmaf(-g0, cc, sub_s[r][1]));
          sub_s[r][2] = __builtin_fmaf(a1, pc, __builtin_fmaf(g1, cs, sub_s[r][2]));
          sub_s[r][3] = __builtin_fmaf(a1, ps, __builtin_fmaf(-g1, cc, sub_s[r][3]));
          sub_s[r][4] = __builtin_fmaf(a2, pc, __builtin_fmaf(g2, cs, sub_s[r][4]));
          sub_s[r][5] = __builtin_fmaf(a2, ps, __builtin_fmaf(-g2, cc, sub_s[r][5]));
          // advance the phasor to the next energy of the run
          float const npc = __builtin_fmaf(pc, wc, -ps * ws);
          float const nps = __builtin_fmaf(ps, wc, pc * ws);
          pc = npc;
          ps = nps;
        }
      }
#pragma unroll
      for (int r = 0; r < RUN; ++r)
#pragma unroll
        for (int q = 0; q < 6; ++q) ex[r][q].add(sub_s[r][q]);
    }
  }
  // combine the waves in wave order, one energy of the run at a time; partial layout
  // part[(c*6 + q)*ne + e] holds the complex field sums (re/im per axis)
#pragma unroll
  for (int r = 0; r < RUN; ++r) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 6; ++q) red[wave][q][lane] = (double)ex[r][q].s - (double)ex[r][q].c;
    __syncthreads();
    if (tid < 6 * 64) {
      int const q = tid / 64, l = tid % 64;
      double v = 0.0;
#pragma unroll
      for (int w = 0; w < SPW; ++w) v += red[w][q][l];
      int const e = (blockIdx.x * 64 + l) * RUN + r;
      if (e < ne) part[((size_t)blockIdx.y * 6 + q) * ne + e] = v;
    }
  }
}

// Mixed power density, FP32 throughout the term.  With u = n - beta and |n| = 1
// (n is the unit vector (n_x, n_y, 1 - (1 - n_z)) whatever the rounding of 1 - n_z):
//   1 - beta.n     = (|u|^2 + 1 - |beta|^2) / 2      (a sum of non-negative terms)
//   n x (u x A)    = u (n.A) - A (1 - beta.n)
// The first form has no cancellation: 1 - beta.n (down to 1/(2 gamma^2) ~ 1.5e-8 at
// 3 GeV) is not a difference of terms ~1e-6, and an error in 1 - n_z moves |n|^2 - 1
// and u_z^2 by amounts that cancel to first order.  1 - |beta|^2 is taken per point in
// FP64 (traj_to_f32_kernel).  1 - n_z = (n_x^2 + n_y^2)/(1 + n_z) for n_z > 0, else
// (1 - n_z)^2/(1 - n_z).  Positions rounded to FP32 keep R = O - X to
// ~1e-7 relative, and 1/D is one v_rsq_f32.
//
// Scaling: n and beta are carried as S n, S beta (S = 2^10, staged scaled in LDS), so
// d2 = S^2 |u|^2 + S^2 (1 - |beta|^2) = 2^21 (1 - beta.n) stays far from the FP32
// range limits when raised to the -5th power, and the term comes out as 2^-55 times
// its value; 2^55 is restored in FP64 when the sum is written.  32-term FP32 sub-sums,
// Kahan-compensated.
// The FP64 power density of one chunk (the reference's terms, every component of N1 as
// its quotient) for chunks whose beam is not along z, where the mixed form's staged
// 1 - beta_z does not apply.  sd: 9 x TILE doubles of LDS.
template <bool DIR>
__device__ __forceinline__ double pd_fp64_chunk(TrajView const& t, double* sd, int64_t i0, int64_t i1, double ox,
                                                double oy, double oz, double qx, double qy, double qz) {
  int const tid = threadIdx.x;
  double sum = 0.0;
  for (int64_t base = i0; base < i1; base += TILE) {
    int const cnt = (i1 - base) < TILE ? (int)(i1 - base) : TILE;
    __syncthreads();
    if (tid < cnt) {
      int64_t const i = base + tid;
      double const v[9] = {t.x[i], t.y[i], t.z[i], t.bx[i], t.by[i], t.bz[i], t.ax[i], t.ay[i], t.az[i]};
#pragma unroll
      for (int c = 0; c < 9; ++c) sd[c * TILE + tid] = v[c];
    }
    __syncthreads();
    for (int k = 0; k < cnt; ++k)
      sum += pd_term<true>(ox, oy, oz, qx, qy, qz, sd[k], sd[TILE + k], sd[2 * TILE + k], sd[3 * TILE + k],
                           sd[4 * TILE + k], sd[5 * TILE + k], sd[6 * TILE + k], sd[7 * TILE + k], sd[8 * TILE + k],
                           DIR ? 1 : 0);
  }
  return sum;
}

template <bool DIR>
__global__ __launch_bounds__(kBlock) void pd_mixed_kernel(TrajView t, const float* __restrict__ tf,
                                                          const int* __restrict__ flags,
                                                          const double* __restrict__ obs, const double* __restrict__ nrm,
                                                          int ns, int chunk, double* __restrict__ part) {
  // (x, y, z, S^2 (1 - |beta|^2)), S (bx, by, 1 - bz), A; or the FP64 tile of pd_fp64_chunk
  __shared__ union {
    struct {
      float4 P[TILE], B[TILE], A[TILE];
    } f;
    double d[9 * TILE];
  } sm;
  float4* const sP = sm.f.P;
  float4* const sB = sm.f.B;
  float4* const sA = sm.f.A;
  int const tid = threadIdx.x;
  int const j = blockIdx.x * kBlock + tid;
  int const jj = j < ns ? j : ns - 1;
  int64_t const i0 = (int64_t)blockIdx.y * chunk;
  int64_t const i1 = (i0 + chunk < t.n) ? i0 + chunk : t.n;
  if (chunk_transverse(flags, i0, i1)) {
    double const sum = pd_fp64_chunk<DIR>(t, sm.d, i0, i1, obs[jj], obs[ns + jj], obs[2 * ns + jj], nrm[jj],
                                          nrm[ns + jj], nrm[2 * ns + jj]);
    if (j < ns) part[(size_t)blockIdx.y * ns + j] = sum;
    return;
  }
  float const ox = (float)obs[jj], oy = (float)obs[ns + jj], oz = (float)obs[2 * ns + jj];
  float const qx = (float)nrm[jj], qy = (float)nrm[ns + jj], qz = (float)nrm[2 * ns + jj];
  int64_t const n = t.n;
  KahanF acc{0.f, 0.f};
  for (int64_t base = i0; base < i1; base += TILE) {
    int const cnt = (i1 - base) < TILE ? (int)(i1 - base) : TILE;
    __syncthreads();
    if (tid < cnt) {
      int64_t const i = base + tid;
      sP[tid] = make_float4((float)t.x[i], (float)t.y[i], (float)t.z[i], tf[6 * n + i] * (kPdS * kPdS));
      sB[tid] = make_float4(tf[i] * kPdS, tf[n + i] * kPdS, tf[2 * n + i] * kPdS, 0.f);
      sA[tid] = make_float4(tf[3 * n + i], tf[4 * n + i], tf[5 * n + i], 0.f);
    }
    __syncthreads();
    for (int k0 = 0; k0 < cnt; k0 += SUB) {
      float sub = 0.f;
      if (cnt - k0 >= SUB) {
#pragma unroll 4
        for (int k = k0; k < k0 + SUB; ++k) sub += pd_mixed_term<DIR>(ox, oy, oz, qx, qy, qz, sP[k], sB[k], sA[k]);
      } else {
        for (int k = k0; k < cnt; ++k) sub += pd_mixed_term<DIR>(ox, oy, oz, qx, qy, qz, sP[k], sB[k], sA[k]);
      }
      acc.add(sub);
    }
  }
  if (j < ns) part[(size_t)blockIdx.y * ns + j] = ((double)acc.s - (double)acc.c) * 0x1p55;
}

inline int nchunks(int64_t n, int chunk) { return (int)((n + chunk - 1) / chunk); }

}  // namespace

// Resident blocks per CU of a kernel (its VGPR, LDS and wave limits), cached.
template <class K>
static int blocks_per_cu(K kernel, int threads, int& cache) {
  if (cache == 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(kernel), threads, 0) !=
            hipSuccess ||
        nb < 1)
      nb = 1;
    cache = nb;
  }
  return cache;
}

int flux_mixed_blocks_per_cu(int ne) {
  static int c[5];
  switch (ne) {
    case 1: return blocks_per_cu(flux_mixed_kernel<1, TILE>, kBlock, c[1]);
    case 2: return blocks_per_cu(flux_mixed_kernel<2, TILE>, kBlock, c[2]);
    case 3: return blocks_per_cu(flux_mixed_kernel<3, TILE>, kBlock, c[3]);
    default: return blocks_per_cu(flux_mixed_kernel<4, TILE>, kBlock, c[4]);
  }
}

int pd_mixed_blocks_per_cu(int directional) {
  static int c[2];
  return directional ? blocks_per_cu(pd_mixed_kernel<true>, kBlock, c[1])
                     : blocks_per_cu(pd_mixed_kernel<false>, kBlock, c[0]);
}

size_t traj_f32_floats(int64_t n, int with_a) { return (size_t)(with_a ? 7 : 3) * n + (n + kFlagSpan - 1) / kFlagSpan; }

hipError_t launch_traj_to_f32(const TrajView& t, int with_a, float* trajf, hipStream_t s) {
  if (with_a && !(t.ax && t.ay && t.az)) return hipErrorInvalidValue;
  traj_to_f32_kernel<<<(unsigned)((t.n + kFlagSpan - 1) / kFlagSpan), kFlagSpan, 0, s>>>(t, with_a, trajf);
  return hipGetLastError();
}

hipError_t launch_flux_mixed(const TrajView& t, const float* trajf, const double* obs_soa, int ns, const EnergySet& es,
                             int chunk, double* part, hipStream_t s) {
  dim3 grid((ns + kBlock - 1) / kBlock, nchunks(t.n, chunk));
  const int* const fl = reinterpret_cast<const int*>(trajf + 3 * t.n);  // staged without accelerations
  switch (es.ne) {
    case 1: flux_mixed_kernel<1, TILE><<<grid, kBlock, 0, s>>>(t, trajf, fl, obs_soa, ns, es, chunk, part); break;
    case 2: flux_mixed_kernel<2, TILE><<<grid, kBlock, 0, s>>>(t, trajf, fl, obs_soa, ns, es, chunk, part); break;
    case 3: flux_mixed_kernel<3, TILE><<<grid, kBlock, 0, s>>>(t, trajf, fl, obs_soa, ns, es, chunk, part); break;
    case 4: flux_mixed_kernel<4, TILE><<<grid, kBlock, 0, s>>>(t, trajf, fl, obs_soa, ns, es, chunk, part); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

int spectrum_mixed_groups(int ne) { return (ne + 64 * RUN - 1) / (64 * RUN); }

hipError_t launch_spectrum_mixed(const TrajView& t, const double obs[3], const double* omega, const double*, int ne,
                                 int chunk, int, double* part, hipStream_t s) {
  dim3 grid(spectrum_mixed_groups(ne), nchunks(t.n, chunk));
  spectrum_mixed_kernel<<<grid, 64 * SPW, 0, s>>>(t, obs[0], obs[1], obs[2], omega, ne, chunk, part);
  return hipGetLastError();
}

hipError_t launch_pd_mixed(const TrajView& t, const float* trajf, const double* obs_soa, const double* nrm_soa, int ns,
                           int directional, int chunk, double* part, hipStream_t s) {
  dim3 grid((ns + kBlock - 1) / kBlock, nchunks(t.n, chunk));
  const int* const fl = reinterpret_cast<const int*>(trajf + 7 * t.n);  // staged with accelerations
  if (directional)
    pd_mixed_kernel<true><<<grid, kBlock, 0, s>>>(t, trajf, fl, obs_soa, nrm_soa, ns, chunk, part);
  else
    pd_mixed_kernel<false><<<grid, kBlock, 0, s>>>(t, trajf, fl, obs_soa, nrm_soa, ns, chunk, part);
  return hipGetLastError();
}

}  // namespace srs
