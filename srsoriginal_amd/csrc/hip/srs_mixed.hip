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

// Whether any point of [i0, i1) does not travel along z, from the span flags (block-uniform,
// NT threads).
template <int NT = kBlock>
__device__ __forceinline__ bool chunk_transverse(const int* __restrict__ flags, int64_t i0, int64_t i1) {
  int f = 0;
  for (int64_t b = i0 / kFlagSpan + threadIdx.x; b <= (i1 - 1) / kFlagSpan; b += NT) f |= flags[b];
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
                                                             ChunkMap cm, float* __restrict__ part) {
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
  int64_t const i0 = cm.begin(blockIdx.y);
  int64_t const i1 = cm.end(blockIdx.y, t.n);
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
        T.raw.t[k] = t.dt * (double)(base + k);
        T.raw.x[k] = t.x[i];
        T.raw.y[k] = t.y[i];
        T.raw.z[k] = t.z[i];
        T.raw.bx[k] = t.bx[i];
        T.raw.by[k] = t.by[i];
        T.raw.bz[k] = t.bz[i];
      }
      __syncthreads();
      flux_terms_fp64<NE, TILE_>(T.raw, cnt, ox, oy, oz, es, d);
    }
    if (j < ns) {
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        size_t const b = ((size_t)(blockIdx.y * NE + e) * 6) * ns + j;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          part[b + (2 * k) * (size_t)ns] = (float)d.re[e][k];
          part[b + (2 * k + 1) * (size_t)ns] = (float)d.im[e][k];
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
        T.raw.t[k] = t.dt * (double)(base + k);
        T.raw.x[k] = t.x[i];
        T.raw.y[k] = t.y[i];
        T.raw.z[k] = t.z[i];
        T.raw.fbx[k] = tf[i];
        T.raw.fby[k] = tf[t.n + i];
        T.raw.fomb[k] = tf[2 * t.n + i];
      }
      __syncthreads();
      flux_terms_mixed<NE, TILE_>(T.raw, cnt, ox, oy, oz, es, cwf, a);
    }
  }
  if (j < ns) {
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      double tc = 1.0, ts = 0.0;  // paraxial sums take the plane's phase omega Z/c
      if (para) para_plane_turn(es.omega[e], pb.z, &tc, &ts);
      size_t const b = ((size_t)(blockIdx.y * NE + e) * 6) * ns + j;
#pragma unroll
      for (int q = 0; q < 6; ++q)
        part[b + (size_t)q * ns] = (float)turned([&](int k) { return a.value(e, k); }, q, tc, ts);
    }
  }
}

// One energy, two observation points per lane: a block of NT lanes owns points j0 + tid
// and j0 + NT + tid (j0 = 2 NT blockIdx.x) and one trajectory chunk; the paraxial records
// are read once for both points (flux_terms_para2).  Chunks whose beam is not along z
// and blocks that fail the paraxial check sum the FP64 / general mixed terms one point
// after the other.  part as flux_mixed_kernel<1, TILE>.
template <int TILE_, int NT>
__global__ __launch_bounds__(NT) void flux_mixed2_kernel(TrajView t, const float* __restrict__ tf,
                                                              const int* __restrict__ flags,
                                                              const double* __restrict__ obs, int ns, EnergySet es,
                                                              ChunkMap cm, float* __restrict__ part) {
  __shared__ MixedTile<TILE_, 1> T;
  __shared__ double red[4 * (NT / 64)];
  int const tid = threadIdx.x;
  int const j0 = blockIdx.x * (2 * NT);
  int const jp[2] = {j0 + tid, j0 + NT + tid};
  double ox[2], oy[2], oz[2];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    int const jj = jp[p] < ns ? jp[p] : j0;  // idle lanes repeat the block's first point
    ox[p] = obs[jj];
    oy[p] = obs[ns + jj];
    oz[p] = obs[2 * ns + jj];
  }
  float cwf = (float)es.cw[0];
  int64_t const i0 = cm.begin(blockIdx.y);
  int64_t const i1 = cm.end(blockIdx.y, t.n);
  if (chunk_transverse<NT>(flags, i0, i1)) {
    // the beam of this chunk is not along z: the FP64 terms, one point after the other
#pragma unroll 1
    for (int p = 0; p < 2; ++p) {
      // selects, not indexing: a runtime index would put the arrays in scratch memory
      double const px = p ? ox[1] : ox[0], py = p ? oy[1] : oy[0], pz = p ? oz[1] : oz[0];
      int const j = p ? jp[1] : jp[0];
      AccFP64<1> d;
      d.zero();
      for (int64_t base = i0; base < i1; base += TILE_) {
        int const cnt = (i1 - base) < TILE_ ? (int)(i1 - base) : TILE_;
        __syncthreads();
        for (int k = tid; k < cnt; k += NT) {
          int64_t const i = base + k;
          T.raw.t[k] = t.dt * (double)(base + k);
          T.raw.x[k] = t.x[i];
          T.raw.y[k] = t.y[i];
          T.raw.z[k] = t.z[i];
          T.raw.bx[k] = t.bx[i];
          T.raw.by[k] = t.by[i];
          T.raw.bz[k] = t.bz[i];
        }
        __syncthreads();
        flux_terms_fp64<1, TILE_>(T.raw, cnt, px, py, pz, es, d);
      }
      if (j < ns) {
        size_t const b = ((size_t)blockIdx.y * 6) * ns + j;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          part[b + (2 * k) * (size_t)ns] = (float)d.re[0][k];
          part[b + (2 * k + 1) * (size_t)ns] = (float)d.im[0][k];
        }
      }
    }
    return;
  }
  double const z0 = obs[2 * ns + j0];
  ParaBlock const pb = para_block_range<NT>(__builtin_fmin(ox[0], ox[1]), __builtin_fmax(ox[0], ox[1]),
                                        __builtin_fmin(oy[0], oy[1]), __builtin_fmax(oy[0], oy[1]),
                                        oz[0] == z0 && oz[1] == z0, z0, red);
  bool const para = para_check<NT>(pb, t.x, t.y, t.z, i0, i1, es);
  AccMixed<1> a[2];
  a[0].zero();
  a[1].zero();
  if (para) {
    for (int64_t base = i0; base < i1; base += TILE_) {
      int const cnt = (i1 - base) < TILE_ ? (int)(i1 - base) : TILE_;
      __syncthreads();
      int const cntp = (cnt + kSub - 1) / kSub * kSub;
      for (int k = tid; k < cntp; k += NT) {
        int64_t const i = base + k;
        if (k < cnt)
          para_stage<TILE_, 1>(T.para, k, t.x[i], t.y[i], t.z[i], tf[i], tf[t.n + i], tf[2 * t.n + i], i, t.dt, pb,
                               es);
        else
          para_pad<TILE_, 1>(T.para, k);
      }
      __syncthreads();
      flux_terms_para2<TILE_>(T.para, cntp, ox, oy, cwf, a);
    }
  } else {
    for (int64_t base = i0; base < i1; base += TILE_) {
      int const cnt = (i1 - base) < TILE_ ? (int)(i1 - base) : TILE_;
      __syncthreads();
      for (int k = tid; k < cnt; k += NT) {
        int64_t const i = base + k;
        T.raw.t[k] = t.dt * (double)(base + k);
        T.raw.x[k] = t.x[i];
        T.raw.y[k] = t.y[i];
        T.raw.z[k] = t.z[i];
        T.raw.fbx[k] = tf[i];
        T.raw.fby[k] = tf[t.n + i];
        T.raw.fomb[k] = tf[2 * t.n + i];
      }
      __syncthreads();
#pragma unroll
      for (int p = 0; p < 2; ++p)
        flux_terms_mixed<1, TILE_>(T.raw, cnt, ox[p], oy[p], oz[p], es, &cwf, a[p]);
    }
  }
  double tc = 1.0, ts = 0.0;  // paraxial sums take the plane's phase omega Z/c
  if (para) para_plane_turn(es.omega[0], pb.z, &tc, &ts);
#pragma unroll
  for (int p = 0; p < 2; ++p)
    if (jp[p] < ns) {
      size_t const b = ((size_t)blockIdx.y * 6) * ns + jp[p];
#pragma unroll
      for (int q = 0; q < 6; ++q)
        part[b + (size_t)q * ns] = (float)turned([&](int k) { return a[p].value(0, k); }, q, tc, ts);
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
    case 1: return blocks_per_cu(flux_mixed2_kernel<kFlux2Threads, kFlux2Threads>, kFlux2Threads, c[1]);
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
                             ChunkMap const& cm, float* part, hipStream_t s) {
  if (cm.c1 < 1 || cm.c2 < 1 || cm.nch < 1 || cm.end(cm.nch - 1, t.n) != t.n) return hipErrorInvalidValue;
  dim3 grid((ns + kBlock - 1) / kBlock, cm.nch);
  const int* const fl = reinterpret_cast<const int*>(trajf + 3 * t.n);  // staged without accelerations
  switch (es.ne) {
    case 1: {
      // one energy: two observation points per lane
      constexpr int ppb = flux_mixed_points_per_block(1);
      dim3 const grid2((ns + ppb - 1) / ppb, cm.nch);
      flux_mixed2_kernel<kFlux2Threads, kFlux2Threads>
          <<<grid2, kFlux2Threads, 0, s>>>(t, trajf, fl, obs_soa, ns, es, cm, part);
      break;
    }
    case 2: flux_mixed_kernel<2, TILE><<<grid, kBlock, 0, s>>>(t, trajf, fl, obs_soa, ns, es, cm, part); break;
    case 3: flux_mixed_kernel<3, TILE><<<grid, kBlock, 0, s>>>(t, trajf, fl, obs_soa, ns, es, cm, part); break;
    case 4: flux_mixed_kernel<4, TILE><<<grid, kBlock, 0, s>>>(t, trajf, fl, obs_soa, ns, es, cm, part); break;
    default: return hipErrorInvalidValue;
  }
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
