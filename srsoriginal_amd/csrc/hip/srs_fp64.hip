// Plain FP64 radiation kernels for gfx950.
//
// These reproduce the reference CPU integrands (OSCARSSR.cc:2054-2092 flux,
// :1522-1573 power density; the spectrum is in srs_spectrum.hip) with the phase path in the
// reference's exact operation order, so results agree with the reference to
// rounding of the amplitude terms (well inside 1e-10 relative):
//   R = O - X;  D = sqrt((Rx*Rx + Ry*Ry) + Rz*Rz);  tau = dt*i + D/C;  N = R/D
//   phi = omega*tau  (all correctly rounded, no contraction: this file is
//   compiled with -ffp-contract=off and FMAs appear only where written).
// The phase reaches 4.6e11 rad at 3 keV, where one ulp of D is 5e-5 rad, so any
// reordering of that chain would break parity; amplitudes tolerate ulp changes.
//
// Work decomposition (VALU-bound, trig/division heavy): a block owns 256 output
// points (observation points) and one chunk of the
// trajectory; the chunk is staged through LDS in 256-point tiles that every lane
// reads by broadcast.  Each chunk writes FP64 partial sums; a second kernel adds
// the chunks in fixed order (deterministic) and applies the prefactors.
#include "srs_kernels.h"
#include "srs_pd.h"
#include "srs_terms.h"

namespace srs {
namespace {

constexpr int TILE = kBlock;

template <int NE, int TILE_>
__global__ __launch_bounds__(kBlock) void flux_fp64_kernel(TrajView t, const double* __restrict__ obs, int ns,
                                                            EnergySet es, ChunkMap cm, double* __restrict__ part) {
  __shared__ FluxTile<TILE_> T;
  int const tid = threadIdx.x;
  int const j = blockIdx.x * kBlock + tid;
  int const jj = j < ns ? j : ns - 1;  // padding lanes compute a valid point, never store
  double const ox = obs[jj], oy = obs[ns + jj], oz = obs[2 * ns + jj];
  int64_t const i0 = cm.begin(blockIdx.y);
  int64_t const i1 = cm.end(blockIdx.y, t.n);
  AccFP64<NE> a;
  a.zero();
  for (int64_t base = i0; base < i1; base += TILE_) {
    int const cnt = (i1 - base) < TILE_ ? (int)(i1 - base) : TILE_;
    __syncthreads();
    for (int k = tid; k < cnt; k += kBlock) {
      T.t[k] = t.dt * (double)(base + k);
      T.x[k] = t.x[base + k];
      T.y[k] = t.y[base + k];
      T.z[k] = t.z[base + k];
      T.bx[k] = t.bx[base + k];
      T.by[k] = t.by[base + k];
      T.bz[k] = t.bz[base + k];
    }
    __syncthreads();
    flux_terms_fp64<NE, TILE_>(T, cnt, ox, oy, oz, es, a);
  }
  if (j < ns) {
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      size_t const b = ((size_t)(blockIdx.y * NE + e) * 6) * ns + j;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        part[b + (2 * k) * (size_t)ns] = a.re[e][k];
        part[b + (2 * k + 1) * (size_t)ns] = a.im[e][k];
      }
    }
  }
}

// Power density: one lane per observation point, the chunk's trajectory points
// staged in LDS, pd_term per pair (reference operation order for N1, N1.n and
// 1 - beta.N1; (Num.N2)^2 + (Num.N3)^2 as |Num|^2).
__global__ __launch_bounds__(kBlock) void pd_fp64_kernel(TrajView t, const double* __restrict__ obs,
                                                         const double* __restrict__ nrm, int ns, int directional,
                                                         int chunk, double* __restrict__ part) {
  __shared__ double sx[TILE], sy[TILE], sz[TILE], sbx[TILE], sby[TILE], sbz[TILE], sax[TILE], say[TILE], saz[TILE];
  int const tid = threadIdx.x;
  int const j = blockIdx.x * kBlock + tid;
  int const jj = j < ns ? j : ns - 1;
  double const ox = obs[jj], oy = obs[ns + jj], oz = obs[2 * ns + jj];
  double const qx = nrm[jj], qy = nrm[ns + jj], qz = nrm[2 * ns + jj];
  int64_t const i0 = (int64_t)blockIdx.y * chunk;
  int64_t const i1 = (i0 + chunk < t.n) ? i0 + chunk : t.n;
  double sum = 0.0;
  for (int64_t base = i0; base < i1; base += TILE) {
    int const cnt = (i1 - base) < TILE ? (int)(i1 - base) : TILE;
    __syncthreads();
    bool tr = false;
    if (tid < cnt) {
      int64_t const i = base + tid;
      sx[tid] = t.x[i];
      sy[tid] = t.y[i];
      sz[tid] = t.z[i];
      sbx[tid] = t.bx[i];
      sby[tid] = t.by[i];
      sbz[tid] = t.bz[i];
      sax[tid] = t.ax[i];
      say[tid] = t.ay[i];
      saz[tid] = t.az[i];
      tr = transverse_beam(sbx[tid], sby[tid]);
    }
    if (__syncthreads_or(tr)) {  // block-uniform: a point of this tile does not travel along z
      for (int k = 0; k < cnt; ++k)
        sum += pd_term<true>(ox, oy, oz, qx, qy, qz, sx[k], sy[k], sz[k], sbx[k], sby[k], sbz[k], sax[k], say[k],
                             saz[k], directional);
    } else {
      for (int k = 0; k < cnt; ++k)
        sum += pd_term(ox, oy, oz, qx, qy, qz, sx[k], sy[k], sz[k], sbx[k], sby[k], sbz[k], sax[k], say[k], saz[k],
                       directional);
    }
  }
  if (j < ns) part[(size_t)blockIdx.y * ns + j] = sum;
}

// Time-domain electric field at one observer (OSCARSSR.cc:2646-2690), one lane per
// trajectory point, the reference's TVector3D operation order:
//   E = C0 (1/(1 - n.beta))^3 [ (1 - |beta|^2)(n - beta)/|R|^2 + (1/c) n x ((n - beta) x A/c)/|R| ],
//   t = i dt + |R|/c,  C0 = q/(4 pi eps0).
__global__ __launch_bounds__(kBlock) void efield_time_kernel(TrajView t, double ox, double oy, double oz, double c0,
                                                             double* __restrict__ e, double* __restrict__ tm) {
  int64_t const i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= t.n) return;
  double const bx = t.bx[i], by = t.by[i], bz = t.bz[i];
  double const ax = t.ax[i], ay = t.ay[i], az = t.az[i];
  double const rx = ox - t.x[i], ry = oy - t.y[i], rz = oz - t.z[i];
  double const r2 = rx * rx + ry * ry + rz * rz;
  double const D = __builtin_sqrt(r2);
  double const nx = rx / D, ny = ry / D, nz = rz / D;
  double const w = 1.0 / (1.0 - (nx * bx + ny * by + nz * bz));
  double const mult = c0 * pow(w, 3.0);
  double const ux = nx - bx, uy = ny - by, uz = nz - bz;
  double const sb = 1.0 - (bx * bx + by * by + bz * bz);
  double const vx = uy * az - ay * uz, vy = uz * ax - az * ux, vz = ux * ay - ax * uy;
  double const fx = ny * vz - vy * nz, fy = nz * vx - vz * nx, fz = nx * vy - vx * ny;
  double const ic = 1.0 / kC;
  double const E[3] = {((ux * sb) / r2 + (fx * ic) / D) * mult, ((uy * sb) / r2 + (fy * ic) / D) * mult,
                       ((uz * sb) / r2 + (fz * ic) / D) * mult};
#pragma unroll
  for (int k = 0; k < 3; ++k) e[3 * i + k] = E[k];
  tm[i] = (double)i * t.dt + D / kC;
}

// Near- plus far-field flux of OSCARSSR::CalculateFlux1 (OSCARSSR.cc:2209-2346): the
// Lienard-Wiechert field with the stored A/c, no integration by parts,
//   E_k = [ (1 - |b|^2)(n - b) / (D D (1 - n.b)^2) + n x ((n - b) x A/c) / (D (1 - n.b)^2) ]
//         e^{-i Omega (k dt + D/c)},
// one lane per observation point, the trajectory staged in LDS tiles.  1 - n.b and
// 1 - |b|^2 lose ~8 digits to cancellation at GeV energies, so n_z = R_z/|R| (IEEE
// quotient), the dot products and the phase keep the reference's TVector3D operation
// order; n_x, n_y and the well-conditioned amplitude quotients use reciprocals.
// part[(c * 6 + 2 q + {0: re, 1: im}) * ns + j]: the chunk's sum for axis q (times dt
// later, with C0, in flux_reduce_kernel).
__global__ __launch_bounds__(kBlock) void flux1_fp64_kernel(TrajView t, const double* __restrict__ obs, int ns,
                                                            double omega, int chunk, double* __restrict__ part) {
  __shared__ double sx[TILE], sy[TILE], sz[TILE], sbx[TILE], sby[TILE], sbz[TILE], sax[TILE], say[TILE], saz[TILE];
  int const tid = threadIdx.x;
  int const j = blockIdx.x * kBlock + tid;
  int const jj = j < ns ? j : ns - 1;
  double const ox = obs[jj], oy = obs[ns + jj], oz = obs[2 * ns + jj];
  int64_t const i0 = (int64_t)blockIdx.y * chunk;
  int64_t const i1 = (i0 + chunk < t.n) ? i0 + chunk : t.n;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t base = i0; base < i1; base += TILE) {
    int const cnt = (i1 - base) < TILE ? (int)(i1 - base) : TILE;
    __syncthreads();
    bool tr = false;
    if (tid < cnt) {
      int64_t const i = base + tid;
      sx[tid] = t.x[i];
      sy[tid] = t.y[i];
      sz[tid] = t.z[i];
      sbx[tid] = t.bx[i];
      sby[tid] = t.by[i];
      sbz[tid] = t.bz[i];
      sax[tid] = t.ax[i];
      say[tid] = t.ay[i];
      saz[tid] = t.az[i];
      tr = transverse_beam(sbx[tid], sby[tid]);
    }
    // block-uniform: a point of this tile does not travel along z -> n_x, n_y as quotients
    bool const exact = __syncthreads_or(tr);
    for (int k = 0; k < cnt; ++k) {
      double const bx = sbx[k], by = sby[k], bz = sbz[k];
      double const rx = ox - sx[k], ry = oy - sy[k], rz = oz - sz[k];
      double const D = __builtin_sqrt(rx * rx + ry * ry + rz * rz);
      double const inv = 1.0 / D;
      // nz correctly rounded (1 - n.b and nz - bz cancel through it); nx, ny by the
      // reciprocal (their last bit is ~1e-15 of 1 - n.b) unless the beam is not along z
      double const nx = exact ? div_rn(rx, D, inv) : rx * inv, ny = exact ? div_rn(ry, D, inv) : ry * inv;
      double const nz = div_rn(rz, D, inv);
      double const omb = 1.0 - (nx * bx + ny * by + nz * bz);
      double const sb = 1.0 - (bx * bx + by * by + bz * bz);
      double const ux = nx - bx, uy = ny - by, uz = nz - bz;
      double const ax = sax[k], ay = say[k], az = saz[k];
      double const vx = uy * az - ay * uz, vy = uz * ax - az * ux, vz = ux * ay - ax * uy;
      double const fx = ny * vz - vy * nz, fy = nz * vx - vz * nx, fz = nx * vy - vx * ny;
      double const ip2 = 1.0 / (omb * omb);
      double const k2 = ip2 * inv;        // 1 / (D (1 - n.b)^2)
      double const k1 = (sb * k2) * inv;  // (1 - |b|^2) / (D D (1 - n.b)^2)
      double const Vx = ux * k1 + fx * k2, Vy = uy * k1 + fy * k2, Vz = uz * k1 + fz * k2;
      double const tau = t.dt * (double)(base + k) + div_rn(D, kC, kInvC);
      double s, c;
      sincos_fp64(omega * tau, &s, &c);  // e^{-i phi} = (c, -s)
      acc[0] = __builtin_fma(Vx, c, acc[0]);
      acc[1] = __builtin_fma(-Vx, s, acc[1]);
      acc[2] = __builtin_fma(Vy, c, acc[2]);
      acc[3] = __builtin_fma(-Vy, s, acc[3]);
      acc[4] = __builtin_fma(Vz, c, acc[4]);
      acc[5] = __builtin_fma(-Vz, s, acc[5]);
    }
  }
  if (j < ns) {
#pragma unroll
    for (int q = 0; q < 6; ++q) part[((size_t)blockIdx.y * 6 + q) * ns + j] = acc[q];
  }
}

// Chunk reductions (fixed order) and the reference's final prefactors:
//   flux:  SumE *= C1*dt (OSCARSSR.cc:2080) -> (re, im) * (0 + i kw) = (-im kw, re kw);
//          F = C2 * Re(SumE . conj(SumE)) summed x, y then z (OSCARSSR.cc:2092).
// A block folds 64 points: wave w adds chunks w, w + 4, w + 8, ... of each point in
// order, then wave 0 adds the four slice sums in order 0, 1, 2, 3 (a fixed order: the
// result does not depend on timing).  Four waves per 64 points keep four times the
// loads in flight of one lane per point.
template <class P>
__global__ __launch_bounds__(256) void flux_reduce_kernel(const P* __restrict__ part, int nchunk, int ns,
                                                          EnergySet es, double c2, double* __restrict__ out) {
  __shared__ double red[3][6][64];
  int const lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int const j = blockIdx.x * 64 + lane;
  int const e = blockIdx.y;
  double v[6] = {0, 0, 0, 0, 0, 0};
  if (j < ns) {
    for (int c = w; c < nchunk; c += 4) {
      size_t const b = ((size_t)(c * es.ne + e) * 6) * ns + j;
#pragma unroll
      for (int k = 0; k < 6; ++k) v[k] += (double)part[b + (size_t)k * ns];
    }
  }
  if (w > 0) {
#pragma unroll
    for (int k = 0; k < 6; ++k) red[w - 1][k][lane] = v[k];
  }
  __syncthreads();
  if (w > 0 || j >= ns) return;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] += red[r][k][lane];
  double const kw = es.kw[e];
  double p[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    double const re = -(v[2 * q + 1] * kw);
    double const im = v[2 * q] * kw;
    p[q] = re * re + im * im;
  }
  out[(size_t)e * ns + j] = c2 * (p[0] + p[1] + p[2]);
}

__global__ void pd_reduce_kernel(const double* __restrict__ part, int nchunk, int ns, double scale, int directional,
                                 double* __restrict__ out) {
  int const j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ns) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * ns + j];
  // OSCARSSR.cc:1563-1573: Sum *= |qI|/(16 pi^2 eps0 c) dt; Sum /= 1e6; |Sum| unless directional
  s *= scale;
  s /= 1e6;
  if (!directional && s < 0) s = -s;
  out[j] = s;
}

inline int nchunks(int64_t n, int chunk) { return (int)((n + chunk - 1) / chunk); }

}  // namespace

// Total power sum over the trajectory (OSCARSSR.cc:1966-1971): one block of 1024 lanes,
// each lane a strided share in order, then the lanes' sums in a fixed tree order.
__global__ __launch_bounds__(1024) void total_power_kernel(TrajView t, double scale, double* __restrict__ out) {
  __shared__ double red[1024];
  double acc = 0.0, comp = 0.0;  // compensated lane sum
  for (int64_t i = threadIdx.x; i < t.n; i += 1024) {
    double const ax = t.ax[i], ay = t.ay[i], az = t.az[i];
    double const bx = t.bx[i], by = t.by[i], bz = t.bz[i];
    double const cx = by * az - bz * ay, cy = bz * ax - bx * az, cz = bx * ay - by * ax;
    double const v = ((ax * ax + ay * ay + az * az) - (cx * cx + cy * cy + cz * cz)) * t.dt;
    double const y = v - comp;
    double const s = acc + y;
    comp = (s - acc) - y;
    acc = s;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0] * scale;
}

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

int flux_fp64_blocks_per_cu(int ne) {
  static int c[5];
  switch (ne) {
    case 1: return blocks_per_cu(flux_fp64_kernel<1, TILE>, kBlock, c[1]);
    case 2: return blocks_per_cu(flux_fp64_kernel<2, TILE>, kBlock, c[2]);
    case 3: return blocks_per_cu(flux_fp64_kernel<3, TILE>, kBlock, c[3]);
    default: return blocks_per_cu(flux_fp64_kernel<4, TILE>, kBlock, c[4]);
  }
}

int pd_fp64_blocks_per_cu() {
  static int c;
  return blocks_per_cu(pd_fp64_kernel, kBlock, c);
}

int flux1_fp64_blocks_per_cu() {
  static int c;
  return blocks_per_cu(flux1_fp64_kernel, kBlock, c);
}

hipError_t launch_flux1_fp64(const TrajView& t, const double* obs_soa, int ns, double omega, int chunk, double* part,
                             hipStream_t s) {
  dim3 grid((ns + kBlock - 1) / kBlock, nchunks(t.n, chunk));
  flux1_fp64_kernel<<<grid, kBlock, 0, s>>>(t, obs_soa, ns, omega, chunk, part);
  return hipGetLastError();
}

hipError_t launch_flux_fp64(const TrajView& t, const double* obs_soa, int ns, const EnergySet& es,
                            ChunkMap const& cm, double* part, hipStream_t s) {
  if (cm.c1 < 1 || cm.c2 < 1 || cm.nch < 1 || cm.end(cm.nch - 1, t.n) != t.n) return hipErrorInvalidValue;
  dim3 grid((ns + kBlock - 1) / kBlock, cm.nch);
  switch (es.ne) {
    case 1: flux_fp64_kernel<1, TILE><<<grid, kBlock, 0, s>>>(t, obs_soa, ns, es, cm, part); break;
    case 2: flux_fp64_kernel<2, TILE><<<grid, kBlock, 0, s>>>(t, obs_soa, ns, es, cm, part); break;
    case 3: flux_fp64_kernel<3, TILE><<<grid, kBlock, 0, s>>>(t, obs_soa, ns, es, cm, part); break;
    case 4: flux_fp64_kernel<4, TILE><<<grid, kBlock, 0, s>>>(t, obs_soa, ns, es, cm, part); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}


hipError_t launch_pd_fp64(const TrajView& t, const double* obs_soa, const double* nrm_soa, int ns, int directional,
                          int chunk, double* part, hipStream_t s) {
  dim3 grid((ns + kBlock - 1) / kBlock, nchunks(t.n, chunk));
  pd_fp64_kernel<<<grid, kBlock, 0, s>>>(t, obs_soa, nrm_soa, ns, directional, chunk, part);
  return hipGetLastError();
}

hipError_t launch_efield_time(const TrajView& t, const double obs[3], double c0, double* e, double* tm, hipStream_t s) {
  efield_time_kernel<<<(unsigned)((t.n + kBlock - 1) / kBlock), kBlock, 0, s>>>(t, obs[0], obs[1], obs[2], c0, e, tm);
  return hipGetLastError();
}

hipError_t launch_total_power(const TrajView& t, double scale, double* out, hipStream_t s) {
  total_power_kernel<<<1, 1024, 0, s>>>(t, scale, out);
  return hipGetLastError();
}

hipError_t launch_flux_reduce(const double* part, int nchunk, int ns, const EnergySet& es, double c2, double* out,
                              hipStream_t s) {
  dim3 grid((ns + 63) / 64, es.ne);
  flux_reduce_kernel<double><<<grid, 256, 0, s>>>(part, nchunk, ns, es, c2, out);
  return hipGetLastError();
}

hipError_t launch_flux_reduce(const float* part, int nchunk, int ns, const EnergySet& es, double c2, double* out,
                              hipStream_t s) {
  dim3 grid((ns + 63) / 64, es.ne);
  flux_reduce_kernel<float><<<grid, 256, 0, s>>>(part, nchunk, ns, es, c2, out);
  return hipGetLastError();
}



hipError_t launch_pd_reduce(const double* part, int nchunk, int ns, double scale, int directional, double* out,
                            hipStream_t s) {
  pd_reduce_kernel<<<(ns + 255) / 256, 256, 0, s>>>(part, nchunk, ns, scale, directional, out);
  return hipGetLastError();
}

}  // namespace srs
