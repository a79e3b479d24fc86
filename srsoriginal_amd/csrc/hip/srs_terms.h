// Per-term inner loops of the spectrum/flux integrand, shared by the
// single-particle (chunked) and the batched (persistent work-queue) kernels.
//
// A lane owns one observation point and NE photon energies and walks a run of
// trajectory points staged in LDS.  The integrand is the reference's
// (OSCARSSR.cc:2054-2092):
//   E_k += [(beta_k - n_k)/D - i (c/omega) n_k / D^2] exp(i omega (dt*i + D/c)).
#pragma once
#include "srs_device.h"
#include "srs_kernels.h"

namespace srs {

// LDS tile of trajectory points for the flux kernels.
template <int TILE>
struct FluxTile {
  double x[TILE], y[TILE], z[TILE];
  double bx[TILE], by[TILE], bz[TILE];  // FP64 form
  float fbx[TILE], fby[TILE], fomb[TILE];  // mixed form: beta_x, beta_y, 1 - beta_z
};

template <int NE>
struct AccFP64 {
  double re[NE][3], im[NE][3];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int e = 0; e < NE; ++e)
#pragma unroll
      for (int k = 0; k < 3; ++k) re[e][k] = im[e][k] = 0.0;
  }
};

// Plain FP64 form: phase chain in the reference's operation order (bitwise),
// N = R/D correctly rounded (beta_z - n_z cancels to ~1e-8).
template <int NE, int TILE>
__device__ __forceinline__ void flux_terms_fp64(FluxTile<TILE> const& T, int cnt, int64_t base, double dt, double ox,
                                                double oy, double oz, EnergySet const& es, AccFP64<NE>& a) {
  for (int k = 0; k < cnt; ++k) {
    double const rx = ox - T.x[k], ry = oy - T.y[k], rz = oz - T.z[k];
    double const d2 = rx * rx + ry * ry + rz * rz;
    double const D = __builtin_sqrt(d2);
    double const tau = dt * (double)(base + k) + div_rn(D, kC, kInvC);
    double const inv = 1.0 / D;
    double const nx = div_rn(rx, D, inv), ny = div_rn(ry, D, inv), nz = div_rn(rz, D, inv);
    double const ax = (T.bx[k] - nx) * inv, ay = (T.by[k] - ny) * inv, az = (T.bz[k] - nz) * inv;
    double const inv2 = inv * inv;
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      double s, c;
      sincos_fp64(es.omega[e] * tau, &s, &c);
      double const g = es.cw[e] * inv2;
      double const gx = nx * g, gy = ny * g, gz = nz * g;
      a.re[e][0] = __builtin_fma(ax, c, __builtin_fma(gx, s, a.re[e][0]));
      a.re[e][1] = __builtin_fma(ay, c, __builtin_fma(gy, s, a.re[e][1]));
      a.re[e][2] = __builtin_fma(az, c, __builtin_fma(gz, s, a.re[e][2]));
      a.im[e][0] = __builtin_fma(ax, s, __builtin_fma(-gx, c, a.im[e][0]));
      a.im[e][1] = __builtin_fma(ay, s, __builtin_fma(-gy, c, a.im[e][1]));
      a.im[e][2] = __builtin_fma(az, s, __builtin_fma(-gz, c, a.im[e][2]));
    }
  }
}

struct KahanF {
  float s, c;
  __device__ __forceinline__ void add(float v) {
    float const y = v - c;
    float const t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

template <int NE>
struct AccMixed {
  KahanF q[NE][6];  // re_x, im_x, re_y, im_y, re_z, im_z
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int e = 0; e < NE; ++e)
#pragma unroll
      for (int k = 0; k < 6; ++k) q[e][k] = KahanF{0.f, 0.f};
  }
  __device__ __forceinline__ double value(int e, int k) const { return (double)q[e][k].s - (double)q[e][k].c; }
};

constexpr int kSub = 32;  // FP32 sub-tile length of the mixed sums

__device__ __forceinline__ void sincos_reduced_f32(double r, float* s, float* c) {
  // r in [-pi, pi] -> revolutions for the hardware sin/cos (abs error ~1.2e-7)
  float const rev = (float)(r * 0.15915494309189535);
  *s = __builtin_amdgcn_sinf(rev);
  *c = __builtin_amdgcn_cosf(rev);
}

// Mixed form: D, tau and omega*tau in FP64 (bitwise as the reference), exact
// reduction to [-pi, pi]; trig, amplitudes and the 32-term sub-tile sums in FP32;
// sub-tile sums folded into Kahan-compensated FP32 accumulators.
template <int NE, int TILE>
__device__ __forceinline__ void flux_terms_mixed(FluxTile<TILE> const& T, int cnt, int64_t base, double dt, double ox,
                                                 double oy, double oz, EnergySet const& es, const float* cwf,
                                                 AccMixed<NE>& a) {
  for (int k0 = 0; k0 < cnt; k0 += kSub) {
    int const k1 = (k0 + kSub < cnt) ? k0 + kSub : cnt;
    float sub[NE][6];
#pragma unroll
    for (int e = 0; e < NE; ++e)
#pragma unroll
      for (int q = 0; q < 6; ++q) sub[e][q] = 0.f;
    for (int k = k0; k < k1; ++k) {
      double const rx = ox - T.x[k], ry = oy - T.y[k], rz = oz - T.z[k];
      double const d2 = rx * rx + ry * ry + rz * rz;
      double const D = __builtin_sqrt(d2);
      double const tau = dt * (double)(base + k) + div_rn(D, kC, kInvC);
      float const inv = __builtin_amdgcn_rcpf((float)D);
      float const nx = (float)rx * inv, ny = (float)ry * inv, nz = (float)rz * inv;
      float const v = __builtin_fmaf(nx, nx, ny * ny);
      float const omnz = nz > 0.f ? v * __builtin_amdgcn_rcpf(1.f + nz) : 1.f - nz;
      float const ax = (T.fbx[k] - nx) * inv, ay = (T.fby[k] - ny) * inv, az = (omnz - T.fomb[k]) * inv;
      float const g = inv * inv;
      float const gx = nx * g, gy = ny * g, gz = nz * g;
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        float s, c;
        sincos_reduced_f32(reduce_two_pi(es.omega[e] * tau), &s, &c);
        float const cs = cwf[e] * s, cc = cwf[e] * c;
        sub[e][0] = __builtin_fmaf(ax, c, __builtin_fmaf(gx, cs, sub[e][0]));
        sub[e][1] = __builtin_fmaf(ax, s, __builtin_fmaf(-gx, cc, sub[e][1]));
        sub[e][2] = __builtin_fmaf(ay, c, __builtin_fmaf(gy, cs, sub[e][2]));
        sub[e][3] = __builtin_fmaf(ay, s, __builtin_fmaf(-gy, cc, sub[e][3]));
        sub[e][4] = __builtin_fmaf(az, c, __builtin_fmaf(gz, cs, sub[e][4]));
        sub[e][5] = __builtin_fmaf(az, s, __builtin_fmaf(-gz, cc, sub[e][5]));
      }
    }
#pragma unroll
    for (int e = 0; e < NE; ++e)
#pragma unroll
      for (int q = 0; q < 6; ++q) a.q[e][q].add(sub[e][q]);
  }
}

// Flux of one energy from the complex field sums (OSCARSSR.cc:2080-2092):
// SumE *= (0 + i kw) -> (-im kw, re kw); F = C2 * sum_k |.|^2 in the order x, y, z.
__device__ __forceinline__ double flux_from_sums(double rex, double imx, double rey, double imy, double rez, double imz,
                                                 double kw, double c2) {
  double const ax = -(imx * kw), bx = rex * kw;
  double const ay = -(imy * kw), by = rey * kw;
  double const az = -(imz * kw), bz = rez * kw;
  return c2 * ((ax * ax + bx * bx) + (ay * ay + by * by) + (az * az + bz * bz));
}

}  // namespace srs
