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
  double t[TILE];  // dt * i, the point's time in the reference's rounding, formed once when staged
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

// |R| = RN(sqrt(d2)) for the squared distance of a term (sqrt_rn from the hardware
// reciprocal square root: the correctly rounded root without the scaling and special
// cases of a general square root, which squared distances never need).
__device__ __forceinline__ double dist_rn(double d2) { return sqrt_rn(d2, __builtin_amdgcn_rsq(d2)); }

// Plain FP64 form: phase chain in the reference's operation order (bitwise),
// n_z = R_z/D correctly rounded (beta_z - n_z cancels to ~1e-8, so its rounding is
// bit-relevant); n_x, n_y by the reciprocal (beta_x - n_x, beta_y - n_y are ~K/gamma,
// no cancellation: an ulp of n_x is 1e-16 of a typical term).
template <int NE, int TILE>
__device__ __forceinline__ void flux_terms_fp64(FluxTile<TILE> const& T, int cnt, double ox, double oy, double oz,
                                                EnergySet const& es, AccFP64<NE>& a) {
  for (int k = 0; k < cnt; ++k) {
    double const rx = ox - T.x[k], ry = oy - T.y[k], rz = oz - T.z[k];
    double const d2 = rx * rx + ry * ry + rz * rz;
    double const D = dist_rn(d2);
    double const tau = T.t[k] + div_rn(D, kC, kInvC);
    double const inv = 1.0 / D;
    double const nx = rx * inv, ny = ry * inv, nz = div_rn(rz, D, inv);
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
__device__ __forceinline__ void flux_terms_mixed(FluxTile<TILE> const& T, int cnt, double ox, double oy, double oz,
                                                 EnergySet const& es, const float* cwf, AccMixed<NE>& a) {
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
      double const D = dist_rn(d2);
      double const tau = T.t[k] + div_rn(D, kC, kInvC);
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

// ---------------------------------------------------------------------------
// Paraxial flat-screen form of the mixed integrand.
//
// For a block whose observation points all lie on one plane z = Z, with
// rx = ox - x_k, ry = oy - y_k, s = rx^2 + ry^2, rz = Z - z_k and eps = s / rz^2:
//   D = rz + s/(2 rz) - s^2/(8 rz^3) + rz*O(eps^3)
//   omega*tau = omega*Z/c + omega*(t_k - z_k/c) + omega*(D - rz)/c.
// omega*Z/c is the same for every term of one observation point and cancels in
// |E|^2.  In revolutions (f = omega/2pi, W = f/c):
//   phase = Phi_k + W*u,  Phi_k = frac(f*(t_k - z_k/c)),  u = s/(2 rz) - s^2/(8 rz^3)
//   (W*u up to ~1e3 rev at 431 eV; FP64).
// Phi_k and 1/(2 rz) are formed once per staged trajectory point (compensated FP64);
// per term only rx, ry, s, u and W*u + Phi_k are FP64 -- no sqrt, no division.
// With 1/D = m/rz, m = 1 - eps/2 + O(eps^2):
//   (beta_x - n_x)/D = m (beta_x/rz - rx/rz^2) (1 + O(eps))
//   (beta_z - n_z)/D = m (s/(2 rz^2) - (1 - beta_z)) / rz (1 + O(eps))
//   n/D^2            = (rx/rz^3, ry/rz^3, 1/rz^2) (1 + O(eps)).
// m varies by < eps*|d rz|/rz over a 32-point sub-tile; it scales the sub-tile sums.
// A block takes this form only when every (point, trajectory point) pair it
// integrates has eps <= kParaEps (relative amplitude error O(eps)),
// W*rz*eps^3/16 <= kParaPhase (dropped phase term, revolutions) and
// (c/omega)/rz <= kParaNear (x, y near field below FP32 resolution); otherwise it
// integrates with flux_terms_mixed.
// ---------------------------------------------------------------------------
constexpr double kParaEps = 1e-6;
constexpr double kParaPhase = 1e-9;
// x, y near field: the amplitude (beta - n)/D - i (c/omega) n/D^2 of the x, y axes is
// beta/D - (n/D)(1 + i (c/omega)/D); the block drops the factor i (c/omega)/D when it
// is below kParaNear = 1e-9 (~2^-30, under FP32 resolution 2^-24: in FP32 it would
// round away).  The z near field -i (c/omega) n_z/D^2 is kept (the z far field
// (beta_z - n_z)/D is itself ~1e-8 of the x, y one).
constexpr double kParaNear = 1e-9;

// The phase the paraxial terms leave out.  Their phase is omega tau - omega Z/c for the
// plane z = Z: omega Z/c is the same for every term of one observation point, so |E|^2
// does not see it, but the sums that the other forms give for the same point (other
// trajectory chunks or time segments, outside the paraxial bounds) carry it.  Every
// paraxial sum is therefore turned by e^{i omega Z/c} before it is stored or combined:
// psi = omega Z / (2 pi c) in revolutions, reduced to [-1/2, 1/2] in double-double
// arithmetic (omega Z as an exact pair, 1/(2 pi c) as a pair); its cosine and sine by the
// hardware (revolutions, ~1e-7 absolute: the sums' own FP32 level), scaled in FP64 to
// unit modulus so that |E|^2 is unchanged.  *c, *s = cos, sin of 2 pi psi.
constexpr double kInv2PiCHi = 0x1.23db3eb8a3fa7p-31, kInv2PiCLo = -0x1.79488f78d104bp-86;  // 1/(2 pi c)
__device__ __forceinline__ double para_plane_phase(double omega, double Z) {
  double const ah = omega * Z, al = __builtin_fma(omega, Z, -ah);
  double const yh = ah * kInv2PiCHi;
  double const yl = __builtin_fma(ah, kInv2PiCHi, -yh) + (ah * kInv2PiCLo + al * kInv2PiCHi);
  return (yh - __builtin_rint(yh)) + yl;
}
__device__ __forceinline__ void para_plane_turn(double omega, double Z, double* c, double* s) {
  float const psi = (float)para_plane_phase(omega, Z);
  double const cf = (double)__builtin_amdgcn_cosf(psi), sf = (double)__builtin_amdgcn_sinf(psi);
  double const r = 1.0 / __builtin_sqrt(__builtin_fma(cf, cf, sf * sf));
  *c = cf * r;
  *s = sf * r;
}

// Component q6 (re_x, im_x, re_y, im_y, re_z, im_z) of the sums v(q6) turned by (c, s).
template <class V>
__device__ __forceinline__ double turned(V const& v, int q6, double c, double s) {
  double const re = v(q6 & ~1), im = v(q6 | 1);
  return (q6 & 1) ? __builtin_fma(re, s, im * c) : __builtin_fma(re, c, -(im * s));
}

// frac(f (t_k - z_k/c)) in revolutions, t_k = RN(dt*ti) as the reference forms it
// (OSCARSSR.cc:2072): v = t_k - z_k/c as an unevaluated sum vh + vl (two-sum, minus
// the error of the quotient z_k/c), then f v with the product's error.
__device__ __forceinline__ double para_point_phase(double f, double t, double zk) {
  double const zc = zk * kInvC;
  double const ez = __builtin_fma(-zc, kC, zk) * kInvC;  // z_k/c - zc
  double const vh = t - zc;
  double const tv = vh - t;
  double const vl = ((t - (vh - tv)) + (-zc - tv)) - ez;
  double const p = f * vh;
  double const pl = __builtin_fma(f, vh, -p) + f * vl;
  return (p - __builtin_rint(p)) + pl;
}

// One staged trajectory point of the paraxial form (one LDS record, so a term
// reads it from a single base address).
//
// With rx = ox - x the x far-field amplitude splits into a part of the trajectory
// point and the observation coordinate times a part shared by x and y:
//   beta_x/rz - rx/rz^2 = ax - ox q,  ax = beta_x/rz + x/rz^2,  q = 1/rz^2,
// and the z near-field amplitude -i cw/rz^2 is the same q: a term accumulates
// sum ax e, sum ay e, sum q e and sum az e (e = e^{i phi}), and the observation
// coordinates and c/omega are applied per sub-tile.  s comes from
// s = (ox^2 + oy^2) + (x^2 + y^2) - 2 (ox x + oy y) (absolute error ~1e-20 m^2 on
// these screens, 1e-13 rev of phase).
template <int NE>
struct alignas(16) ParaPoint {
  double x, y;        // trajectory position
  double bsq;         // x^2 + y^2
  double h;           // 1 / (2 rz)      (one energy: W / (2 rz), W folded in)
  double t3;          // 1 / (8 rz^3)    (one energy: W / (8 rz^3))
  double phi[NE];     // Phi_k per energy, revolutions in [-0.5, 0.5]
  float ax, ay;       // beta_x / rz + x / rz^2, beta_y / rz + y / rz^2
  float ombr;         // (1 - beta_z) / rz
  float q;            // 1 / rz^2
  float gh;           // 1 / (2 rz^3)
};

template <int TILE, int NE>
struct ParaTile {
  ParaPoint<NE> p[TILE];
};

// LDS of the mixed kernels: the paraxial tile or, for tiles that fail the check,
// the plain mixed tile.
template <int TILE, int NE>
union MixedTile {
  FluxTile<TILE> raw;
  ParaTile<TILE, NE> para;
};

// Plane and bounding box of a block's observation points.
struct ParaBlock {
  double z;
  double xlo, xhi, ylo, yhi;
  bool flat;  // every point of the block has oz == z
};

// Called by every lane of the block (NT threads) with the range [xlo, xhi] x [ylo, yhi]
// of its points and whether they all lie on z = z0; lanes without a point pass a point
// of the block.  red holds 4 * (NT / 64) doubles of LDS.
template <int NT = kBlock>
__device__ __forceinline__ ParaBlock para_block_range(double xlo, double xhi, double ylo, double yhi, bool on_plane,
                                                      double z0, double* red) {
  double v[4] = {xlo, -xhi, ylo, -yhi};  // minima of these give xlo, xhi, ylo, yhi
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = __builtin_fmin(v[q], __shfl_xor(v[q], o));
  int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < 4; ++q) red[wave * 4 + q] = v[q];
  bool const flat = __syncthreads_and(on_plane) != 0;  // also orders the LDS writes above
  ParaBlock b;
  b.z = z0;
  b.flat = flat;
  double m[4] = {red[0], red[1], red[2], red[3]};
  for (int w = 1; w < NT / 64; ++w)
#pragma unroll
    for (int q = 0; q < 4; ++q) m[q] = __builtin_fmin(m[q], red[w * 4 + q]);
  b.xlo = m[0];
  b.xhi = -m[1];
  b.ylo = m[2];
  b.yhi = -m[3];
  return b;
}

// One point per lane.
__device__ __forceinline__ ParaBlock para_block(double ox, double oy, double oz, double z0, double* red) {
  return para_block_range(ox, ox, oy, oy, oz == z0, z0, red);
}

// True when trajectory point (xk, yk, zk) satisfies the expansion bounds for every
// observation point of the block.
__device__ __forceinline__ bool para_point_ok(ParaBlock const& b, double xk, double yk, double zk, double fwmax,
                                              double cwmax) {
  double const rz = b.z - zk;
  double const dx = __builtin_fmax(__builtin_fabs(b.xlo - xk), __builtin_fabs(b.xhi - xk));
  double const dy = __builtin_fmax(__builtin_fabs(b.ylo - yk), __builtin_fabs(b.yhi - yk));
  double const eps = __builtin_fma(dx, dx, dy * dy) / (rz * rz);
  return rz > 0.0 && eps <= kParaEps && fwmax * rz * (eps * eps * eps) <= 16.0 * kParaPhase &&
         cwmax <= kParaNear * rz;  // D >= rz
}

// Block-uniform decision for trajectory points [i0, i1) of arrays x, y, z: every
// lane checks a strided share, the block (NT threads) takes the paraxial form only if
// all pass.  Called by every lane of the block.
template <int NT = kBlock>
__device__ __forceinline__ bool para_check(ParaBlock const& b, const double* __restrict__ x,
                                           const double* __restrict__ y, const double* __restrict__ z, int64_t i0,
                                           int64_t i1, EnergySet const& es) {
  double fwmax = 0.0, cwmax = 0.0;
  for (int e = 0; e < es.ne; ++e) {
    fwmax = __builtin_fmax(fwmax, es.fw[e]);
    cwmax = __builtin_fmax(cwmax, es.cw[e]);
  }
  bool ok = b.flat;
  if (ok)
    for (int64_t i = i0 + threadIdx.x; i < i1; i += NT)
      ok = ok && para_point_ok(b, x[i], y[i], z[i], fwmax, cwmax);
  return __syncthreads_and(ok) != 0;
}

// Stage trajectory point k (time index ti) of the paraxial tile.
template <int TILE, int NE>
__device__ __forceinline__ void para_stage(ParaTile<TILE, NE>& P, int k, double xk, double yk, double zk, float fbx,
                                           float fby, float fomb, int64_t ti, double dt, ParaBlock const& b,
                                           EnergySet const& es) {
  double const rz = b.z - zk;
  double const ir = 1.0 / rz;
  float const irf = (float)ir;
  ParaPoint<NE>& r = P.p[k];
  r.x = xk;
  r.y = yk;
  r.bsq = __builtin_fma(xk, xk, yk * yk);
  double const fold = NE == 1 ? es.fw[0] : 1.0;
  r.h = fold * (0.5 * ir);
  r.t3 = fold * (0.125 * ir * ir * ir);
  double const q = ir * ir;
  r.ax = (float)__builtin_fma(xk, q, (double)fbx * ir);
  r.ay = (float)__builtin_fma(yk, q, (double)fby * ir);
  r.ombr = fomb * irf;
  r.q = (float)q;
  r.gh = (float)(0.5 * q * ir);
#pragma unroll
  for (int e = 0; e < NE; ++e) r.phi[e] = para_point_phase(es.fr[e], dt * (double)ti, zk);
}

// A padding record past the tile's last point: zero amplitudes, so its terms add
// nothing and the 32-term sub-tile loop needs no remainder.
template <int TILE, int NE>
__device__ __forceinline__ void para_pad(ParaTile<TILE, NE>& P, int k) {
  ParaPoint<NE>& r = P.p[k];
  r.x = r.y = r.bsq = r.h = r.t3 = 0.0;
#pragma unroll
  for (int e = 0; e < NE; ++e) r.phi[e] = 0.0;
  r.ax = r.ay = r.ombr = r.q = r.gh = 0.f;
}

// The z near-field sum (-i (c/omega) n_z/D^2, through sum q e) is folded in with
// c/omega at the sub-tile's end, so a term needs no per-energy scaling; the x, y near
// field is below kParaNear on these blocks (see above).  cnt is a multiple of kSub (the
// tile is padded with zero-amplitude records), so the sub-tile loop has a fixed trip
// count.
template <int NE, int TILE>
__device__ __forceinline__ void flux_terms_para(ParaTile<TILE, NE> const& P, int cnt, double ox, double oy,
                                                EnergySet const& es, const float* cwf, AccMixed<NE>& a) {
  double const asq = __builtin_fma(ox, ox, oy * oy);
  double const mx = -2.0 * ox, my = -2.0 * oy;
  float const oxf = (float)ox, oyf = (float)oy;
  for (int k0 = 0; k0 < cnt; k0 += kSub) {
    // per energy: sum ax c, ax s, ay c, ay s, q c, q s, az c, az s
    float sub[NE][8];
#pragma unroll
    for (int e = 0; e < NE; ++e)
#pragma unroll
      for (int q = 0; q < 8; ++q) sub[e][q] = 0.f;
    // unrolled by 4 for one energy (one address computation per 4 records: 7 %
    // faster, measured) and by 2 for several (same occupancy; the batched three-energy
    // flux 1.7 % faster than rolled, measured)
#pragma unroll NE == 1 ? 4 : 2
    for (int kk = 0; kk < kSub; ++kk) {
      ParaPoint<NE> const& r = P.p[k0 + kk];
      double const s = __builtin_fma(mx, r.x, __builtin_fma(my, r.y, asq + r.bsq));
      // D - rz = s (h - t3 s)  (times W for one energy, folded into h and t3)
      double const hs = __builtin_fma(-r.t3, s, r.h);
      float const az = __builtin_fmaf((float)s, r.gh, -r.ombr);
      double u = 0.0;
      if constexpr (NE > 1) u = s * hs;
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        // revolutions W (D - rz) + Phi_k, reduced to one turn
        double p;
        if constexpr (NE == 1) p = __builtin_fma(s, hs, r.phi[0]);
        else p = __builtin_fma(es.fw[e], u, r.phi[e]);
        // one exact FP64 instruction reduces to [0, 1) (p - rint(p) takes two)
        float const ph = (float)__builtin_amdgcn_fract(p);
        float const sn = __builtin_amdgcn_sinf(ph), cn = __builtin_amdgcn_cosf(ph);
        float* const b = sub[e];
        b[0] = __builtin_fmaf(r.ax, cn, b[0]);
        b[1] = __builtin_fmaf(r.ax, sn, b[1]);
        b[2] = __builtin_fmaf(r.ay, cn, b[2]);
        b[3] = __builtin_fmaf(r.ay, sn, b[3]);
        b[4] = __builtin_fmaf(r.q, cn, b[4]);
        b[5] = __builtin_fmaf(r.q, sn, b[5]);
        b[6] = __builtin_fmaf(az, cn, b[6]);
        b[7] = __builtin_fmaf(az, sn, b[7]);
      }
    }
    // m = 1 - eps/2 at the sub-tile's first point (always a trajectory point)
    ParaPoint<NE> const& rl = P.p[k0];
    double const sl = __builtin_fma(mx, rl.x, __builtin_fma(my, rl.y, asq + rl.bsq));
    float const m = __builtin_fmaf(-0.5f * (float)sl, rl.q, 1.f);
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      float const* const b = sub[e];
      float const cw = cwf[e];
      // field sums re/im of x, y, z:  x = sum ax e - ox sum q e,  z = sum (az - i cw q) e
      a.q[e][0].add(__builtin_fmaf(-oxf, b[4], b[0]) * m);
      a.q[e][1].add(__builtin_fmaf(-oxf, b[5], b[1]) * m);
      a.q[e][2].add(__builtin_fmaf(-oyf, b[4], b[2]) * m);
      a.q[e][3].add(__builtin_fmaf(-oyf, b[5], b[3]) * m);
      a.q[e][4].add(__builtin_fmaf(cw, b[5], b[6]) * m);
      a.q[e][5].add(__builtin_fmaf(-cw, b[4], b[7]) * m);
    }
  }
}

// The paraxial terms at one energy for two observation points per lane: each staged
// record is read from LDS once and integrated for both points (two independent
// dependency chains per record; half the LDS reads per term).
template <int TILE>
__device__ __forceinline__ void flux_terms_para2(ParaTile<TILE, 1> const& P, int cnt, double const ox[2],
                                                 double const oy[2], float cw, AccMixed<1> a[2]) {
  double asq[2], mx[2], my[2];
  float oxf[2], oyf[2];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    asq[p] = __builtin_fma(ox[p], ox[p], oy[p] * oy[p]);
    mx[p] = -2.0 * ox[p];
    my[p] = -2.0 * oy[p];
    oxf[p] = (float)ox[p];
    oyf[p] = (float)oy[p];
  }
  for (int k0 = 0; k0 < cnt; k0 += kSub) {
    float sub[2][8];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int q = 0; q < 8; ++q) sub[p][q] = 0.f;
    // four records per iteration (eight independent chains): measured 0.8 % faster than
    // two and than one, three or eight
#pragma unroll 4
    for (int kk = 0; kk < kSub; ++kk) {
      ParaPoint<1> const& r = P.p[k0 + kk];
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        double const s = __builtin_fma(mx[p], r.x, __builtin_fma(my[p], r.y, asq[p] + r.bsq));
        double const hs = __builtin_fma(-r.t3, s, r.h);
        float const az = __builtin_fmaf((float)s, r.gh, -r.ombr);
        float const ph = (float)__builtin_amdgcn_fract(__builtin_fma(s, hs, r.phi[0]));
        float const sn = __builtin_amdgcn_sinf(ph), cn = __builtin_amdgcn_cosf(ph);
        float* const b = sub[p];
        b[0] = __builtin_fmaf(r.ax, cn, b[0]);
        b[1] = __builtin_fmaf(r.ax, sn, b[1]);
        b[2] = __builtin_fmaf(r.ay, cn, b[2]);
        b[3] = __builtin_fmaf(r.ay, sn, b[3]);
        b[4] = __builtin_fmaf(r.q, cn, b[4]);
        b[5] = __builtin_fmaf(r.q, sn, b[5]);
        b[6] = __builtin_fmaf(az, cn, b[6]);
        b[7] = __builtin_fmaf(az, sn, b[7]);
      }
    }
    ParaPoint<1> const& rl = P.p[k0];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      double const sl = __builtin_fma(mx[p], rl.x, __builtin_fma(my[p], rl.y, asq[p] + rl.bsq));
      float const m = __builtin_fmaf(-0.5f * (float)sl, rl.q, 1.f);
      float const* const b = sub[p];
      a[p].q[0][0].add(__builtin_fmaf(-oxf[p], b[4], b[0]) * m);
      a[p].q[0][1].add(__builtin_fmaf(-oxf[p], b[5], b[1]) * m);
      a[p].q[0][2].add(__builtin_fmaf(-oyf[p], b[4], b[2]) * m);
      a[p].q[0][3].add(__builtin_fmaf(-oyf[p], b[5], b[3]) * m);
      a[p].q[0][4].add(__builtin_fmaf(cw, b[5], b[6]) * m);
      a[p].q[0][5].add(__builtin_fmaf(-cw, b[4], b[7]) * m);
    }
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
