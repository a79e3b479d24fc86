// Device core of the spectrum kernels (one observer, many photon energies), shared by
// the single-trajectory kernel (srs_spectrum.hip) and the batched one (srs_batch.hip).
//
// The reference evaluates, per energy and trajectory point (OSCARSSR.cc:1202-1246),
//   E += [(beta - n)/D - i (c/omega) n/D^2] exp(i omega tau),  tau = dt*i + D/c,
// with the phase p = RN(omega * tau) rounded once in FP64.  At p ~ 1e11-1e12 rad one
// ulp of p is 1e-5-2e-4 rad: that rounding is part of the reference's result (its
// spectrum differs from the exactly evaluated integral by up to 1e-4 relative at
// 10 keV), so both forms here reproduce RN(omega tau) for every energy.
//
// A lane owns a run of RUN consecutive energies omega_0..omega_{RUN-1} (a uniform grid
// up to rounding) and, per trajectory point:
//   h_0 = RN(omega_0 tau), psi = RN(dw tau)      (dw: the mean spacing of the block's
//                                                 64 runs, the same for all its lanes)
//   p_0 = exp(i h_0),  w = exp(i psi)           (one sin/cos pair each, FP64 reduction;
//                                                 psi and w once per point per block,
//                                                 by the lane that stages the point)
//   p_r = p_{r-1} w                              (the recurrence: angle h_0 + r psi)
//   theta_r = RN(omega_r tau) - h_0 - r psi      (three FP64 operations, exactly: the
//             difference of the two products is exact by Sterbenz's lemma while
//             omega_r / omega_0 stays in [0.6, 1.6], and one FMA rounds the rest)
//   u_r = p_r exp(i theta_r),  |theta_r| ~ 1e-4  (a short series)
// so u_r = exp(i RN(omega_r tau)) up to the rounding of the recurrence.  |theta_r| is at
// most |tau| K for a constant K of the run (its grid's deviation from a line plus the
// three products' rounding bounds); a kernel that meets |tau| > tau_lim = range / K
// reports it and the caller evaluates directly instead.  A run of length 1
// (non-uniform energy lists, or that fallback) is the direct evaluation.
//
// FP64 form: every step in FP64 (6 sums per energy, the near-field factor c/omega
// applied per term).  Mixed form: the phases h_0, psi, omega_r tau and theta_r in
// FP64, everything after them in FP32: sin/cos of the anchor (hardware, on
// revolutions), sin/cos of the step (an FP32 polynomial after an FP64 quadrant
// reduction: its error is multiplied by r), the recurrence, the correction, the
// amplitudes and 32-term FP32 sub-sums, which are added into FP64 accumulators.
#pragma once
#include <type_traits>

#include "srs_device.h"
#include "srs_terms.h"

namespace srs {

// Geometry of one trajectory point for observer o, in the reference's FP64 operation
// order: tau = dt*i + D/c, A = (beta - n)/D, G = n/D^2 (ti = dt * i).
struct SpecPoint {
  double tau;
  double a[3], g[3];
  double d;  // |R|
};

// x, y near field of the mixed form: the amplitude (beta - n)/D - i (c/omega) n/D^2 of
// the x and y axes is ((beta - n) - i ((c/omega)/D) n)/D; a tile drops the second part
// when (c/omega)/D is below kSpecNear = 1e-9 (~2^-30, under FP32 resolution 2^-24) for
// every energy of the wave and every point of the tile, as the paraxial flux terms do.
// The z near field -i (c/omega) n_z/D^2 is always kept (the z far field (beta_z - n_z)/D
// is itself ~1e-8 of the x, y one).  The FP64 form keeps every term.
constexpr double kSpecNear = 1e-9;

// Orders the LDS accesses of one wave's lanes, for a tile that the wave writes and reads
// back by itself (no other wave touches it), in place of a block barrier.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v = __builtin_fmin(v, __shfl_xor(v, m, 64));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v = __builtin_fmax(v, __shfl_xor(v, m, 64));
  return v;
}

__device__ __forceinline__ void spec_geometry(double ox, double oy, double oz, double x, double y, double z,
                                              double bx, double by, double bz, double ti, SpecPoint& P) {
  double const rx = ox - x, ry = oy - y, rz = oz - z;
  double const d2 = rx * rx + ry * ry + rz * rz;
  double const D = __builtin_sqrt(d2);
  P.d = D;
  P.tau = ti + div_rn(D, kC, kInvC);
  double const inv = 1.0 / D;
  double const nx = div_rn(rx, D, inv), ny = div_rn(ry, D, inv), nz = div_rn(rz, D, inv);
  P.a[0] = (bx - nx) * inv;
  P.a[1] = (by - ny) * inv;
  P.a[2] = (bz - nz) * inv;
  double const inv2 = inv * inv;
  P.g[0] = nx * inv2;
  P.g[1] = ny * inv2;
  P.g[2] = nz * inv2;
}

// Forward declaration (defined below).
__device__ __forceinline__ void sincos_quadrant_f32(double x, float* s_out, float* c_out);

// The step of a run for one trajectory point: psi = RN(dw tau) and exp(i psi) (FP32 to
// ~1 ulp for the mixed form, sincos_fp64 for the FP64 form).
template <bool MIXED>
__device__ __forceinline__ void spec_step(double dw, double tau, double& psi,
                                          typename std::conditional<MIXED, float, double>::type& wc,
                                          typename std::conditional<MIXED, float, double>::type& ws) {
  psi = dw * tau;
  if constexpr (MIXED)
    sincos_quadrant_f32(psi, &ws, &wc);
  else
    sincos_fp64(psi, &ws, &wc);
}

// sin and cos of an FP64 angle to FP32 accuracy (~1 ulp): FP64 reduction to
// [-pi/4, pi/4] and the quadrant (as sincos_fp64), then Taylor polynomials through
// r^9 / r^8 in FP32 (truncation < 2.5e-8 on the interval).
__device__ __forceinline__ void sincos_quadrant_f32(double x, float* s_out, float* c_out) {
  double const k = __builtin_rint(x * kTwoOverPi);
  double r = __builtin_fma(-k, kPio2Hi, x);
  r = __builtin_fma(-k, kPio2Lo, r);
  double const kb = k + 0x1.8p52;
  uint64_t bits;
  __builtin_memcpy(&bits, &kb, sizeof bits);
  int const q = (int)(bits & 3u);
  float const f = (float)r, f2 = f * f;
  float ps = __builtin_fmaf(f2, 2.7557319e-06f, -1.9841270e-04f);
  ps = __builtin_fmaf(ps, f2, 8.3333333e-03f);
  ps = __builtin_fmaf(ps, f2, -1.6666667e-01f);
  float const sn = __builtin_fmaf(f * f2, ps, f);
  float pc = __builtin_fmaf(f2, 2.4801587e-05f, -1.3888889e-03f);
  pc = __builtin_fmaf(pc, f2, 4.1666668e-02f);
  pc = __builtin_fmaf(pc, f2, -0.5f);
  float const cs = __builtin_fmaf(f2, pc, 1.0f);
  float const s1 = (q & 1) ? cs : sn;
  float const c1 = (q & 1) ? sn : cs;
  *s_out = (q & 2) ? -s1 : s1;
  *c_out = ((q + 1) & 2) ? -c1 : c1;
}

// Series ranges of the correction exp(i theta): FP32 keeps 1 - theta^2/2 and theta
// (dropped terms < 1e-8 for |theta| <= 4e-3); FP64 keeps 1 - theta^2/2 + theta^4/24 and
// theta - theta^3/6 (dropped terms < 1e-12 for |theta| <= 1e-2).  Where every |theta|
// of a tile is within kThetaF32First = kThetaF32 / 10 the mixed form keeps only
// 1 + i theta (the dropped theta^2/2 < 8e-8, FP32 rounding level); the same bound K
// gives that range at |tau| <= tau_lim / 10.
constexpr float kThetaF32 = 4e-3f;
constexpr double kThetaF64 = 1e-2;
constexpr double kFirstOrderFrac = 0.1;

// The sums of one lane: RUN energies x 6 reals (re/im of the x, y, z field sums).
template <bool MIXED, int RUN>
struct SpecLane {
  using T = typename std::conditional<MIXED, float, double>::type;
  double w[RUN];   // omega of each energy (FP64, as the reference)
  double dw;       // run step
  double tau_lim;  // |tau| below which every |theta_r| of this run is in its series' range
  T cw[RUN];       // c / omega
  double acc[RUN][6];
  T sub[RUN][6];   // mixed: the current FP32 sub-sums; FP64: unused

  // Energies e0 .. e0 + RUN - 1 of the tables omega[ne], cw[ne] (= c/omega) and of the
  // run's step rdw[e0 / RUN] (its block's) and bound rtl[e0 / RUN] (spectrum_run_tables);
  // entries beyond ne continue the run's line (their sums are not written).
  __device__ __forceinline__ void init(const double* __restrict__ omega, const double* __restrict__ cwt,
                                       const double* __restrict__ rdw, const double* __restrict__ rtl, int ne,
                                       int e0) {
    int const a = e0 < ne ? e0 : ne - 1;
    int const len = (e0 < ne) ? ((ne - e0) < RUN ? ne - e0 : RUN) : 1;
    w[0] = omega[a];
    dw = 0.0;
    tau_lim = 1e300;
    if constexpr (RUN > 1) {
      if (e0 < ne) {
        dw = rdw[e0 / RUN];
        tau_lim = rtl[e0 / RUN];
      }
    }
#pragma unroll
    for (int r = 1; r < RUN; ++r) w[r] = r < len ? omega[a + r] : __builtin_fma((double)r, dw, w[0]);
#pragma unroll
    for (int r = 0; r < RUN; ++r) {
      cw[r] = (T)cwt[r < len ? a + r : a + len - 1];
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        acc[r][q] = 0.0;
        sub[r][q] = (T)0;
      }
    }
  }

  // One trajectory point: tau (FP64), the step angle psi = RN(dw tau) and its phasor
  // (wc, ws) = exp(i psi) (spec_step; unused for RUN 1), and the amplitudes A, G.
  // XY_NEAR false (mixed form, tiles within kSpecNear): the x, y near-field terms are
  // left out; SECOND false (mixed form, tiles within tau_lim / 10): the correction
  // keeps its first order only.
  template <bool XY_NEAR = true, bool SECOND = true>
  __device__ __forceinline__ void add(double tau, double psi, T wc, T ws, T a0, T a1, T a2, T g0, T g1, T g2) {
    static_assert((XY_NEAR && SECOND) || MIXED, "the FP64 form keeps every term");
    double const h0 = w[0] * tau;
    T pc, ps;
    if constexpr (MIXED) {
      float s, c;
      sincos_reduced_f32(reduce_two_pi(h0), &s, &c);
      ps = s;
      pc = c;
    } else {
      sincos_fp64(h0, &ps, &pc);
    }
#pragma unroll
    for (int r = 0; r < RUN; ++r) {
      T uc = pc, us = ps;
      if (r > 0) {
        // exp(i theta_r) by its series (the caller checks |tau| <= tau_lim)
        double const th = __builtin_fma(-(double)r, psi, w[r] * tau - h0);
        if constexpr (MIXED && SECOND) {
          float const f = (float)th;
          float const ct = __builtin_fmaf(f * f, -0.5f, 1.0f);
          uc = __builtin_fmaf(-ps, f, pc * ct);
          us = __builtin_fmaf(pc, f, ps * ct);
        } else if constexpr (MIXED) {
          float const f = (float)th;
          uc = __builtin_fmaf(-ps, f, pc);
          us = __builtin_fmaf(pc, f, ps);
        } else {
          double const t2 = th * th;
          double const ct = __builtin_fma(t2, __builtin_fma(t2, 1.0 / 24.0, -0.5), 1.0);
          double const st = th * __builtin_fma(t2, -1.0 / 6.0, 1.0);
          uc = __builtin_fma(-ps, st, pc * ct);
          us = __builtin_fma(pc, st, ps * ct);
        }
      }
      if constexpr (MIXED && !XY_NEAR) {
        float* const x = sub[r];
        float const gz = g2 * cw[r];
        x[0] = __builtin_fmaf(a0, uc, x[0]);
        x[1] = __builtin_fmaf(a0, us, x[1]);
        x[2] = __builtin_fmaf(a1, uc, x[2]);
        x[3] = __builtin_fmaf(a1, us, x[3]);
        x[4] = __builtin_fmaf(a2, uc, __builtin_fmaf(gz, us, x[4]));
        x[5] = __builtin_fmaf(a2, us, __builtin_fmaf(-gz, uc, x[5]));
      } else if constexpr (MIXED) {
        T const cs = cw[r] * us, cc = cw[r] * uc;
        float* const x = sub[r];
        x[0] = __builtin_fmaf(a0, uc, __builtin_fmaf(g0, cs, x[0]));
        x[1] = __builtin_fmaf(a0, us, __builtin_fmaf(-g0, cc, x[1]));
        x[2] = __builtin_fmaf(a1, uc, __builtin_fmaf(g1, cs, x[2]));
        x[3] = __builtin_fmaf(a1, us, __builtin_fmaf(-g1, cc, x[3]));
        x[4] = __builtin_fmaf(a2, uc, __builtin_fmaf(g2, cs, x[4]));
        x[5] = __builtin_fmaf(a2, us, __builtin_fmaf(-g2, cc, x[5]));
      } else {
        T const cs = cw[r] * us, cc = cw[r] * uc;
        double* const y = acc[r];
        y[0] = __builtin_fma(a0, uc, __builtin_fma(g0, cs, y[0]));
        y[1] = __builtin_fma(a0, us, __builtin_fma(-g0, cc, y[1]));
        y[2] = __builtin_fma(a1, uc, __builtin_fma(g1, cs, y[2]));
        y[3] = __builtin_fma(a1, us, __builtin_fma(-g1, cc, y[3]));
        y[4] = __builtin_fma(a2, uc, __builtin_fma(g2, cs, y[4]));
        y[5] = __builtin_fma(a2, us, __builtin_fma(-g2, cc, y[5]));
      }
      if (r + 1 < RUN) {
        T npc, nps;
        if constexpr (MIXED) {
          npc = __builtin_fmaf(pc, wc, -(ps * ws));
          nps = __builtin_fmaf(ps, wc, pc * ws);
        } else {
          npc = __builtin_fma(pc, wc, -(ps * ws));
          nps = __builtin_fma(ps, wc, pc * ws);
        }
        pc = npc;
        ps = nps;
      }
    }
  }

  // The largest c/omega of this lane's run (its lowest energy; entries past ne repeat
  // the last one).
  __device__ __forceinline__ double cw_max() const {
    T m = cw[0];
#pragma unroll
    for (int r = 1; r < RUN; ++r) m = m > cw[r] ? m : cw[r];
    return (double)m;
  }

  // Mixed: move the FP32 sub-sums into the FP64 accumulators.
  __device__ __forceinline__ void flush() {
    if constexpr (MIXED) {
#pragma unroll
      for (int r = 0; r < RUN; ++r)
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          acc[r][q] += (double)sub[r][q];
          sub[r][q] = 0.f;
        }
    }
  }
};

}  // namespace srs
