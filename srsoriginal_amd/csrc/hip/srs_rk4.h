// Device magnetic field and RK4 of the trajectory (reference:
// TField3D_IdealUndulator.cc:106-139, TFieldContainer.cc:107-117,
// OSCARSSR.cc:741-874; same operation order, device sine), shared by the batched
// forward pipeline and the adjoint.
//
// One lane integrates one particle; the RK4 chain is strictly sequential (20k steps
// x 4 field evaluations), so its cost is the latency of that chain.  It is kept
// short: the undulator records are loaded into registers once (NU <= 2 at compile
// time), every division of the reference (2 pi D / lam, h / 6, v / c) is a
// correctly rounded Markstein quotient from a reciprocal formed once (bitwise the
// division), and the regions of the field are selected without branches.
#pragma once
#include "srs_batch.h"
#include "srs_device.h"

namespace srs {

constexpr double kInvSix = 0x1.5555555555555p-3;  // RN(1/6)

struct GridV {
  double v[3];
};

// Field map at the rotated, translated position r, TField3D_Grid::GetF (.cc:124-250):
// zero outside the open interval of every axis with more than one point; linear steps
// a + (d (b - a)) / step, x first, then y, then z.  Not inlined: the field maps are the
// rare element, and inlining their code into the RK4 step would cost every step
// registers.
__device__ __noinline__ GridV grid_eval(const double* __restrict__ gd, int n0, int n1, int n2, double st0, double st1,
                                        double st2, double sp0, double sp1, double sp2, double so0, double so1,
                                        double so2, double r0, double r1, double r2) {
  GridV out{{0, 0, 0}};
  int const n[3] = {n0, n1, n2};
  double const start[3] = {st0, st1, st2}, step[3] = {sp0, sp1, sp2}, stop[3] = {so0, so1, so2};
  double const r[3] = {r0, r1, r2};
  int64_t i[3] = {0, 0, 0};
  double d[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (n[k] <= 1) continue;
    if (!(r[k] > start[k] && r[k] < stop[k])) return out;  // (also for a NaN position)
    int64_t ik = (int64_t)((r[k] - start[k]) / step[k]);
    ik = ik > n[k] - 2 ? n[k] - 2 : ik;  // the upper edge within rounding: last cell
    i[k] = ik;
    d[k] = (r[k] - start[k]) - (double)ik * step[k];
  }
  bool const hx = n0 > 1, hy = n1 > 1, hz = n2 > 1;
  // the data index of (ix, iy, iz) is (ix n1 + iy) n2 + iz; axes with one point index 0
  int64_t const s1 = n2, s0 = (int64_t)n1 * n2;
  auto at = [&](int64_t ix, int64_t iy, int64_t iz, int c) { return gd[3 * (ix * s0 + iy * s1 + iz) + c]; };
  auto lerp = [](double a, double b, double dd, double st) { return a + (dd * (b - a)) / st; };
  int64_t const x0 = i[0], y0 = i[1], z0 = i[2];
  for (int c = 0; c < 3; ++c) {
    double v;
    if (hx && hy && hz) {
      double const v00 = lerp(at(x0, y0, z0, c), at(x0 + 1, y0, z0, c), d[0], step[0]);
      double const v10 = lerp(at(x0, y0 + 1, z0, c), at(x0 + 1, y0 + 1, z0, c), d[0], step[0]);
      double const v01 = lerp(at(x0, y0, z0 + 1, c), at(x0 + 1, y0, z0 + 1, c), d[0], step[0]);
      double const v11 = lerp(at(x0, y0 + 1, z0 + 1, c), at(x0 + 1, y0 + 1, z0 + 1, c), d[0], step[0]);
      v = lerp(lerp(v00, v10, d[1], step[1]), lerp(v01, v11, d[1], step[1]), d[2], step[2]);
    } else if (hx && hy) {
      v = lerp(lerp(at(x0, y0, 0, c), at(x0 + 1, y0, 0, c), d[0], step[0]),
               lerp(at(x0, y0 + 1, 0, c), at(x0 + 1, y0 + 1, 0, c), d[0], step[0]), d[1], step[1]);
    } else if (hx && hz) {
      v = lerp(lerp(at(x0, 0, z0, c), at(x0 + 1, 0, z0, c), d[0], step[0]),
               lerp(at(x0, 0, z0 + 1, c), at(x0 + 1, 0, z0 + 1, c), d[0], step[0]), d[2], step[2]);
    } else if (hy && hz) {
      v = lerp(lerp(at(0, y0, z0, c), at(0, y0 + 1, z0, c), d[1], step[1]),
               lerp(at(0, y0, z0 + 1, c), at(0, y0 + 1, z0 + 1, c), d[1], step[1]), d[2], step[2]);
    } else if (hx) {
      v = lerp(at(x0, 0, 0, c), at(x0 + 1, 0, 0, c), d[0], step[0]);
    } else if (hy) {
      v = lerp(at(0, y0, 0, c), at(0, y0 + 1, 0, c), d[1], step[1]);
    } else {
      v = lerp(at(0, 0, z0, c), at(0, 0, z0 + 1, c), d[2], step[2]);
    }
    out.v[c] = v;
  }
  return out;
}

// Gaussian (kind 1), uniform box (kind 2) or field map (kind 3) at (X, Y, Z), added to
// B: TField3D_Gaussian::GetF (.cc:88-113), TField3D_UniformBox::GetF (.cc:94-117),
// TField3D_Grid::GetF (.cc:124-250), the position rotated about X, then Y, then Z as
// TVector3D::RotateSelfX/Y/Z.
__device__ __forceinline__ void field_add_local(DevUndulator const& f, double X, double Y, double Z, double* B) {
  double x = X, y = Y, z = Z;
  {
    double const yy = y;
    y = f.rc[0] * yy - f.rs[0] * z;
    z = f.rs[0] * yy + f.rc[0] * z;
  }
  {
    double const zz = z;
    z = f.rc[1] * zz - f.rs[1] * x;
    x = f.rs[1] * zz + f.rc[1] * x;
  }
  {
    double const xx = x;
    x = f.rc[2] * xx - f.rs[2] * y;
    y = f.rs[2] * xx + f.rc[2] * y;
  }
  double const r[3] = {x - f.center[0], y - f.center[1], z - f.center[2]};
  if (f.kind == 3) {
    GridV const g = grid_eval(f.gdata, f.gn[0], f.gn[1], f.gn[2], f.gstart[0], f.gstart[1], f.gstart[2], f.gstep[0],
                              f.gstep[1], f.gstep[2], f.gstop[0], f.gstop[1], f.gstop[2], r[0], r[1], r[2]);
    B[0] += g.v[0];
    B[1] += g.v[1];
    B[2] += g.v[2];
    return;
  }
  if (f.kind == 1) {
    double frac = 1;
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (f.w[k] > 0) {
        double const t = r[k] / f.w[k];
        frac *= exp(-(t * t) / 2.);
      }
#pragma unroll
    for (int k = 0; k < 3; ++k) B[k] += frac * f.field[k];
  } else {
    bool out = false;
#pragma unroll
    for (int k = 0; k < 3; ++k) out = out || (f.w[k] > 0 && __builtin_fabs(r[k]) > f.w[k]);
#pragma unroll
    for (int k = 0; k < 3; ++k) B[k] += out ? 0.0 : f.field[k];
  }
}

// Field of one undulator at (X, Y, Z), added to B.  poles = this particle's pole
// deltas (used when f.pole_off >= 0), entry k at poles[k * ps] (ps = kPoleLane when the
// kernel staged them in LDS, a column per lane; 1 in global memory).
constexpr int kPoleLane = 64;
__device__ __forceinline__ void field_add_und(DevUndulator const& f, const double* __restrict__ poles, int ps,
                                              double X, double Y, double Z, double* B) {
  double const D = (X - f.center[0]) * f.u[0] + (Y - f.center[1]) * f.u[1] + (Z - f.center[2]) * f.u[2];
  double const T = 1 + D * f.taper;
  bool const inside = !(D > f.hi0 || D < f.lo0);
  double s, c;
  sincos_fp64_short(div_rn(2.0 * kPi * (D - f.ps), f.lam, f.rlam), &s, &c);
  bool const term = D < f.lo1 || D > f.hi1;
  double const fct = (D < f.lo2 || D > f.hi2) ? 0.25 : 0.75;
  double g = T;
  if (f.pole_off >= 0) {  // uniform across the wave
    int p = (int)__builtin_floor(div_rn(D - f.lo1, f.lam / 2., 2.0 * f.rlam));
    p = p < 0 ? 0 : (p > 2 * f.nperiods - 1 ? 2 * f.nperiods - 1 : p);
    g = 1.0 + poles[(f.pole_off + p) * ps];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double const b = term ? ((f.field[k] * fct) * s) * T : (f.field[k] * s) * g;
    B[k] += inside ? b : 0.0;
  }
}

// Field of any element at (X, Y, Z), added to B.
__device__ __forceinline__ void field_add(DevUndulator const& f, const double* __restrict__ poles, int ps, double X,
                                          double Y, double Z, double* B) {
  if (f.kind != 0)
    field_add_local(f, X, Y, Z, B);
  else
    field_add_und(f, poles, ps, X, Y, Z, B);
}

// Field, its derivative along the period direction and its derivative with
// respect to the scale of the pole the point lies in, for the adjoint.
//   dB/dD = field * fct * (cos(theta) * 2 pi / lam * T + sin(theta) * taper)    (terminations)
//         = field * (cos(theta) * 2 pi / lam * g + sin(theta) * taper)          (interior)
// with g = T, or 1 + delta_p for the pole-scaled form (taper 0 there);
// dB/d delta_p = field * sin(theta) in pole p's interior half period.
struct FieldJet {
  double B[3];
  double dBdD[3];  // summed over undulators (all share one period direction)
  double sp;       // sin(theta) at a point inside pole p of the scaled undulator, else 0
  int p;           // that pole, or -1
};

__device__ __forceinline__ void field_add_jet(DevUndulator const& f, const double* __restrict__ poles, int ps,
                                              double X, double Y, double Z, FieldJet& J) {
  double const D = (X - f.center[0]) * f.u[0] + (Y - f.center[1]) * f.u[1] + (Z - f.center[2]) * f.u[2];
  double const T = 1 + D * f.taper;
  bool const inside = !(D > f.hi0 || D < f.lo0);
  double s, c;
  sincos_fp64(div_rn(2.0 * kPi * (D - f.ps), f.lam, f.rlam), &s, &c);
  double const k = 2.0 * kPi * f.rlam;
  bool const term = D < f.lo1 || D > f.hi1;
  double const fct = (D < f.lo2 || D > f.hi2) ? 0.25 : 0.75;
  double g = T;
  int p = -1;
  if (f.pole_off >= 0) {
    p = (int)__builtin_floor(div_rn(D - f.lo1, f.lam / 2., 2.0 * f.rlam));
    p = p < 0 ? 0 : (p > 2 * f.nperiods - 1 ? 2 * f.nperiods - 1 : p);
    g = 1.0 + poles[(f.pole_off + p) * ps];
  }
  double const dt = term ? fct * (c * k * T + s * f.taper) : c * k * g + s * f.taper;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    double const b = term ? ((f.field[q] * fct) * s) * T : (f.field[q] * s) * g;
    J.B[q] += inside ? b : 0.0;
    J.dBdD[q] += inside ? f.field[q] * dt : 0.0;
  }
  if (f.pole_off >= 0 && inside && !term) {
    J.sp = s;
    J.p = p;
  }
}

// Field records: NU > 0 undulators held in registers (lists of undulators only: the
// step then carries no code of the other kinds), NU == 0 any elements read from memory.
template <int NU>
struct Fields {
  DevUndulator u[NU];
  const double* poles;
  int ps;  // stride of the pole deltas (field_add_und)
  __device__ __forceinline__ Fields(const DevUndulator* U, int, const double* p, int stride = 1)
      : poles(p), ps(stride) {
#pragma unroll
    for (int k = 0; k < NU; ++k) u[k] = U[k];
  }
  __device__ __forceinline__ DevUndulator const& at(int k) const { return u[k]; }
  __device__ __forceinline__ int count() const { return NU; }
  __device__ __forceinline__ void eval(double X, double Y, double Z, double* B) const {
    B[0] = B[1] = B[2] = 0.0;
#pragma unroll
    for (int k = 0; k < NU; ++k) field_add_und(u[k], poles, ps, X, Y, Z, B);
  }
  __device__ __forceinline__ void jet(double X, double Y, double Z, FieldJet& J) const {
    J = FieldJet{{0, 0, 0}, {0, 0, 0}, 0.0, -1};
#pragma unroll
    for (int k = 0; k < NU; ++k) field_add_jet(u[k], poles, ps, X, Y, Z, J);
  }
};
template <>
struct Fields<0> {
  const DevUndulator* U;
  int nu;
  const double* poles;
  int ps;
  __device__ __forceinline__ Fields(const DevUndulator* U_, int n, const double* p, int stride = 1)
      : U(U_), nu(n), poles(p), ps(stride) {}
  __device__ __forceinline__ DevUndulator const& at(int k) const { return U[k]; }
  __device__ __forceinline__ int count() const { return nu; }
  __device__ __forceinline__ void eval(double X, double Y, double Z, double* B) const {
    B[0] = B[1] = B[2] = 0.0;
    for (int k = 0; k < nu; ++k) field_add(U[k], poles, ps, X, Y, Z, B);
  }
  // magnetic and electric sums, each element into the field of its target
  __device__ __forceinline__ void eval_eb(double X, double Y, double Z, double* B, double* E) const {
    B[0] = B[1] = B[2] = 0.0;
    E[0] = E[1] = E[2] = 0.0;
    for (int k = 0; k < nu; ++k) field_add(U[k], poles, ps, X, Y, Z, U[k].target ? E : B);
  }
  __device__ __forceinline__ void jet(double X, double Y, double Z, FieldJet& J) const {
    J = FieldJet{{0, 0, 0}, {0, 0, 0}, 0.0, -1};
    for (int k = 0; k < nu; ++k) field_add_jet(U[k], poles, ps, X, Y, Z, J);
  }
};

// The pole deltas of one lane's particle, npl of them from global memory at g, copied
// into the lane's column of the block's LDS table (kPoleLane lanes a row) when npl > 0:
// the field then reads them at LDS latency, not at the global memory latency that
// every RK4 stage would otherwise wait for.  Returns the table to read and its stride.
__device__ __forceinline__ const double* stage_poles(const double* g, double* lds, int npl, int* ps) {
  if (npl <= 0) {
    *ps = 1;
    return g;
  }
  int const lane = threadIdx.x % kPoleLane;
  for (int k = 0; k < npl; ++k) lds[k * kPoleLane + lane] = g[k];
  *ps = kPoleLane;
  return lds + lane;
}

// Derivative form, as OSCARSSR::SetDerivativesFunction picks it (OSCARSSR.cc:690-703):
// magnetic elements only (DerivativesB), electric only (DerivativesE) or both
// (DerivativesEB); the electric forms take the general records (NU == 0).
enum class Force { B, E, EB };

template <int NU, Force FORCE = Force::B>
struct RK4Ctx {
  Fields<NU> F;
  double qmg;
  // state x = [x, vx, y, vy, z, vz]; d = dx/dt with dv/dt = (q/m gamma) v x B
  // (OSCARSSR::DerivativesB, OSCARSSR.cc:741-765)
  __device__ __forceinline__ void derivs_from(const double* x, const double* B, double* d) const {
    d[0] = x[1];
    d[1] = qmg * (-x[5] * B[1] + x[3] * B[2]);
    d[2] = x[3];
    d[3] = qmg * (x[5] * B[0] - x[1] * B[2]);
    d[4] = x[5];
    d[5] = qmg * (x[1] * B[1] - x[3] * B[0]);
  }
  // dv/dt = (q/m gamma)(E + v x B), OSCARSSR::DerivativesEB (OSCARSSR.cc:769-794), and
  // (q/m gamma) E, DerivativesE (:713-737)
  __device__ __forceinline__ void derivs_from_eb(const double* x, const double* B, const double* E,
                                                 double* d) const {
    d[0] = x[1];
    d[1] = qmg * (E[0] - x[5] * B[1] + x[3] * B[2]);
    d[2] = x[3];
    d[3] = qmg * (E[1] + x[5] * B[0] - x[1] * B[2]);
    d[4] = x[5];
    d[5] = qmg * (E[2] + x[1] * B[1] - x[3] * B[0]);
  }
  __device__ __forceinline__ void derivs_from_e(const double* x, const double* E, double* d) const {
    d[0] = x[1];
    d[1] = qmg * E[0];
    d[2] = x[3];
    d[3] = qmg * E[1];
    d[4] = x[5];
    d[5] = qmg * E[2];
  }
  __device__ __forceinline__ void derivs(const double* x, double* d) const {
    double B[3];
    if constexpr (FORCE == Force::B) {
      F.eval(x[0], x[2], x[4], B);
      derivs_from(x, B, d);
    } else {
      double E[3];
      F.eval_eb(x[0], x[2], x[4], B, E);
      if constexpr (FORCE == Force::E) derivs_from_e(x, E, d);
      else derivs_from_eb(x, B, E, d);
    }
  }
  // OSCARSSR::RK4 (OSCARSSR.cc:826-874); hh = h/2 and h6 = h/6 are formed by the caller
  __device__ __forceinline__ void step(double* y, const double* dydx, double h, double hh, double h6) const {
    double dym[6], dyt[6], yt[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) yt[i] = y[i] + hh * dydx[i];
    derivs(yt, dyt);
#pragma unroll
    for (int i = 0; i < 6; ++i) yt[i] = y[i] + hh * dyt[i];
    derivs(yt, dym);
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      yt[i] = y[i] + h * dym[i];
      dym[i] += dyt[i];
    }
    derivs(yt, dyt);
#pragma unroll
    for (int i = 0; i < 6; ++i) y[i] = y[i] + h6 * (dydx[i] + dyt[i] + 2.0 * dym[i]);
  }
};

// Stores point k; returns whether it does not travel along z (transverse_beam).
__device__ __forceinline__ bool store_point(BatchTraj const& bt, int64_t k, const double* x, const double* d) {
  bt.x[k] = x[0];
  bt.y[k] = x[2];
  bt.z[k] = x[4];
  double const bx = div_rn(x[1], kC, kInvC), by = div_rn(x[3], kC, kInvC), bz = div_rn(x[5], kC, kInvC);
  bt.bx[k] = bx;
  bt.by[k] = by;
  bt.bz[k] = bz;
  bt.fbx[k] = (float)bx;
  bt.fby[k] = (float)by;
  bt.fomb[k] = (float)(1.0 - bz);
  if (bt.ax) {
    bt.ax[k] = div_rn(d[1], kC, kInvC);
    bt.ay[k] = div_rn(d[3], kC, kInvC);
    bt.az[k] = div_rn(d[5], kC, kInvC);
  }
  return transverse_beam(bx, by);
}

}  // namespace srs
