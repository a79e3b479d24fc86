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

// Gaussian (kind 1) or uniform box (kind 2) at (X, Y, Z), added to B:
// TField3D_Gaussian::GetF (.cc:88-113), TField3D_UniformBox::GetF (.cc:94-117), the
// position rotated about X, then Y, then Z as TVector3D::RotateSelfX/Y/Z.
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

// Field of one element at (X, Y, Z), added to B.  poles = this particle's pole
// deltas (used when f.pole_off >= 0).
__device__ __forceinline__ void field_add(DevUndulator const& f, const double* __restrict__ poles, double X, double Y,
                                          double Z, double* B) {
  if (f.kind != 0) {
    field_add_local(f, X, Y, Z, B);
    return;
  }
  double const D = (X - f.center[0]) * f.u[0] + (Y - f.center[1]) * f.u[1] + (Z - f.center[2]) * f.u[2];
  double const T = 1 + D * f.taper;
  bool const inside = !(D > f.hi0 || D < f.lo0);
  double s, c;
  sincos_fp64(div_rn(2.0 * kPi * (D - f.ps), f.lam, f.rlam), &s, &c);
  bool const term = D < f.lo1 || D > f.hi1;
  double const fct = (D < f.lo2 || D > f.hi2) ? 0.25 : 0.75;
  double g = T;
  if (f.pole_off >= 0) {  // uniform across the wave
    int p = (int)__builtin_floor(div_rn(D - f.lo1, f.lam / 2., 2.0 * f.rlam));
    p = p < 0 ? 0 : (p > 2 * f.nperiods - 1 ? 2 * f.nperiods - 1 : p);
    g = 1.0 + poles[f.pole_off + p];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double const b = term ? ((f.field[k] * fct) * s) * T : (f.field[k] * s) * g;
    B[k] += inside ? b : 0.0;
  }
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

__device__ __forceinline__ void field_add_jet(DevUndulator const& f, const double* __restrict__ poles, double X,
                                              double Y, double Z, FieldJet& J) {
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
    g = 1.0 + poles[f.pole_off + p];
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

// Undulator records: NU > 0 held in registers, NU == 0 read from memory (any count).
template <int NU>
struct Fields {
  DevUndulator u[NU];
  const double* poles;
  __device__ __forceinline__ Fields(const DevUndulator* U, int, const double* p) : poles(p) {
#pragma unroll
    for (int k = 0; k < NU; ++k) u[k] = U[k];
  }
  __device__ __forceinline__ DevUndulator const& at(int k) const { return u[k]; }
  __device__ __forceinline__ int count() const { return NU; }
  __device__ __forceinline__ void eval(double X, double Y, double Z, double* B) const {
    B[0] = B[1] = B[2] = 0.0;
#pragma unroll
    for (int k = 0; k < NU; ++k) field_add(u[k], poles, X, Y, Z, B);
  }
  __device__ __forceinline__ void jet(double X, double Y, double Z, FieldJet& J) const {
    J = FieldJet{{0, 0, 0}, {0, 0, 0}, 0.0, -1};
#pragma unroll
    for (int k = 0; k < NU; ++k) field_add_jet(u[k], poles, X, Y, Z, J);
  }
};
template <>
struct Fields<0> {
  const DevUndulator* U;
  int nu;
  const double* poles;
  __device__ __forceinline__ Fields(const DevUndulator* U_, int n, const double* p) : U(U_), nu(n), poles(p) {}
  __device__ __forceinline__ DevUndulator const& at(int k) const { return U[k]; }
  __device__ __forceinline__ int count() const { return nu; }
  __device__ __forceinline__ void eval(double X, double Y, double Z, double* B) const {
    B[0] = B[1] = B[2] = 0.0;
    for (int k = 0; k < nu; ++k) field_add(U[k], poles, X, Y, Z, B);
  }
  __device__ __forceinline__ void jet(double X, double Y, double Z, FieldJet& J) const {
    J = FieldJet{{0, 0, 0}, {0, 0, 0}, 0.0, -1};
    for (int k = 0; k < nu; ++k) field_add_jet(U[k], poles, X, Y, Z, J);
  }
};

template <int NU>
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
  __device__ __forceinline__ void derivs(const double* x, double* d) const {
    double B[3];
    F.eval(x[0], x[2], x[4], B);
    derivs_from(x, B, d);
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

__device__ __forceinline__ void store_point(BatchTraj const& bt, int64_t k, const double* x, const double* d) {
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
}

}  // namespace srs
