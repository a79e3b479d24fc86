// CPU oracle: clean-room FP64 restatement of the reference OSCARS SR CPU path.
// TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline); the product never
// links it.  Compile with the reference's flags (oracle/Makefile) so that the
// arithmetic is bit-identical to the reference engine.
//
// Pinning (tests/test_oracle_golden.py, tests/test_oracle_closed_form.py):
//   constants ............ closed form (CODATA values of TOSCARSSR.h:20-69)
//   oracle_bfield ........ closed form sine + 0.25/0.75 terminations
//   oracle_trajectory .... bitwise vs reference digests (C0 20001 pts, EPU 40001 pts)
//                          + uniform-field circle radius
//   oracle_energy_grid ... bitwise vs reference spectra energies (shifted grid)
//   oracle_rect_surface .. bitwise vs reference container coordinates
//   oracle_spectrum ...... bitwise vs reference (C0 full 2000 E, small, off-axis list)
//                          + first-harmonic resonance formula
//   oracle_flux .......... bitwise vs reference (C1 full 201x201, routes, point list)
//                          + mirror symmetry
//   oracle_power_density . bitwise vs reference (C2 full 301x301, directional tilt)
//                          + integral vs total power
//   oracle_total_power ... bitwise vs reference (C0, C2)
//   oracle_flux1 ......... bitwise vs the reference's CalculateFlux1 on the trajectory it
//                          used; converges to the integration-by-parts identity (integral
//                          + endpoint terms) as the step shrinks
//   random particles ..... bitwise vs reference (seeded mt19937/normal stream)
//   oracle_pole_loss_gradient  dual-number forward mode; vs the reference's own central
//                          differences (h = 1e-6) of the pole-error flux loss
//   oracle_pole_pd_loss_gradient  the same for the power-density loss; vs the
//                          reference's power-density map and its central differences

#include "oracle.h"

#include <cmath>
#include <complex>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

namespace {

// Physical constants, TOSCARSSR.h:20-69 / TSRS.h:20-69 (dated CODATA values).
const double kPi = 3.14159265358979323846;
const double kTwoPi = 2.0 * kPi;
const double kFourPi = 4.0 * kPi;
const double kPi2 = kPi * kPi;
const double kSqrt2Pi = 2.5066282746310002416123552393401;
const double kC = 2.99792458e8;
const double kH = 6.62606876e-34;
const double kHbar = 1.054571596e-34;
const double kQe = 1.602176462e-19;
const double kMe = 9.10938356e-31;
const double kEps0 = 8.854187817E-12;
const double kMu0 = 1.2566370614E-6;

inline double ev_to_omega(double e) { return e * kQe / kHbar; }             // TOSCARSSR.h:66
inline double kg_to_gev(double m) { return 1e-9 * m * kC * kC / kQe; }       // TOSCARSSR.h:68

// 3-vector helpers in the reference's operation order (TVector3D.cc:98-207).
struct V3 {
  double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double mag2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 divs(V3 a, double d) { return V3{a.x / d, a.y / d, a.z / d}; }
inline V3 muls(V3 a, double s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline V3 unit(V3 a) { return divs(a, std::sqrt(mag2(a))); }
inline V3 cross(V3 a, V3 b) { return V3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
// TVector3D::Orthogonal, TVector3D.h:120-131
inline V3 orthogonal(V3 a) {
  double xx = a.x < 0.0 ? -a.x : a.x;
  double yy = a.y < 0.0 ? -a.y : a.y;
  double zz = a.z < 0.0 ? -a.z : a.z;
  if (xx < yy) return xx < zz ? V3{0, a.z, -a.y} : V3{a.y, -a.x, 0};
  return yy < zz ? V3{-a.z, 0, a.x} : V3{a.y, -a.x, 0};
}
inline void rot_x(V3& v, double a) {  // TVector3D::RotateSelfX
  double s = std::sin(a), c = std::cos(a), yy = v.y;
  v.y = c * yy - s * v.z;
  v.z = s * yy + c * v.z;
}
inline void rot_y(V3& v, double a) {  // TVector3D::RotateSelfY
  double s = std::sin(a), c = std::cos(a), zz = v.z;
  v.z = c * zz - s * v.x;
  v.x = s * zz + c * v.x;
}
inline void rot_z(V3& v, double a) {  // TVector3D::RotateSelfZ
  double s = std::sin(a), c = std::cos(a), xx = v.x;
  v.x = c * xx - s * v.y;
  v.y = s * xx + c * v.y;
}
inline void rot_xyz(V3& v, V3 r) {
  rot_x(v, r.x);
  rot_y(v, r.y);
  rot_z(v, r.z);
}
inline V3 v3(const double* p) { return V3{p[0], p[1], p[2]}; }

// Ideal undulator field, TField3D_IdealUndulator.cc:50-77 (Init) and :106-139 (GetF).
// With pole_scales != NULL the 2N interior half-period poles are scaled by
// (1 + delta_p), p = floor((D - (-L/2 + lambda)) / (lambda/2)) (phase and taper 0).
// TField3D_Gaussian::GetF (TField3D_Gaussian.cc:88-113) and TField3D_UniformBox::GetF
// (TField3D_UniformBox.cc:94-117): the position rotated by TVector3D::RotateSelfXYZ
// (sin and cos per call), the center subtracted, then the Gaussian envelope or the box.
V3 local_field(const oracle_undulator& u, V3 X) {
  double const* r = u.rotations;
  {
    double const s = std::sin(r[0]), c = std::cos(r[0]), yy = X.y;
    X.y = c * yy - s * X.z;
    X.z = s * yy + c * X.z;
  }
  {
    double const s = std::sin(r[1]), c = std::cos(r[1]), zz = X.z;
    X.z = c * zz - s * X.x;
    X.x = s * zz + c * X.x;
  }
  {
    double const s = std::sin(r[2]), c = std::cos(r[2]), xx = X.x;
    X.x = c * xx - s * X.y;
    X.y = s * xx + c * X.y;
  }
  V3 const RX = sub(X, v3(u.center));
  V3 const F = v3(u.field);
  double const w[3] = {u.width[0], u.width[1], u.width[2]};
  double const d[3] = {RX.x, RX.y, RX.z};
  if (u.kind == 1) {
    double frac = 1;
    for (int k = 0; k < 3; ++k)
      if (w[k] > 0) frac *= std::exp(-std::pow(d[k] / w[k], 2) / 2.);
    return V3{frac * F.x, frac * F.y, frac * F.z};
  }
  for (int k = 0; k < 3; ++k)
    if (w[k] > 0 && std::fabs(d[k]) > std::fabs(w[k] / 2.)) return V3{0, 0, 0};
  return F;
}

V3 undulator_field(const oracle_undulator& u, const double* pole_scales, V3 X) {
  if (u.kind != 0) return local_field(u, X);
  V3 F = v3(u.field), P = v3(u.period), C = v3(u.center);
  double const lam = std::sqrt(mag2(P));
  V3 const pu = divs(P, std::sqrt(mag2(P)));
  double const len = std::sqrt(mag2(P)) * (u.nperiods + 2);
  if (pole_scales) {
    double const D = dot(X, pu);
    double const s = std::sin(2.0 * kPi * D / lam);
    if (D > len / 2. || D < -len / 2.) return V3{0, 0, 0};
    if (D < -len / 2. + lam || D > len / 2. - lam) {
      if (D < -len / 2. + lam / 2. || D > len / 2. - lam / 2.) return V3{(0.25 * F.x) * s, (0.25 * F.y) * s, (0.25 * F.z) * s};
      return V3{(0.75 * F.x) * s, (0.75 * F.y) * s, (0.75 * F.z) * s};
    }
    int p = (int)std::floor((D - (-len / 2. + lam)) / (lam / 2.));
    if (p < 0) p = 0;
    if (p > 2 * u.nperiods - 1) p = 2 * u.nperiods - 1;
    double const sc = 1.0 + pole_scales[p];
    return V3{(F.x * s) * sc, (F.y * s) * sc, (F.z * s) * sc};
  }
  double const D = dot(sub(X, C), pu);
  double const T = 1 + D * u.taper;
  double const PS = u.phase * std::sqrt(mag2(P)) / kTwoPi;
  if (D > len / 2. + PS || D < -len / 2. + PS) return V3{0, 0, 0};
  double const s = std::sin(kTwoPi * (D - PS) / lam);
  if (D < -len / 2. + PS + lam || D > len / 2. + PS - lam) {
    if (D < -len / 2. + PS + lam / 2. || D > len / 2. + PS - lam / 2.)
      return V3{((F.x * 0.25) * s) * T, ((F.y * 0.25) * s) * T, ((F.z * 0.25) * s) * T};
    return V3{((F.x * 0.75) * s) * T, ((F.y * 0.75) * s) * T, ((F.z * 0.75) * s) * T};
  }
  return V3{(F.x * s) * T, (F.y * s) * T, (F.z * s) * T};
}

// TFieldContainer::GetF, TFieldContainer.cc:107-117: sum from (0,0,0).
V3 field_sum(const oracle_fields* f, V3 X) {
  V3 S{0, 0, 0};
  for (int i = 0; i < f->nund; ++i) {
    V3 b = undulator_field(f->und[i], 0, X);
    S.x += b.x;
    S.y += b.y;
    S.z += b.z;
  }
  if (f->pole_und) {
    V3 b = undulator_field(*f->pole_und, f->pole_scales, X);
    S.x += b.x;
    S.y += b.y;
    S.z += b.z;
  }
  return S;
}

// The electric container's sum (TFieldContainer::GetF over the electric elements).
V3 efield_sum(const oracle_fields* f, V3 X) {
  V3 S{0, 0, 0};
  for (int i = 0; i < f->neund; ++i) {
    V3 e = undulator_field(f->eund[i], 0, X);
    S.x += e.x;
    S.y += e.y;
    S.z += e.z;
  }
  return S;
}

// OSCARSSR::SetDerivativesFunction (OSCARSSR.cc:690-703): DerivativesE when only electric
// fields exist, DerivativesB when only magnetic ones, else DerivativesEB.
inline void derivs(const oracle_fields* f, double qmg, const double* x, double* dxdt) {
  bool const has_b = f->nund > 0 || f->pole_und != 0;
  bool const has_e = f->neund > 0;
  dxdt[0] = x[1];
  dxdt[2] = x[3];
  dxdt[4] = x[5];
  if (!has_b && has_e) {  // DerivativesE, :713-737
    V3 E = efield_sum(f, V3{x[0], x[2], x[4]});
    dxdt[1] = qmg * E.x;
    dxdt[3] = qmg * E.y;
    dxdt[5] = qmg * E.z;
  } else if (has_b && !has_e) {  // DerivativesB, :741-765
    V3 B = field_sum(f, V3{x[0], x[2], x[4]});
    dxdt[1] = qmg * (-x[5] * B.y + x[3] * B.z);
    dxdt[3] = qmg * (x[5] * B.x - x[1] * B.z);
    dxdt[5] = qmg * (x[1] * B.y - x[3] * B.x);
  } else {  // DerivativesEB, :769-794
    V3 B = field_sum(f, V3{x[0], x[2], x[4]});
    V3 E = efield_sum(f, V3{x[0], x[2], x[4]});
    dxdt[1] = qmg * (E.x - x[5] * B.y + x[3] * B.z);
    dxdt[3] = qmg * (E.y + x[5] * B.x - x[1] * B.z);
    dxdt[5] = qmg * (E.z + x[1] * B.y - x[3] * B.x);
  }
}

// OSCARSSR::RK4, OSCARSSR.cc:826-874 (in place: y == yout).
inline void rk4(const oracle_fields* f, double qmg, double* y, const double* dydx, double h) {
  double dym[6], dyt[6], yt[6];
  double const hh = h * 0.5;
  double const h6 = h / 6.0;
  for (int i = 0; i < 6; ++i) yt[i] = y[i] + hh * dydx[i];
  derivs(f, qmg, yt, dyt);
  for (int i = 0; i < 6; ++i) yt[i] = y[i] + hh * dyt[i];
  derivs(f, qmg, yt, dym);
  for (int i = 0; i < 6; ++i) {
    yt[i] = y[i] + h * dym[i];
    dym[i] += dyt[i];
  }
  derivs(f, qmg, yt, dyt);
  for (int i = 0; i < 6; ++i) y[i] = y[i] + h6 * (dydx[i] + dyt[i] + 2.0 * dym[i]);
}

// Spectrum/flux integrand for one (observation point, omega), OSCARSSR.cc:1202-1246
// (spectrum), :2054-2092 (flux).  Returns C2 * |C1 dt SumE|^2 * weight.
double flux_point(const double* traj, int64_t n, double dt, V3 obs, double omega, double q, double current, double weight) {
  const double *X = traj, *Y = traj + n, *Z = traj + 2 * n, *BX = traj + 3 * n, *BY = traj + 4 * n, *BZ = traj + 5 * n;
  double const C0 = q / (kFourPi * kC * kEps0 * kSqrt2Pi);
  double const C2 = kFourPi * current / (kH * std::fabs(q) * kMu0 * kC) * 1e-6 * 0.001;
  std::complex<double> const I(0, 1);
  std::complex<double> const One(1, 0);
  std::complex<double> const ICoverOmega = I * kC / omega;
  std::complex<double> const C1(0, C0 * omega);
  std::complex<double> sx(0, 0), sy(0, 0), sz(0, 0);
  for (int iT = 0; iT != n; ++iT) {
    V3 const R = sub(obs, V3{X[iT], Y[iT], Z[iT]});
    V3 const N = unit(R);
    double const D = std::sqrt(mag2(R));
    std::complex<double> const e = std::exp(std::complex<double>(0, omega * (dt * iT + D / kC)));
    std::complex<double> const w = One + (ICoverOmega / D);
    sx += (std::complex<double>(BX[iT], 0) - std::complex<double>(N.x, 0) * w) / D * e;
    sy += (std::complex<double>(BY[iT], 0) - std::complex<double>(N.y, 0) * w) / D * e;
    sz += (std::complex<double>(BZ[iT], 0) - std::complex<double>(N.z, 0) * w) / D * e;
  }
  std::complex<double> const k = C1 * dt;
  sx *= k;
  sy *= k;
  sz *= k;
  std::complex<double> const d = sx * std::conj(sx) + sy * std::conj(sy) + sz * std::conj(sz);
  return C2 * d.real() * weight;
}

// Power-density integrand for one point, OSCARSSR.cc:1522-1573 (and :1726-1771).
double pd_point(const double* traj, int64_t n, double dt, V3 obs, V3 nrm, bool directional, double q, double current) {
  const double *X = traj, *Y = traj + n, *Z = traj + 2 * n, *BX = traj + 3 * n, *BY = traj + 4 * n, *BZ = traj + 5 * n;
  const double *AX = traj + 6 * n, *AY = traj + 7 * n, *AZ = traj + 8 * n;
  double Sum = 0;
  for (int iT = 0; iT != n; ++iT) {
    V3 const Xp{X[iT], Y[iT], Z[iT]};
    V3 const B{BX[iT], BY[iT], BZ[iT]};
    V3 const A{AX[iT], AY[iT], AZ[iT]};
    V3 const N1 = unit(sub(obs, Xp));
    V3 const N2 = unit(orthogonal(N1));
    V3 const N3 = unit(cross(N1, N2));
    double const N1DotNormal = dot(N1, nrm);
    if (directional && N1DotNormal <= 0) continue;
    V3 const Num = cross(N1, cross(sub(N1, B), A));
    double const Den = std::pow(1 - dot(B, N1), 5);
    Sum += std::pow(dot(Num, N2), 2) / Den / mag2(sub(obs, Xp)) * N1DotNormal;
    Sum += std::pow(dot(Num, N3), 2) / Den / mag2(sub(obs, Xp)) * N1DotNormal;
  }
  Sum *= std::fabs(q * current) / (16 * kPi2 * kEps0 * kC) * dt;
  Sum /= 1e6;
  if (!directional) {
    if (Sum < 0) Sum *= -1;
  }
  return Sum;
}

// The reference threading model: one std::thread per output point, NThreads
// concurrent, joined in order (OSCARSSR.cc:1109-1132, :1823-1846, :2533-2556).
template <class F>
void per_point_threads(int64_t npts, int nthreads, F fn) {
  if (nthreads <= 1) {
    for (int64_t i = 0; i < npts; ++i) fn(i);
    return;
  }
  std::vector<std::thread> th;
  int64_t const first = npts > nthreads ? nthreads : npts;
  for (int64_t i = 0; i < first; ++i) th.push_back(std::thread(fn, i));
  for (int64_t i = first; i != npts + first; ++i) {
    int64_t const it = i % first;
    th[it].join();
    if (i < npts) th[it] = std::thread(fn, i);
  }
}

// Process-global generator, TRandomA.cc:12-64.
std::mt19937* gMT = new std::mt19937(5489u);
std::normal_distribution<double> gNormal(0, 1);
std::uniform_real_distribution<double> gUniform(0, 1);

void finish_particle(oracle_particle* p) {
  // TParticleA::SetGamma / SetQoverMGamma, TParticleA.cc:353-382
  double const b2 = mag2(v3(p->beta0));
  if (b2 == 1) return;
  p->gamma = 1.0 / std::sqrt(1.0 - b2);
  if (p->m == 0 || p->gamma == 0) return;
  p->qmg = p->q / p->m / p->gamma;
}


// ---- forward-mode derivatives for the pole-error gradient -------------------------
// Dual numbers (value, derivative along one direction) carried through the same field,
// RK4 and flux arithmetic as above: an FP64 CPU gradient that shares no hand-derived
// derivative formula with the adjoint kernels it checks.
struct Dual {
  double v, d;
};
inline Dual operator+(Dual a, Dual b) { return Dual{a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return Dual{a.v - b.v, a.d - b.d}; }
inline Dual operator*(Dual a, Dual b) { return Dual{a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) { return Dual{a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
inline Dual operator-(Dual a) { return Dual{-a.v, -a.d}; }
inline Dual cst(double x) { return Dual{x, 0.0}; }
inline Dual dsqrt(Dual a) {
  double const r = std::sqrt(a.v);
  return Dual{r, a.d / (2.0 * r)};
}
inline Dual dsin(Dual a) { return Dual{std::sin(a.v), std::cos(a.v) * a.d}; }
inline Dual dcos(Dual a) { return Dual{std::cos(a.v), -std::sin(a.v) * a.d}; }

// undulator_field with dual position and dual pole scales
void undulator_field_dual(const oracle_undulator& u, const Dual* pole_scales, const Dual* X, Dual* B) {
  double const lam = std::sqrt(mag2(v3(u.period)));
  V3 const pu = divs(v3(u.period), lam);
  double const len = lam * (u.nperiods + 2);
  V3 const F = v3(u.field);
  B[0] = B[1] = B[2] = cst(0);
  if (pole_scales) {
    Dual const D = X[0] * cst(pu.x) + X[1] * cst(pu.y) + X[2] * cst(pu.z);
    Dual const s = dsin(cst(2.0 * kPi) * D / cst(lam));
    if (D.v > len / 2. || D.v < -len / 2.) return;
    Dual sc = cst(1.0);
    if (D.v < -len / 2. + lam || D.v > len / 2. - lam) {
      sc = cst((D.v < -len / 2. + lam / 2. || D.v > len / 2. - lam / 2.) ? 0.25 : 0.75);
    } else {
      int p = (int)std::floor((D.v - (-len / 2. + lam)) / (lam / 2.));
      if (p < 0) p = 0;
      if (p > 2 * u.nperiods - 1) p = 2 * u.nperiods - 1;
      sc = cst(1.0) + pole_scales[p];
    }
    B[0] = cst(F.x) * s * sc;
    B[1] = cst(F.y) * s * sc;
    B[2] = cst(F.z) * s * sc;
    return;
  }
  V3 const C = v3(u.center);
  Dual const D = (X[0] - cst(C.x)) * cst(pu.x) + (X[1] - cst(C.y)) * cst(pu.y) + (X[2] - cst(C.z)) * cst(pu.z);
  Dual const T = cst(1) + D * cst(u.taper);
  double const PS = u.phase * lam / kTwoPi;
  if (D.v > len / 2. + PS || D.v < -len / 2. + PS) return;
  Dual const s = dsin(cst(kTwoPi) * (D - cst(PS)) / cst(lam));
  double f = 1.0;
  if (D.v < -len / 2. + PS + lam || D.v > len / 2. + PS - lam)
    f = (D.v < -len / 2. + PS + lam / 2. || D.v > len / 2. + PS - lam / 2.) ? 0.25 : 0.75;
  B[0] = cst(F.x * f) * s * T;
  B[1] = cst(F.y * f) * s * T;
  B[2] = cst(F.z * f) * s * T;
}

void derivs_dual(const oracle_fields* f, const Dual* pole_scales, double qmg, const Dual* x, Dual* d) {
  Dual const X[3] = {x[0], x[2], x[4]};
  Dual B[3] = {cst(0), cst(0), cst(0)}, b[3];
  for (int i = 0; i < f->nund; ++i) {
    undulator_field_dual(f->und[i], 0, X, b);
    for (int k = 0; k < 3; ++k) B[k] = B[k] + b[k];
  }
  if (f->pole_und) {
    undulator_field_dual(*f->pole_und, pole_scales, X, b);
    for (int k = 0; k < 3; ++k) B[k] = B[k] + b[k];
  }
  Dual const q = cst(qmg);
  d[0] = x[1];
  d[1] = q * (-x[5] * B[1] + x[3] * B[2]);
  d[2] = x[3];
  d[3] = q * (x[5] * B[0] - x[1] * B[2]);
  d[4] = x[5];
  d[5] = q * (x[1] * B[1] - x[3] * B[0]);
}

void rk4_dual(const oracle_fields* f, const Dual* ps, double qmg, Dual* y, const Dual* dydx, double h) {
  Dual dym[6], dyt[6], yt[6];
  Dual const hh = cst(h * 0.5), h6 = cst(h / 6.0), hd = cst(h);
  for (int i = 0; i < 6; ++i) yt[i] = y[i] + hh * dydx[i];
  derivs_dual(f, ps, qmg, yt, dyt);
  for (int i = 0; i < 6; ++i) yt[i] = y[i] + hh * dyt[i];
  derivs_dual(f, ps, qmg, yt, dym);
  for (int i = 0; i < 6; ++i) {
    yt[i] = y[i] + hd * dym[i];
    dym[i] = dym[i] + dyt[i];
  }
  derivs_dual(f, ps, qmg, yt, dyt);
  for (int i = 0; i < 6; ++i) y[i] = y[i] + h6 * (dydx[i] + dyt[i] + cst(2.0) * dym[i]);
}

// Trajectory of the pole-scaled field in dual numbers, with pole direction m
// (m < 0: none): tr[c*n + k] for c = x, y, z, beta x, y, z, A/c x, y, z.  Point layout
// and the lagged A/c as CalculateTrajectory (OSCARSSR.cc:600-667): the forward leg
// stores each point with the derivative of the previous step (0 at T0), the backward
// leg stores each point with the derivative of the step that led to it.
void pole_trajectory_dual(const oracle_fields* f, const oracle_particle* p, double ct_start, double ct_stop,
                          int64_t npoints, int m, std::vector<Dual>& tr, int64_t* n_out, double* dt_out) {
  int const npoles = 2 * f->pole_und->nperiods;
  std::vector<Dual> ps((size_t)npoles);
  for (int i = 0; i < npoles; ++i) ps[i] = Dual{f->pole_scales[i], i == m ? 1.0 : 0.0};
  double const dt = ((ct_stop - ct_start) / kC / (size_t)(npoints - 1));
  size_t const nf = 1 + (ct_stop - p->t0) / kC / dt;
  size_t const nb = (p->t0 - ct_start) / kC / dt;
  int64_t const n = (int64_t)(nf + nb);
  tr.assign((size_t)9 * n, cst(0));
  Dual x[6], d[6];
  auto init = [&]() {
    x[0] = cst(p->x0[0]);
    x[1] = cst(p->beta0[0] * kC);
    x[2] = cst(p->x0[1]);
    x[3] = cst(p->beta0[1] * kC);
    x[4] = cst(p->x0[2]);
    x[5] = cst(p->beta0[2] * kC);
    for (int c = 0; c < 6; ++c) d[c] = cst(0);
  };
  auto put = [&](int64_t k) {
    tr[0 * n + k] = x[0];
    tr[1 * n + k] = x[2];
    tr[2 * n + k] = x[4];
    tr[3 * n + k] = x[1] / cst(kC);
    tr[4 * n + k] = x[3] / cst(kC);
    tr[5 * n + k] = x[5] / cst(kC);
    tr[6 * n + k] = d[1] / cst(kC);
    tr[7 * n + k] = d[3] / cst(kC);
    tr[8 * n + k] = d[5] / cst(kC);
  };
  init();
  for (size_t i = 0; i != nf; ++i) {
    put((int64_t)(nb + i));
    derivs_dual(f, ps.data(), p->qmg, x, d);
    rk4_dual(f, ps.data(), p->qmg, x, d, dt);
  }
  init();
  for (size_t i = 0; i != nb; ++i) {
    derivs_dual(f, ps.data(), p->qmg, x, d);
    rk4_dual(f, ps.data(), p->qmg, x, d, -dt);
    put((int64_t)(nb - 1 - i));
  }
  *n_out = n;
  *dt_out = dt;
}

// Flux loss and its derivative along pole direction m (m < 0: no direction).
Dual pole_loss_dual(const oracle_fields* f, const oracle_particle* p, double ct_start, double ct_stop,
                    int64_t npoints, const double* obs, int64_t ns, double energy_ev, const double* target, int m) {
  std::vector<Dual> tr;
  int64_t n;
  double dt;
  pole_trajectory_dual(f, p, ct_start, ct_stop, npoints, m, tr, &n, &dt);
  double const omega = ev_to_omega(energy_ev);
  double const q = p->q, cur = p->current;
  double const C0 = q / (kFourPi * kC * kEps0 * kSqrt2Pi);
  double const C2 = kFourPi * cur / (kH * std::fabs(q) * kMu0 * kC) * 1e-6 * 0.001;
  double const kw = (C0 * omega) * dt;
  double const cw = kC / omega;
  Dual L = cst(0);
  for (int64_t j = 0; j < ns; ++j) {
    Dual sr[3] = {cst(0), cst(0), cst(0)}, si[3] = {cst(0), cst(0), cst(0)};
    for (int64_t k = 0; k < n; ++k) {
      Dual const R[3] = {cst(obs[3 * j]) - tr[k], cst(obs[3 * j + 1]) - tr[n + k], cst(obs[3 * j + 2]) - tr[2 * n + k]};
      Dual const D = dsqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
      Dual const ph = cst(omega) * (cst(dt * k) + D / cst(kC));
      Dual const c = dcos(ph), s = dsin(ph);
      for (int a = 0; a < 3; ++a) {
        Dual const na = R[a] / D;
        Dual const re = (tr[(3 + a) * n + k] - na) / D;   // real amplitude
        Dual const im = -(cst(cw) * na) / (D * D);        // imaginary amplitude
        sr[a] = sr[a] + (re * c - im * s);
        si[a] = si[a] + (re * s + im * c);
      }
    }
    Dual F = cst(0);
    for (int a = 0; a < 3; ++a) F = F + cst(kw * kw) * (sr[a] * sr[a] + si[a] * si[a]);
    F = cst(C2) * F;
    Dual const e = F - cst(target[j]);
    L = L + e * e;
  }
  return L;
}

// Power-density loss L = sum_j (P_j - target_j)^2 and its derivative along pole
// direction m.  P_j is the integrand of OSCARSSR.cc:1522-1573 with
// (Num.N2)^2 + (Num.N3)^2 taken as |Num|^2 (N2, N3 orthonormal and orthogonal to
// N1, Num orthogonal to N1), the skip of points behind the surface when directional,
// the prefactor |qI|/(16 pi^2 eps0 c) dt / 1e6 and |P| when not directional.
Dual pole_pd_loss_dual(const oracle_fields* f, const oracle_particle* p, double ct_start, double ct_stop,
                       int64_t npoints, const double* obs, const double* nrm, int64_t ns, bool directional,
                       const double* target, int m, double* pd_out) {
  std::vector<Dual> tr;
  int64_t n;
  double dt;
  pole_trajectory_dual(f, p, ct_start, ct_stop, npoints, m, tr, &n, &dt);
  double const scale = std::fabs(p->q * p->current) / (16 * kPi2 * kEps0 * kC) * dt;
  Dual L = cst(0);
  for (int64_t j = 0; j < ns; ++j) {
    Dual sum = cst(0);
    for (int64_t k = 0; k < n; ++k) {
      Dual const R[3] = {cst(obs[3 * j]) - tr[k], cst(obs[3 * j + 1]) - tr[n + k], cst(obs[3 * j + 2]) - tr[2 * n + k]};
      Dual const d2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
      Dual const D = dsqrt(d2);
      Dual const N[3] = {R[0] / D, R[1] / D, R[2] / D};
      Dual const nd = N[0] * cst(nrm[3 * j]) + N[1] * cst(nrm[3 * j + 1]) + N[2] * cst(nrm[3 * j + 2]);
      if (directional && nd.v <= 0) continue;
      Dual const B[3] = {tr[3 * n + k], tr[4 * n + k], tr[5 * n + k]};
      Dual const A[3] = {tr[6 * n + k], tr[7 * n + k], tr[8 * n + k]};
      Dual const u[3] = {N[0] - B[0], N[1] - B[1], N[2] - B[2]};
      Dual const v[3] = {u[1] * A[2] - u[2] * A[1], u[2] * A[0] - u[0] * A[2], u[0] * A[1] - u[1] * A[0]};
      Dual const w[3] = {N[1] * v[2] - N[2] * v[1], N[2] * v[0] - N[0] * v[2], N[0] * v[1] - N[1] * v[0]};
      Dual const num2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
      Dual const b = cst(1) - (B[0] * N[0] + B[1] * N[1] + B[2] * N[2]);
      Dual const b2 = b * b;
      sum = sum + num2 * nd / (b2 * b2 * b * d2);
    }
    Dual P = cst(scale) * sum / cst(1e6);
    if (!directional && P.v < 0) P = -P;
    if (pd_out) pd_out[j] = P.v;
    Dual const e = P - cst(target[j]);
    L = L + e * e;
  }
  return L;
}

}  // namespace

extern "C" {

void oracle_ideal_electron(double energy_gev, const double x0[3], const double d0[3], double current, oracle_particle* out) {
  // TParticleBeam ctor (.cc:58-69), SetParticleType electron (TParticleA.cc:71-73),
  // SetCurrent (:254-266), GetNewParticle("ideal") (TParticleBeam.cc:200-222)
  std::memset(out, 0, sizeof(*out));
  out->q = -kQe;
  out->m = kMe;
  out->current = current != 0 ? current : out->q;
  V3 const u0 = unit(v3(d0));
  double const gamma = energy_gev / kg_to_gev(out->m);
  double const beta = std::sqrt(1.0 - 1.0 / (gamma * gamma));
  V3 const b0 = muls(u0, beta);
  out->x0[0] = x0[0];
  out->x0[1] = x0[1];
  out->x0[2] = x0[2];
  out->beta0[0] = b0.x;
  out->beta0[1] = b0.y;
  out->beta0[2] = b0.z;
  out->t0 = 0;
  finish_particle(out);
}

void oracle_rng_seed(int seed, int reset_normal) {
  // TRandomA::SetSeed recreates the engine but keeps the normal distribution's cached
  // value (TRandomA.cc:42-48); reset_normal clears it as a fresh process would have it.
  delete gMT;
  gMT = new std::mt19937(seed);
  if (reset_normal) gNormal.reset();
}
double oracle_rng_normal(void) { return gNormal(*gMT); }
double oracle_rng_uniform(void) { return gUniform(*gMT); }

// TParticleBeam::GetNewParticle (TParticleBeam.cc:228-265) for given draws n[0..4] in
// the reference's order: energy, horizontal, vertical position, horizontal, vertical
// angle.  next() supplies them one at a time (the reference draws as it goes).
}  // extern "C"
template <class Next>
static void beam_particle(double energy_gev, double sigma_e_gev, V3 const& x0, V3 const& u0, V3 const& h, V3 const& v,
                          double su0, double su1, double sp0, double sp1, V3 const& lattice_ref, Next next,
                          oracle_particle* out) {
  double const enew = energy_gev + sigma_e_gev * next();
  double const gamma = enew / kg_to_gev(out->m);
  double const b = std::sqrt(1.0 - 1.0 / (gamma * gamma));
  double const dmid = dot(sub(lattice_ref, x0), u0);
  V3 X = x0;
  double const n1 = next();
  V3 const dx = muls(muls(muls(h, su0), 1 + dmid), n1);
  X.x += dx.x;
  X.y += dx.y;
  X.z += dx.z;
  double const n2 = next();
  V3 const dy = muls(muls(muls(v, su1), 1 + dmid), n2);
  X.x += dy.x;
  X.y += dy.y;
  X.z += dy.z;
  V3 B = muls(u0, b);
  rot_y(B, sp0 * next());
  rot_x(B, -sp1 * next());
  out->x0[0] = X.x;
  out->x0[1] = X.y;
  out->x0[2] = X.z;
  out->beta0[0] = B.x;
  out->beta0[1] = B.y;
  out->beta0[2] = B.z;
  finish_particle(out);
}
extern "C" {

void oracle_random_electron(double energy_gev, double sigma_e_gev, const double x0[3], const double d0[3],
                            const double hdir[3], const double beta[2], const double emit[2],
                            const double lattice_ref[3], double current, oracle_particle* out) {
  // Sigmas from emittance and beta (OSCARSSR_Python.cc:1822-1829), directions from
  // TParticleBeam::SetSigma (TParticleBeam.cc:137-158), draw order E, x, y, x', y'
  // (TParticleBeam.cc:228-265).
  std::memset(out, 0, sizeof(*out));
  out->q = -kQe;
  out->m = kMe;
  out->current = current != 0 ? current : out->q;
  V3 const u0 = unit(v3(d0));
  V3 const h = unit(unit(v3(hdir)));
  V3 const v = unit(cross(u0, h));
  double const su0 = std::sqrt(emit[0] * beta[0]), su1 = std::sqrt(emit[1] * beta[1]);
  double const sp0 = std::sqrt(emit[0] / beta[0]), sp1 = std::sqrt(emit[1] / beta[1]);
  beam_particle(energy_gev, sigma_e_gev, v3(x0), u0, h, v, su0, su1, sp0, sp1, v3(lattice_ref),
                [] { return gNormal(*gMT); }, out);
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2,
// 3", SC'11): 10 rounds of two 32x32->64 multiplications with the published multipliers
// and Weyl key increments.
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]}, k[2] = {key[0], key[1]};
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k[0] += 0x9E3779B9u;
      k[1] += 0xBB67AE85u;
    }
    uint64_t const p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t const n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k[0], n1 = (uint32_t)p1;
    uint32_t const n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k[1], n3 = (uint32_t)p0;
    c[0] = n0;
    c[1] = n1;
    c[2] = n2;
    c[3] = n3;
  }
  for (int i = 0; i < 4; ++i) out[i] = c[i];
}

// Six standard normals of counter idx under key seed: three Philox blocks (counter
// (lo idx, hi idx, j, 0), j = 0, 1, 2), each two 53-bit uniforms u1 in (0, 1], u2 in
// [0, 1) and the Box-Muller pair sqrt(-2 ln u1) (cos, sin)(2 pi u2).
void oracle_counter_normals(uint64_t seed, uint64_t idx, double out[6]) {
  uint32_t const key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int j = 0; j < 3; ++j) {
    uint32_t const ctr[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), (uint32_t)j, 0u};
    uint32_t w[4];
    oracle_philox4x32_10(ctr, key, w);
    uint64_t const a = (((uint64_t)w[0] << 32) | w[1]) >> 11, b = (((uint64_t)w[2] << 32) | w[3]) >> 11;
    double const u1 = std::ldexp((double)(a + 1), -53), u2 = std::ldexp((double)b, -53);
    double const r = std::sqrt(-2.0 * std::log(u1));
    out[2 * j] = r * std::cos(2.0 * kPi * u2);
    out[2 * j + 1] = r * std::sin(2.0 * kPi * u2);
  }
}

// Beam index for counter idx by weight (TParticleBeamContainer::GetRandomBeamIndexByWeight,
// TParticleBeamContainer.cc:109-142): the first k with u * W < w_0 + ... + w_k, u the
// 53-bit uniform of a fourth Philox block (counter (lo, hi, 3, 0)); 0 for one beam.
int oracle_beam_index(const double* weights, int nb, uint64_t seed, uint64_t idx) {
  if (nb <= 1) return 0;
  double sum = 0;
  for (int k = 0; k < nb; ++k) sum += weights[k];
  uint32_t const key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t const ctr[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), 3u, 0u};
  uint32_t w[4];
  oracle_philox4x32_10(ctr, key, w);
  uint64_t const a = (((uint64_t)w[0] << 32) | w[1]) >> 11;
  double const r = std::ldexp((double)a, -53) * sum;
  double acc = 0;
  for (int k = 0; k < nb; ++k) {
    acc += weights[k];
    if (r < acc) return k;
  }
  return nb - 1;
}

void oracle_sample_beam(const oracle_beam* bm, uint64_t seed, uint64_t idx, oracle_particle* out) {
  // the reference's particle transform on the counter-based draws n[0..4]
  std::memset(out, 0, sizeof(*out));
  out->q = bm->charge;
  out->m = bm->mass;
  out->current = bm->current;
  double n[6];
  oracle_counter_normals(seed, idx, n);
  int k = 0;
  beam_particle(bm->e0, bm->sigma_e, v3(bm->x0), v3(bm->u0), v3(bm->hdir), v3(bm->vdir), bm->sigma_u[0],
                bm->sigma_u[1], bm->sigma_up[0], bm->sigma_up[1], v3(bm->sigma_at), [&] { return n[k++]; }, out);
  out->t0 = bm->t0;
}

int64_t oracle_trajectory_npoints(double ct_start, double ct_stop, int64_t npoints, double t0) {
  double const dt = ((ct_stop - ct_start) / kC / (size_t)(npoints - 1));
  size_t const nf = 1 + (ct_stop - t0) / kC / dt;
  size_t const nb = (t0 - ct_start) / kC / dt;
  return (int64_t)(nf + nb);
}

void oracle_efield(const oracle_fields* f, const double x[3], double out[3]) {
  V3 e = efield_sum(f, v3(x));
  out[0] = e.x;
  out[1] = e.y;
  out[2] = e.z;
}

void oracle_bfield(const oracle_fields* f, const double x[3], double out[3]) {
  V3 b = field_sum(f, v3(x));
  out[0] = b.x;
  out[1] = b.y;
  out[2] = b.z;
}

int64_t oracle_trajectory(const oracle_fields* f, const oracle_particle* p, double ct_start, double ct_stop,
                          int64_t npoints, double* traj, double* dt_out) {
  // OSCARSSR::CalculateTrajectory(TParticleA&), OSCARSSR.cc:562-673.  The forward
  // pass stores the state *before* each step with the derivative of the previous
  // step (A/c lags one point); the first point's A/c, uninitialised in the reference,
  // is stored as 0.
  double const dt = ((ct_stop - ct_start) / kC / (size_t)(npoints - 1));
  size_t const nf = 1 + (ct_stop - p->t0) / kC / dt;
  size_t const nb = (p->t0 - ct_start) / kC / dt;
  int64_t const n = (int64_t)(nf + nb);
  *dt_out = dt;
  double x[6], dxdt[6] = {0, 0, 0, 0, 0, 0};
  x[0] = p->x0[0];
  x[1] = p->beta0[0] * kC;
  x[2] = p->x0[1];
  x[3] = p->beta0[1] * kC;
  x[4] = p->x0[2];
  x[5] = p->beta0[2] * kC;
  auto put = [&](int64_t k) {
    traj[0 * n + k] = x[0];
    traj[1 * n + k] = x[2];
    traj[2 * n + k] = x[4];
    traj[3 * n + k] = x[1] / kC;
    traj[4 * n + k] = x[3] / kC;
    traj[5 * n + k] = x[5] / kC;
    traj[6 * n + k] = dxdt[1] / kC;
    traj[7 * n + k] = dxdt[3] / kC;
    traj[8 * n + k] = dxdt[5] / kC;
  };
  for (size_t i = 0; i != nf; ++i) {
    put((int64_t)(nb + i));
    derivs(f, p->qmg, x, dxdt);
    rk4(f, p->qmg, x, dxdt, dt);
  }
  x[0] = p->x0[0];
  x[1] = p->beta0[0] * kC;
  x[2] = p->x0[1];
  x[3] = p->beta0[1] * kC;
  x[4] = p->x0[2];
  x[5] = p->beta0[2] * kC;
  double const mdt = -dt;
  for (size_t i = 0; i != nb; ++i) {
    derivs(f, p->qmg, x, dxdt);
    rk4(f, p->qmg, x, dxdt, mdt);
    put((int64_t)(nb - 1 - i));
  }
  return n;
}

void oracle_energy_grid(int64_t n, double e0, double e1, double* out) {
  // TSpectrumContainer::Init(N, EFirst, ELast), TSpectrumContainer.cc:55-86
  if (n == 1) {
    out[0] = e0;
    return;
  }
  for (int64_t i = 0; i < n; ++i) out[i] = e0 + (e1 - e0) / (size_t)(n - 1) * (double)(i + 1);
}

}  // extern "C"
// Points, normal and frame coordinates of a rectangle from its start corner, step
// vectors and step sizes (TSurfacePoints_Rectangle::GetXYZ :222-230: start + i1 a1 + i2 a2;
// normal a1 x a2 / |a1 x a2|, reversed for Normal = -1; GetX1/X2 :244-259).
static int fill_rect(V3 st, V3 a1, V3 a2, double s1, double s2, int nx1, int nx2, int normal, double* xyz,
                     double* nrm, double* x1x2) {
  V3 nv = unit(cross(a1, a2));
  if (normal == -1) nv = muls(nv, -1);
  else if (normal != 0 && normal != 1) return -2;
  int64_t const np = (int64_t)nx1 * nx2;
  for (int64_t i = 0; i < np; ++i) {
    int const i1 = (int)(i / nx2), i2 = (int)(i % nx2);
    V3 p = st;
    V3 t1 = muls(a1, (double)i1);
    p = V3{p.x + t1.x, p.y + t1.y, p.z + t1.z};
    V3 t2 = muls(a2, (double)i2);
    p = V3{p.x + t2.x, p.y + t2.y, p.z + t2.z};
    xyz[3 * i] = p.x;
    xyz[3 * i + 1] = p.y;
    xyz[3 * i + 2] = p.z;
    nrm[3 * i] = nv.x;
    nrm[3 * i + 1] = nv.y;
    nrm[3 * i + 2] = nv.z;
    x1x2[2 * i] = s1 * i1 - std::sqrt(mag2(a1)) * (nx1 - 1) / 2.;
    x1x2[2 * i + 1] = s2 * i2 - std::sqrt(mag2(a2)) * (nx2 - 1) / 2.;
  }
  return 0;
}
extern "C" {

int oracle_rect_surface(const char* plane, int nx1, int nx2, double w1, double w2, const double rot[3],
                        const double trans[3], int normal, double* xyz, double* nrm, double* x1x2) {
  // TSurfacePoints_Rectangle::Init(plane,...) :130-210, GetXYZ :222-230, GetX1/X2 :244-259
  std::string P(plane);
  for (auto& ch : P) ch = (char)toupper(ch);
  double const s1 = w1 / (nx1 - 1), s2 = w2 / (nx2 - 1);
  V3 st, a1, a2;
  if (P == "XY") {
    st = V3{-w1 / 2., -w2 / 2., 0}; a1 = V3{s1, 0, 0}; a2 = V3{0, s2, 0};
  } else if (P == "YX") {
    st = V3{-w2 / 2., -w1 / 2., 0}; a2 = V3{s2, 0, 0}; a1 = V3{0, s1, 0};
  } else if (P == "XZ") {
    st = V3{-w1 / 2., 0, -w2 / 2.}; a1 = V3{s1, 0, 0}; a2 = V3{0, 0, s2};
  } else if (P == "ZX") {
    st = V3{-w2 / 2., 0, -w1 / 2.}; a2 = V3{s2, 0, 0}; a1 = V3{0, 0, s1};
  } else if (P == "YZ") {
    st = V3{0, -w1 / 2., -w2 / 2.}; a1 = V3{0, s1, 0}; a2 = V3{0, 0, s2};
  } else if (P == "ZY") {
    st = V3{0, -w2 / 2., -w1 / 2.}; a2 = V3{0, s2, 0}; a1 = V3{0, 0, s1};
  } else {
    return -1;
  }
  V3 const r = v3(rot);
  rot_xyz(st, r);
  rot_xyz(a1, r);
  rot_xyz(a2, r);
  st.x += trans[0];
  st.y += trans[1];
  st.z += trans[2];
  return fill_rect(st, a1, a2, s1, s2, nx1, nx2, normal, xyz, nrm, x1x2);
}

int oracle_rect3_surface(int nx1, int nx2, const double x0[3], const double x1[3], const double x2[3], int normal,
                         double* xyz, double* nrm, double* x1x2) {
  // TSurfacePoints_Rectangle::Init(NX1, NX2, X0, X1, X2, Normal) :76-123 (corners as given,
  // already rotated and translated by the caller), GetXYZ :222-230, GetX1/X2 :244-259
  V3 const c0 = v3(x0), d1 = sub(v3(x1), c0), d2 = sub(v3(x2), c0);
  double const s1 = std::sqrt(mag2(d1)) / (nx1 - 1), s2 = std::sqrt(mag2(d2)) / (nx2 - 1);
  return fill_rect(c0, divs(d1, nx1 - 1), divs(d2, nx2 - 1), s1, s2, nx1, nx2, normal, xyz, nrm, x1x2);
}

void oracle_spectrum(const double* traj, int64_t n, double dt, const double obs[3], const double* energies_ev,
                     int64_t ne, double q, double current, double weight, int nthreads, double* out) {
  V3 const o = v3(obs);
  per_point_threads(ne, nthreads, [=](int64_t i) { out[i] = flux_point(traj, n, dt, o, ev_to_omega(energies_ev[i]), q, current, weight); });
}

void oracle_flux(const double* traj, int64_t n, double dt, const double* obs, int64_t ns, double energy_ev,
                 double q, double current, double weight, int nthreads, double* out) {
  double const omega = ev_to_omega(energy_ev);
  per_point_threads(ns, nthreads, [=](int64_t i) { out[i] = flux_point(traj, n, dt, v3(obs + 3 * i), omega, q, current, weight); });
}

void oracle_power_density(const double* traj, int64_t n, double dt, const double* obs, const double* normals,
                          int64_t ns, int directional, double q, double current, int nthreads, double* out) {
  per_point_threads(ns, nthreads, [=](int64_t i) { out[i] = pd_point(traj, n, dt, v3(obs + 3 * i), v3(normals + 3 * i), directional != 0, q, current); });
}

void oracle_efield_time(const double* traj, int64_t n, double dt, const double obs[3], double q, double* e_out,
                        double* t_out) {
  // OSCARSSR.cc:2646-2690 (Duke, Synchrotron Radiation, p74), TVector3D operation order
  const double *X = traj, *Y = traj + n, *Z = traj + 2 * n, *BX = traj + 3 * n, *BY = traj + 4 * n, *BZ = traj + 5 * n;
  const double *AX = traj + 6 * n, *AY = traj + 7 * n, *AZ = traj + 8 * n;
  double const C0 = q / (kFourPi * kEps0);
  V3 const O = v3(obs);
  for (int64_t i = 0; i < n; ++i) {
    V3 const B{BX[i], BY[i], BZ[i]};
    V3 const A{AX[i], AY[i], AZ[i]};
    V3 const R = sub(O, V3{X[i], Y[i], Z[i]});
    V3 const N = unit(R);
    double const D = std::sqrt(mag2(R));
    double const Mult = C0 * std::pow(1.0 / (1.0 - dot(N, B)), 3);
    V3 const NB = sub(N, B);
    V3 const Near = divs(muls(NB, 1.0 - mag2(B)), mag2(R));
    V3 const Far = divs(muls(cross(N, cross(NB, A)), 1.0 / kC), D);
    V3 const E = muls(V3{Near.x + Far.x, Near.y + Far.y, Near.z + Far.z}, Mult);
    e_out[3 * i] = E.x;
    e_out[3 * i + 1] = E.y;
    e_out[3 * i + 2] = E.z;
    t_out[i] = (double)i * dt + D / kC;
  }
}

void oracle_flux1(const double* traj, int64_t n, double dt, const double* obs, int64_t ns, double energy_ev, double q,
                  double current, double* out) {
  // OSCARSSR::CalculateFlux1, OSCARSSR.cc:2209-2346: the Lienard-Wiechert field with the
  // stored A/c (velocity term + acceleration term, no integration by parts), summed over
  // the trajectory in order; TVector3D / TVector3DC / std::complex operation order.
  const double *X = traj, *Y = traj + n, *Z = traj + 2 * n, *BX = traj + 3 * n, *BY = traj + 4 * n, *BZ = traj + 5 * n;
  const double *AX = traj + 6 * n, *AY = traj + 7 * n, *AZ = traj + 8 * n;
  double const C0 = q / (4 * kPi * kC * kEps0 * kSqrt2Pi);                                         // :2245
  double const C2 = kFourPi * current / (kH * std::fabs(q) * kMu0 * kC) * 1e-6 * 0.001;             // :2249
  double const Omega = energy_ev * kTwoPi / 4.1357e-15;                                            // :2255
  for (int64_t ip = 0; ip < ns; ++ip) {
    V3 const O = v3(obs + 3 * ip);
    std::complex<double> Sx(0, 0), Sy(0, 0), Sz(0, 0);
    for (int64_t iT = 0; iT < n; ++iT) {
      V3 const B{BX[iT], BY[iT], BZ[iT]};
      V3 const A{AX[iT], AY[iT], AZ[iT]};
      V3 const R = sub(O, V3{X[iT], Y[iT], Z[iT]});
      V3 const N = unit(R);
      double const D = std::sqrt(mag2(R));
      std::complex<double> const Exponent(0, -Omega * (dt * (double)(int)iT + D / kC));
      double const p2 = std::pow(1 - dot(N, B), 2);
      V3 const NB = sub(N, B);
      V3 const Near = divs(muls(NB, 1 - mag2(B)), D * D * p2);
      V3 const Far = divs(cross(N, cross(NB, A)), D * p2);
      std::complex<double> const e = std::exp(Exponent);
      Sx += std::complex<double>(Near.x + Far.x, 0) * e * dt;
      Sy += std::complex<double>(Near.y + Far.y, 0) * e * dt;
      Sz += std::complex<double>(Near.z + Far.z, 0) * e * dt;
    }
    Sx *= C0;
    Sy *= C0;
    Sz *= C0;
    out[ip] = C2 * (Sx * std::conj(Sx) + Sy * std::conj(Sy) + Sz * std::conj(Sz)).real();
  }
}

double oracle_total_power(const double* traj, int64_t n, double dt, double q, double current, double gamma) {
  // OSCARSSR::CalculateTotalPower(TParticleA&), OSCARSSR.cc:1939-1978
  double tp = 0;
  for (int64_t i = 0; i != n; ++i) {
    V3 const B{traj[3 * n + i], traj[4 * n + i], traj[5 * n + i]};
    V3 const A{traj[6 * n + i], traj[7 * n + i], traj[8 * n + i]};
    tp += (mag2(A) - mag2(cross(B, A))) * dt;
  }
  tp *= std::fabs(q * current) * std::pow(gamma, 6) / (6 * kPi * kEps0 * kC);
  return tp;
}

void oracle_kahan_add(double* sum, double* comp, const double* v, int64_t n) {
  // T3DScalarContainer::AddToPoint, T3DScalarContainer.cc:31-47
  for (int64_t i = 0; i < n; ++i) {
    double const s = sum[i];
    double const y = v[i] - comp[i];
    double const t = s + y;
    comp[i] = (t - s) - y;
    sum[i] = t;
  }
}


int oracle_pole_loss_gradient(const oracle_fields* f, const oracle_particle* p, double ct_start, double ct_stop,
                              int64_t npoints, const double* obs, int64_t ns, double energy_ev, const double* target,
                              double* loss_out, double* grad_out, int nthreads) {
  if (!f || !f->pole_und || !f->pole_scales || f->pole_und->nperiods < 1) return -1;
  int const npoles = 2 * f->pole_und->nperiods;
  *loss_out = pole_loss_dual(f, p, ct_start, ct_stop, npoints, obs, ns, energy_ev, target, -1).v;
  per_point_threads(npoles, nthreads, [&](int64_t m) {
    grad_out[m] = pole_loss_dual(f, p, ct_start, ct_stop, npoints, obs, ns, energy_ev, target, (int)m).d;
  });
  return 0;
}

int oracle_pole_pd_loss_gradient(const oracle_fields* f, const oracle_particle* p, double ct_start, double ct_stop,
                                 int64_t npoints, const double* obs, const double* normals, int64_t ns,
                                 int directional, const double* target, double* pd_out, double* loss_out,
                                 double* grad_out, int nthreads) {
  if (!f || !f->pole_und || !f->pole_scales || f->pole_und->nperiods < 1) return -1;
  int const npoles = 2 * f->pole_und->nperiods;
  *loss_out =
      pole_pd_loss_dual(f, p, ct_start, ct_stop, npoints, obs, normals, ns, directional != 0, target, -1, pd_out).v;
  per_point_threads(npoles, nthreads, [&](int64_t m) {
    grad_out[m] = pole_pd_loss_dual(f, p, ct_start, ct_stop, npoints, obs, normals, ns, directional != 0, target,
                                    (int)m, nullptr).d;
  });
  return 0;
}

}  // extern "C"
