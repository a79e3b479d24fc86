// Power-density integrand (OSCARSSR.cc:1522-1573) for one (observation point,
// trajectory point) pair, and its derivative, shared by the single-trajectory
// kernel, the batched realisation kernel and the adjoint.
//
//   R = O - X,  d2 = |R|^2,  N = R/|R|,  nd = N.n,  b = 1 - beta.N
//   Num = N x ((N - beta) x A)          (|Num|^2 = (Num.N2)^2 + (Num.N3)^2, N2, N3
//                                         orthonormal and orthogonal to N)
//   term = |Num|^2 nd / (b^5 d2), skipped when directional and nd <= 0.
// The forward keeps the reference's operation order for N, b and the quotient
// (b cancels to ~1e-8, so it must round as the reference's does).
#pragma once
#include "srs_device.h"

namespace srs {

// EXACT: every component of N1 as the reference's quotient (for points whose beam is not
// along z, transverse_beam); else n1z correctly rounded (1 - beta.N1 and n1z - bz cancel
// to ~1e-8 through it) and n1x, n1y by the reciprocal: beta_x n1x is ~1e-7 of 1, so
// their last bit is ~1e-15 of 1 - beta.N1.
template <bool EXACT = false>
__device__ __forceinline__ double pd_term(double ox, double oy, double oz, double qx, double qy, double qz, double X,
                                          double Y, double Z, double bx, double by, double bz, double Ax, double Ay,
                                          double Az, int directional) {
  double const rx = ox - X, ry = oy - Y, rz = oz - Z;
  double const d2 = rx * rx + ry * ry + rz * rz;
  double const D = __builtin_sqrt(d2);
  double const inv = 1.0 / D;
  double const n1x = EXACT ? div_rn(rx, D, inv) : rx * inv, n1y = EXACT ? div_rn(ry, D, inv) : ry * inv;
  double const n1z = div_rn(rz, D, inv);
  double const ndot = n1x * qx + n1y * qy + n1z * qz;
  if (directional && ndot <= 0) return 0.0;
  double const den = 1 - (bx * n1x + by * n1y + bz * n1z);
  double const ux = n1x - bx, uy = n1y - by, uz = n1z - bz;
  double const vx = uy * Az - Ay * uz, vy = uz * Ax - Az * ux, vz = ux * Ay - Ax * uy;
  double const mx = n1y * vz - vy * n1z, my = n1z * vx - vz * n1x, mz = n1x * vy - vx * n1y;
  double const num2 = __builtin_fma(mx, mx, __builtin_fma(my, my, mz * mz));
  double const den2 = den * den;
  double const den5 = den2 * den2 * den;
  return num2 * ndot / (den5 * d2);
}

// Cotangents of one term scaled by G = dL/d(term): adds dL/dX, dL/dbeta, dL/dA.
// With u = N - beta, NA = N.A:  Num = u NA - b A (the triple product expanded with
// |N| = 1; the extra radial part this drops is removed by the projection below).
//   s = |Num|^2, m = 2 G nd/(b^5 d2) Num, val = s nd/(b^5 d2):
//   dL/dA    = N (u.m) - b m
//   dL/dbeta = -NA m + N (A.m) + 5 G val/b N
//   dL/dN    = NA m + A (u.m) + beta (A.m) + G s/(b^5 d2) n + 5 G val/b beta
//   dL/dR    = (I - N N^T) dL/dN / D - 2 G val/D N,   dL/dX = -dL/dR.
template <bool EXACT = false>
__device__ __forceinline__ void pd_term_adjoint(double G, double ox, double oy, double oz, double qx, double qy,
                                                double qz, const double* X, const double* B, const double* A,
                                                int directional, double* gX, double* gB, double* gA) {
  double const r[3] = {ox - X[0], oy - X[1], oz - X[2]};
  double const d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  double const D = __builtin_sqrt(d2);
  double const inv = 1.0 / D;
  // N and b rounded exactly as in pd_term<EXACT>: b = 1 - beta.N cancels to ~1e-8, and
  // its rounding enters the derivative through 1/b^6
  double const N[3] = {EXACT ? div_rn(r[0], D, inv) : r[0] * inv, EXACT ? div_rn(r[1], D, inv) : r[1] * inv,
                       div_rn(r[2], D, inv)};
  double const nd = N[0] * qx + N[1] * qy + N[2] * qz;
  if (directional && nd <= 0) return;
  double const b = 1 - (B[0] * N[0] + B[1] * N[1] + B[2] * N[2]);
  double const NA = N[0] * A[0] + N[1] * A[1] + N[2] * A[2];
  double const u[3] = {N[0] - B[0], N[1] - B[1], N[2] - B[2]};
  double num[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) num[c] = u[c] * NA - b * A[c];
  double const s = num[0] * num[0] + num[1] * num[1] + num[2] * num[2];
  double const ib = 1.0 / b;
  double const ib2 = ib * ib;
  double const w = ib2 * ib2 * ib * (inv * inv);  // 1/(b^5 d2)
  double const val = s * nd * w;
  double const g2 = 2.0 * G * nd * w;
  double m[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) m[c] = g2 * num[c];
  double const um = u[0] * m[0] + u[1] * m[1] + u[2] * m[2];
  double const Am = A[0] * m[0] + A[1] * m[1] + A[2] * m[2];
  double const g5 = 5.0 * G * val * ib;
  double const gn = G * s * w;
  double gN[3];
  double const q[3] = {qx, qy, qz};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    gA[c] += N[c] * um - b * m[c];
    gB[c] += -NA * m[c] + N[c] * Am + g5 * N[c];
    gN[c] = NA * m[c] + A[c] * um + B[c] * Am + gn * q[c] + g5 * B[c];
  }
  double const Ng = N[0] * gN[0] + N[1] * gN[1] + N[2] * gN[2];
  double const gD = -2.0 * G * val * inv;
#pragma unroll
  for (int c = 0; c < 3; ++c) gX[c] -= (gN[c] - N[c] * Ng) * inv + N[c] * gD;
}

// Mixed (FP32) power-density term, shared by pd_mixed_kernel and the batched kernel;
// the formulation (cancellation-free 1 - beta.n, scaling by S = 2^10) is described at
// pd_mixed_kernel.  P = (x, y, z, S^2 (1 - |beta|^2)), B = S (bx, by, 1 - bz), A = A/c;
// the result is 2^-55 times the term.
constexpr float kPdS = 0x1p10f;

template <bool DIR>
__device__ __forceinline__ float pd_mixed_term(float ox, float oy, float oz, float qx, float qy, float qz, float4 P,
                                               float4 B, float4 A) {
  float const rx = ox - P.x, ry = oy - P.y, rz = oz - P.z;
  float const inv = __builtin_amdgcn_rsqf(__builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, rz * rz)));
  float const is = inv * kPdS;
  float const nx = rx * is, ny = ry * is, nz = rz * is;  // S n
  float const v = __builtin_fmaf(nx, nx, ny * ny);
  // S (1 - n_z) = w / a with a = S + |S n_z|: w = S^2 (n_x^2 + n_y^2) for n_z > 0, a^2 else;
  // the select picks between two plain values, so no branch is generated
  float const a = kPdS + __builtin_fabsf(nz);
  float const omnz = (nz > 0.f ? v : a * a) * __builtin_amdgcn_rcpf(a);
  float const ux = nx - B.x, uy = ny - B.y, uz = B.z - omnz;                       // S u
  float const d2 = __builtin_fmaf(ux, ux, __builtin_fmaf(uy, uy, __builtin_fmaf(uz, uz, P.w)));
  float const na = __builtin_fmaf(nx, A.x, __builtin_fmaf(ny, A.y, nz * A.z));     // S n.A
  float const hd = 0.5f * d2;                                                      // S^2 (1 - beta.n)
  float const mx = __builtin_fmaf(ux, na, -(A.x * hd)), my = __builtin_fmaf(uy, na, -(A.y * hd));
  float const mz = __builtin_fmaf(uz, na, -(A.z * hd));                            // S^2 n x (u x A)
  float const num2 = __builtin_fmaf(mx, mx, __builtin_fmaf(my, my, mz * mz));
  float const ndot = __builtin_fmaf(nx, qx, __builtin_fmaf(ny, qy, nz * qz));      // S n.normal
  float const rd = __builtin_amdgcn_rcpf(d2);
  float const rd2 = rd * rd;
  float const term = num2 * ndot * (inv * inv) * (rd2 * rd2 * rd);
  return (DIR && ndot <= 0.f) ? 0.f : term;
}

}  // namespace srs
