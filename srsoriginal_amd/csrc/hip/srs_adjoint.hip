// Adjoint of the flux and power-density losses with respect to per-pole field errors.
//
// Forward: realisation r is a particle whose undulator poles carry their own deltas;
// rk4_batch_kernel (srs_batch.hip) integrates it, checkpointing the state every C
// steps.  Flux: batch_flux_kernel (FP64) returns its flux map F_r and the six field
// sums S_r of every point.  Power density: pd_batch_kernel + pd_batch_reduce_kernel
// (srs_batch.hip) return its map P_r and the signed sums before |.|.  Loss L_r = sum_j (F_rj - t_j)^2
// (or P_rj).
//
// Backward:
//   flux_adjoint_kernel  lanes over (r, trajectory point k), loop over observation
//                        points staged in LDS: dL/dX_k and dL/dbeta_k of the
//                        radiation integral (analytic derivative of every term);
//   pd_adjoint_kernel    the same for the power density: dL/dX_k, dL/dbeta_k and
//                        dL/dA_k (the stored acceleration, whose cotangent joins the
//                        first RK4 stage of the step that produced it);
//   rk4_record_kernel    recomputes checkpoint segments from their checkpoints, all
//                        (segment, realisation) pairs in parallel, recording each RK4
//                        stage's velocity, B, dB/dD and pole sine;
//   rk4_segment_sweep_kernel  sweeps every segment in reverse in parallel (the
//                        transposed RK4 step is linear in the cotangent: a particular
//                        sweep plus six unit sweeps per segment), collecting dL/d delta_p;
//   rk4_combine_kernel   chains the segments per realisation.  Segments are processed
//                        in passes (last first) of as many as the record budget holds.
//
// Radiation-integral term (OSCARSSR.cc:2054-2092), R = O - X, D = |R|, n = R/D:
//   T_c = a_c P,  a_c = (beta_c - n_c)/D - i cw n_c/D^2,  P = exp(i omega (dt k + D/c)).
// With G_c = dL/dRe S_c + i dL/dIm S_c and H_c = conj(G_c) P:
//   dL/dbeta_c = Re(H_c)/D
//   dL/dR_d    = -(Re H_d + n_d S1)/D^2 + cw (Im H_d - 3 n_d S2)/D^3 - (omega/c) n_d S3,
//   S1 = sum_c Re H_c (beta_c - 2 n_c),  S2 = sum_c Im H_c n_c,  S3 = sum_c Im(H_c a_c),
// and dL/dX = -dL/dR.  F = C2 kw^2 sum_c |S_c|^2, so G_c = 2 (F - target) 2 C2 kw^2 S_c.
//
// RK4 step y' = y + h/6 ((k1 + k4) + 2 (k3 + k2)) with k_s = f(y_s), f = (v, qmg v x B(X)):
// for a cotangent w of f at a stage, with w_a = (w1, w3, w5) and c = qmg (w_a x v),
//   d(w.f)/dX = u (c . dB/dD),  d(w.f)/dv = (w0, w2, w4) + qmg (B x w_a),
//   d(w.f)/d delta_p = sin(theta) (c . field)   (point inside pole p),
// u being the undulators' common period direction.
#include "srs_batch.h"
#include "srs_pd.h"
#include "srs_rk4.h"
#include "srs_terms.h"

namespace srs {
namespace {

constexpr int kAdjTile = 256;
constexpr int kStageRec = 11;            // v[3], B[3], dBdD[3], sin at pole, pole index
// + the direct dL/dy of the step's point and dL/dk1 (velocity part) of its stored acceleration
constexpr int kStepRec = 4 * kStageRec + 9;
constexpr int kSweeps = 7;                 // per segment: particular + 6 unit end cotangents

__global__ void zero_kernel(double* __restrict__ a, int64_t n) {
  int64_t const i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

// L_r = sum_j (F_rj - target_j)^2, one block per realisation, fixed-order tree.
__global__ __launch_bounds__(256) void loss_kernel(const double* __restrict__ flux_pp,
                                                   const double* __restrict__ target, int ns,
                                                   double* __restrict__ loss) {
  __shared__ double red[256];
  int const r = blockIdx.x, tid = threadIdx.x;
  double s = 0.0;
  for (int j = tid; j < ns; j += 256) {
    double const d = flux_pp[(size_t)r * ns + j] - target[j];
    s += d * d;
  }
  red[tid] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) loss[r] = red[0];
}

// The FP64 terms of one block of flux_adjoint_kernel (also the fallback of the mixed
// kernel for a realisation whose beam is not along z).
__device__ __forceinline__ void flux_adjoint_fp64_block(BatchTraj const& bt, BatchParticle const& q, int r,
                                                        const double* __restrict__ obs, int ns, double omega,
                                                        double cw, double c0w, double c2,
                                                        const double* __restrict__ field_pp,
                                                        const double* __restrict__ flux_pp,
                                                        const double* __restrict__ target,
                                                        const double* __restrict__ dloss, double* __restrict__ gy,
                                                        int64_t total) {
  __shared__ double sx[kAdjTile], sy[kAdjTile], sz[kAdjTile];
  __shared__ double sgr[3][kAdjTile], sgi[3][kAdjTile];
  int const tid = threadIdx.x;
  int64_t const n = q.nb + q.nf;
  int64_t const k = (int64_t)blockIdx.x * kBlock + tid;
  bool const valid = k < n;
  int64_t const kk = valid ? k : n - 1;
  int64_t const i = q.off + kk;
  double const X = bt.x[i], Y = bt.y[i], Z = bt.z[i];
  double const b[3] = {bt.bx[i], bt.by[i], bt.bz[i]};
  double const t0 = q.dt * (double)kk;
  double const kw = c0w * q.dt;
  double const gsc = dloss[r] * 4.0 * c2 * kw * kw;
  double const woc = omega * kInvC;
  double gR[3] = {0, 0, 0}, gb[3] = {0, 0, 0};
  for (int j0 = 0; j0 < ns; j0 += kAdjTile) {
    int const cnt = (ns - j0) < kAdjTile ? ns - j0 : kAdjTile;
    __syncthreads();
    for (int t = tid; t < cnt; t += kBlock) {
      int const j = j0 + t;
      sx[t] = obs[j];
      sy[t] = obs[ns + j];
      sz[t] = obs[2 * ns + j];
      double const w = gsc * (flux_pp[(size_t)r * ns + j] - target[j]);
      size_t const base = (size_t)r * 6 * ns + j;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        sgr[c][t] = w * field_pp[base + (size_t)(2 * c) * ns];
        sgi[c][t] = w * field_pp[base + (size_t)(2 * c + 1) * ns];
      }
    }
    __syncthreads();
    for (int t = 0; t < cnt; ++t) {
      double const rr[3] = {sx[t] - X, sy[t] - Y, sz[t] - Z};
      // D and tau bitwise as the forward (correctly rounded root and quotient); 1/D only
      // scales the derivative's amplitudes, so the hardware reciprocal refined by two
      // Newton steps (within an ulp) serves in place of a full division
      double const D = dist_rn(rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2]);
      double const tau = t0 + div_rn(D, kC, kInvC);
      double inv = __builtin_amdgcn_rcp(D);
      inv = __builtin_fma(inv, __builtin_fma(-D, inv, 1.0), inv);
      inv = __builtin_fma(inv, __builtin_fma(-D, inv, 1.0), inv);
      double const inv2 = inv * inv;
      double sn, cs;
      sincos_fp64(omega * tau, &sn, &cs);
      double nn[3], hr[3], hi[3], ar[3], ai[3];
      double S1 = 0, S2 = 0, S3 = 0;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        nn[c] = rr[c] * inv;
        hr[c] = __builtin_fma(sgr[c][t], cs, sgi[c][t] * sn);
        hi[c] = __builtin_fma(sgr[c][t], sn, -(sgi[c][t] * cs));
        ar[c] = (b[c] - nn[c]) * inv;
        ai[c] = -(cw * nn[c]) * inv2;
        gb[c] = __builtin_fma(hr[c], inv, gb[c]);
        S1 = __builtin_fma(hr[c], b[c] - 2.0 * nn[c], S1);
        S2 = __builtin_fma(hi[c], nn[c], S2);
        S3 = __builtin_fma(hr[c], ai[c], __builtin_fma(hi[c], ar[c], S3));
      }
      double const cw3 = cw * inv2 * inv;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double const t1 = -inv2 * __builtin_fma(nn[d], S1, hr[d]);
        double const t2 = cw3 * __builtin_fma(-3.0 * nn[d], S2, hi[d]);
        gR[d] += t1 + t2 - woc * nn[d] * S3;
      }
    }
  }
  if (valid) {
    gy[0 * total + i] = -gR[0];
    gy[1 * total + i] = gb[0] * kInvC;
    gy[2 * total + i] = -gR[1];
    gy[3 * total + i] = gb[1] * kInvC;
    gy[4 * total + i] = -gR[2];
    gy[5 * total + i] = gb[2] * kInvC;
  }
}

__global__ __launch_bounds__(kBlock) void flux_adjoint_kernel(BatchTraj bt, const BatchParticle* __restrict__ P,
                                                              const double* __restrict__ obs, int ns, double omega,
                                                              double cw, double c0w, double c2,
                                                              const double* __restrict__ field_pp,
                                                              const double* __restrict__ flux_pp,
                                                              const double* __restrict__ target,
                                                              const double* __restrict__ dloss,
                                                              double* __restrict__ gy, int64_t total) {
  int const r = blockIdx.y;
  BatchParticle const q = P[r];
  if ((int64_t)blockIdx.x * kBlock >= q.nb + q.nf) return;  // whole block past this realisation's points
  flux_adjoint_fp64_block(bt, q, r, obs, ns, omega, cw, c0w, c2, field_pp, flux_pp, target, dloss, gy, total);
}

// Mixed form of flux_adjoint_kernel: the phase as the mixed flux terms take it (D, tau
// and omega tau in FP64, bitwise as the reference; the reduction to [-pi, pi] in FP64,
// sin/cos in FP32), everything after it in FP32: n, the amplitudes a_c with
// beta_z - n_z = (1 - n_z) - (1 - beta_z) (1 - beta_z staged by the RK4 in FP32 from
// FP64), H_c, S1..S3 and the term's dL/dR, dL/dbeta, summed over 32 observation points
// in FP32 and added into FP64 per lane.  The seeds G_c of a tile of 256 points are
// carried as G_c 2^-e (e: the exponent of the tile's largest |G|) so that the FP32
// products stay in range; 2^e is restored in FP64 with each sub-sum.  A realisation
// whose beam is not along z takes the FP64 terms.
__global__ __launch_bounds__(kBlock) void flux_adjoint_mixed_kernel(BatchTraj bt, const BatchParticle* __restrict__ P,
                                                                    const double* __restrict__ obs, int ns,
                                                                    double omega, double cw, double c0w, double c2,
                                                                    const double* __restrict__ field_pp,
                                                                    const double* __restrict__ flux_pp,
                                                                    const double* __restrict__ target,
                                                                    const double* __restrict__ dloss,
                                                                    double* __restrict__ gy, int64_t total,
                                                                    AdjScreen sc) {
  __shared__ double sx[kAdjTile], sy[kAdjTile], sz[kAdjTile];
  __shared__ float sgr[3][kAdjTile], sgi[3][kAdjTile];
  __shared__ double smax[kBlock / 64];
  static_assert(kAdjTile == kBlock, "one observation point of a tile per thread");
  int const r = blockIdx.y, tid = threadIdx.x;
  BatchParticle const q = P[r];
  int64_t const n = q.nb + q.nf;
  if ((int64_t)blockIdx.x * kBlock >= n) return;  // whole block past this realisation's points
  if (bt.tflag && bt.tflag[r]) {
    flux_adjoint_fp64_block(bt, q, r, obs, ns, omega, cw, c0w, c2, field_pp, flux_pp, target, dloss, gy, total);
    return;
  }
  int64_t const k = (int64_t)blockIdx.x * kBlock + tid;
  bool const valid = k < n;
  int64_t const kk = valid ? k : n - 1;
  int64_t const i = q.off + kk;
  double const X = bt.x[i], Y = bt.y[i], Z = bt.z[i];
  float const fbx = bt.fbx[i], fby = bt.fby[i], fomb = bt.fomb[i];  // beta_x, beta_y, 1 - beta_z
  double const t0 = q.dt * (double)kk;
  double const kw = c0w * q.dt;
  double const gsc = dloss[r] * 4.0 * c2 * kw * kw;
  float const woc = (float)(omega * kInvC), cwf = (float)cw;
  // The paraxial phase of the mixed flux terms, on a screen in one plane z = sc.z when
  // every (point, trajectory point of this block) pair meets their bounds (the same
  // as para_point_ok's, with the screen's extent): phase = frac(f (t_k - z_k/c) +
  // omega Z/(2 pi c)) + f u, u = s/(2 r_z) - s^2/(8 r_z^3), in revolutions, without the
  // root and quotients of the general form; D ~ r_z + s/(2 r_z) for the FP32 amplitudes.
  double const fr = omega / (2.0 * kPi), fw = fr * kInvC;
  double const rzd = sc.z - Z;
  bool ok = false;
  if (sc.flat) {
    double const dx = __builtin_fmax(__builtin_fabs(sc.xlo - X), __builtin_fabs(sc.xhi - X));
    double const dy = __builtin_fmax(__builtin_fabs(sc.ylo - Y), __builtin_fabs(sc.yhi - Y));
    double const eps = __builtin_fma(dx, dx, dy * dy) / (rzd * rzd);
    ok = rzd > 0.0 && eps <= kParaEps && fw * rzd * (eps * eps * eps) <= 16.0 * kParaPhase;
  }
  bool const para = __syncthreads_and(ok) != 0;
  double ph_h = 0.0, ph_t3 = 0.0, ph0 = 0.0;
  float rzf = 1.f, hzf = 0.f;
  if (para) {
    double const ir = 1.0 / rzd;
    ph_h = fw * (0.5 * ir);
    ph_t3 = fw * (0.125 * ir * ir * ir);
    double const p0 = para_point_phase(fr, t0, Z) + para_plane_phase(omega, sc.z);
    ph0 = p0 - __builtin_rint(p0);
    rzf = (float)rzd;
    hzf = (float)(0.5 * ir);
  }
  double gR[3] = {0, 0, 0}, gb[3] = {0, 0, 0};
  for (int j0 = 0; j0 < ns; j0 += kAdjTile) {
    int const cnt = (ns - j0) < kAdjTile ? ns - j0 : kAdjTile;
    __syncthreads();
    // this thread's seeds (one point of the tile), their largest magnitude over the block
    double g[6] = {0, 0, 0, 0, 0, 0};
    double m = 0.0;
    if (tid < cnt) {
      int const j = j0 + tid;
      sx[tid] = obs[j];
      sy[tid] = obs[ns + j];
      sz[tid] = obs[2 * ns + j];
      double const w = gsc * (flux_pp[(size_t)r * ns + j] - target[j]);
      size_t const base = (size_t)r * 6 * ns + j;
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        g[c] = w * field_pp[base + (size_t)c * ns];
        m = __builtin_fmax(m, __builtin_fabs(g[c]));
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = __builtin_fmax(m, __shfl_xor(m, o, 64));
    if ((tid & 63) == 0) smax[tid >> 6] = m;
    __syncthreads();
    double tm = smax[0];
#pragma unroll
    for (int w = 1; w < kBlock / 64; ++w) tm = __builtin_fmax(tm, smax[w]);
    int const e = tm > 0.0 ? ilogb(tm) : 0;  // tile seeds carried as G 2^-e
    if (tid < cnt) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        sgr[c][tid] = (float)ldexp(g[2 * c], -e);
        sgi[c][tid] = (float)ldexp(g[2 * c + 1], -e);
      }
    }
    __syncthreads();
    // the terms of points [t0s, t1s) of the tile: paraxial (PARA) or general phase
    auto terms = [&](auto PARA, int t0s, int t1s, float* fR, float* fb) {
      for (int t = t0s; t < t1s; ++t) {
        double const rx = sx[t] - X, ry = sy[t] - Y;
        float sn, cs, inv, nz;
        if constexpr (decltype(PARA)::value) {
          double const sq = __builtin_fma(rx, rx, ry * ry);
          double const hs = __builtin_fma(-ph_t3, sq, ph_h);
          float const ph = (float)__builtin_amdgcn_fract(__builtin_fma(sq, hs, ph0));
          sn = __builtin_amdgcn_sinf(ph);
          cs = __builtin_amdgcn_cosf(ph);
          inv = __builtin_amdgcn_rcpf(__builtin_fmaf((float)sq, hzf, rzf));
          nz = rzf * inv;
        } else {
          double const rz = sz[t] - Z;
          double const D = dist_rn(rx * rx + ry * ry + rz * rz);
          double const tau = t0 + div_rn(D, kC, kInvC);
          sincos_reduced_f32(reduce_two_pi(omega * tau), &sn, &cs);
          inv = __builtin_amdgcn_rcpf((float)D);
          nz = (float)rz * inv;
        }
        float const inv2 = inv * inv;
        float const nx = (float)rx * inv, ny = (float)ry * inv;
        float const v = __builtin_fmaf(nx, nx, ny * ny);
        float const omnz = nz > 0.f ? v * __builtin_amdgcn_rcpf(1.f + nz) : 1.f - nz;  // 1 - n_z
        float const nn[3] = {nx, ny, nz};
        float const bmn[3] = {fbx - nx, fby - ny, omnz - fomb};  // beta - n
        float hr[3], hi[3];
        float S1 = 0.f, S2 = 0.f, S3 = 0.f;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          hr[c] = __builtin_fmaf(sgr[c][t], cs, sgi[c][t] * sn);
          hi[c] = __builtin_fmaf(sgr[c][t], sn, -(sgi[c][t] * cs));
          float const ar = bmn[c] * inv, ai = -(cwf * nn[c]) * inv2;
          fb[c] = __builtin_fmaf(hr[c], inv, fb[c]);
          S1 = __builtin_fmaf(hr[c], bmn[c] - nn[c], S1);  // beta_c - 2 n_c
          S2 = __builtin_fmaf(hi[c], nn[c], S2);
          S3 = __builtin_fmaf(hr[c], ai, __builtin_fmaf(hi[c], ar, S3));
        }
        float const cw3 = cwf * inv2 * inv;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          float const t1 = -inv2 * __builtin_fmaf(nn[d], S1, hr[d]);
          float const t2 = cw3 * __builtin_fmaf(-3.f * nn[d], S2, hi[d]);
          fR[d] += t1 + t2 - woc * nn[d] * S3;
        }
      }
    };
    for (int t0s = 0; t0s < cnt; t0s += kSub) {
      int const t1s = (t0s + kSub < cnt) ? t0s + kSub : cnt;
      float fR[3] = {0.f, 0.f, 0.f}, fb[3] = {0.f, 0.f, 0.f};
      if (para) terms(std::true_type{}, t0s, t1s, fR, fb);
      else terms(std::false_type{}, t0s, t1s, fR, fb);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        gR[c] += ldexp((double)fR[c], e);
        gb[c] += ldexp((double)fb[c], e);
      }
    }
  }
  if (valid) {
    gy[0 * total + i] = -gR[0];
    gy[1 * total + i] = gb[0] * kInvC;
    gy[2 * total + i] = -gR[1];
    gy[3 * total + i] = gb[1] * kInvC;
    gy[4 * total + i] = -gR[2];
    gy[5 * total + i] = gb[2] * kInvC;
  }
}

// Power-density adjoint: lanes over (r, trajectory point k), observation points with
// their seeds G_rj = dloss_r 2 (P_rj - t_j) scale/1e6 sign(raw_rj) staged through LDS
// (sign 1 when directional: no |.| there).  Writes dL/dy of every point to gy (x, vx,
// y, vy, z, vz; dL/dv = dL/dbeta / c) and dL/dA to ga.
__global__ __launch_bounds__(kBlock) void pd_adjoint_kernel(BatchTraj bt, const BatchParticle* __restrict__ P,
                                                            const double* __restrict__ obs,
                                                            const double* __restrict__ nrm, int ns, int directional,
                                                            double scale, const double* __restrict__ raw,
                                                            const double* __restrict__ pd,
                                                            const double* __restrict__ target,
                                                            const double* __restrict__ dloss, double* __restrict__ gy,
                                                            double* __restrict__ ga, int64_t total) {
  __shared__ double so[6][kAdjTile], sg[kAdjTile];
  int const r = blockIdx.y, tid = threadIdx.x;
  BatchParticle const q = P[r];
  int64_t const n = q.nb + q.nf;
  int64_t const k = (int64_t)blockIdx.x * kBlock + tid;
  if ((int64_t)blockIdx.x * kBlock >= n) return;  // whole block past this realisation's points
  bool const valid = k < n;
  int64_t const i = q.off + (valid ? k : n - 1);
  double const X[3] = {bt.x[i], bt.y[i], bt.z[i]};
  double const B[3] = {bt.bx[i], bt.by[i], bt.bz[i]};
  double const A[3] = {bt.ax[i], bt.ay[i], bt.az[i]};
  double const gsc = dloss[r] * 2.0 * scale / 1e6;
  double gX[3] = {0, 0, 0}, gB[3] = {0, 0, 0}, gA[3] = {0, 0, 0};
  for (int j0 = 0; j0 < ns; j0 += kAdjTile) {
    int const cnt = (ns - j0) < kAdjTile ? ns - j0 : kAdjTile;
    __syncthreads();
    for (int t = tid; t < cnt; t += kBlock) {
      int const j = j0 + t;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        so[c][t] = obs[c * ns + j];
        so[3 + c][t] = nrm[c * ns + j];
      }
      size_t const rj = (size_t)r * ns + j;
      double const sgn = (!directional && raw[rj] < 0) ? -1.0 : 1.0;
      sg[t] = gsc * sgn * (pd[rj] - target[j]);
    }
    __syncthreads();
    if (transverse_beam(B[0], B[1])) {  // this lane's point does not travel along z
      for (int t = 0; t < cnt; ++t)
        pd_term_adjoint<true>(sg[t], so[0][t], so[1][t], so[2][t], so[3][t], so[4][t], so[5][t], X, B, A, directional,
                              gX, gB, gA);
    } else {
      for (int t = 0; t < cnt; ++t)
        pd_term_adjoint(sg[t], so[0][t], so[1][t], so[2][t], so[3][t], so[4][t], so[5][t], X, B, A, directional, gX,
                        gB, gA);
    }
  }
  if (valid) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      gy[(2 * c) * total + i] = gX[c];
      gy[(2 * c + 1) * total + i] = gB[c] * kInvC;
      ga[c * total + i] = gA[c];
    }
  }
}

// One RK4 step from y recording the four stages (same arithmetic as RK4Ctx::step).
template <int NU>
__device__ __forceinline__ void step_record(RK4Ctx<NU> const& R, double* y, double h, double hh, double h6,
                                            double* __restrict__ rec, int np) {
  double k1[6], k2[6], k3[6], k4[6], yt[6];
  FieldJet J;
  auto put = [&](int s, const double* ys) {
    double v[kStageRec] = {ys[1], ys[3], ys[5], J.B[0], J.B[1], J.B[2], J.dBdD[0], J.dBdD[1], J.dBdD[2], J.sp,
                           (double)J.p};
#pragma unroll
    for (int f = 0; f < kStageRec; ++f) rec[(size_t)(s * kStageRec + f) * np] = v[f];
  };
  R.F.jet(y[0], y[2], y[4], J);
  R.derivs_from(y, J.B, k1);
  put(0, y);
#pragma unroll
  for (int c = 0; c < 6; ++c) yt[c] = y[c] + hh * k1[c];
  R.F.jet(yt[0], yt[2], yt[4], J);
  R.derivs_from(yt, J.B, k2);
  put(1, yt);
#pragma unroll
  for (int c = 0; c < 6; ++c) yt[c] = y[c] + hh * k2[c];
  R.F.jet(yt[0], yt[2], yt[4], J);
  R.derivs_from(yt, J.B, k3);
  put(2, yt);
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    yt[c] = y[c] + h * k3[c];
    k3[c] += k2[c];
  }
  R.F.jet(yt[0], yt[2], yt[4], J);
  R.derivs_from(yt, J.B, k4);
  put(3, yt);
#pragma unroll
  for (int c = 0; c < 6; ++c) y[c] = y[c] + h6 * (k1[c] + k4[c] + 2.0 * k3[c]);
}

struct PoleAcc {
  double* out;  // this lane's gradient entries, pole p at out[p * stride]
  double acc;
  int cur;
  int stride;
  __device__ __forceinline__ void add(int p, double v) {
    if (p != cur) {
      if (cur >= 0) out[(size_t)cur * stride] += acc;
      cur = p;
      acc = 0.0;
    }
    acc += v;
  }
  __device__ __forceinline__ void flush() {
    if (cur >= 0) out[(size_t)cur * stride] += acc;
    cur = -1;
    acc = 0.0;
  }
};

// Transposed RK4 step: lam = dL/dy' in, dL/dy out; dL/d delta collected in pa.  gk1 =
// direct cotangent of the first stage's acceleration (dL/dk1 for k1[1], k1[3], k1[5]).
__device__ __forceinline__ void step_adjoint(const double* __restrict__ st, double* lam, double h, double qmg,
                                             const double* u, const double* fpole, const double* gk1,
                                             PoleAcc& pa) {
  double const hh = h * 0.5, h6 = div_rn(h, 6.0, kInvSix);
  double kb[4][6], ybar[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    ybar[c] = lam[c];
    kb[0][c] = h6 * lam[c];
    kb[3][c] = h6 * lam[c];
    kb[1][c] = 2.0 * h6 * lam[c];
    kb[2][c] = 2.0 * h6 * lam[c];
  }
  kb[0][1] += gk1[0];
  kb[0][3] += gk1[1];
  kb[0][5] += gk1[2];
  double const back[4] = {0.0, hh, hh, h};  // stage s state = y + back[s] * k_{s-1}
#pragma unroll
  for (int s = 3; s >= 0; --s) {
    const double* S = st + s * kStageRec;
    double const* w = kb[s];
    double const wa[3] = {w[1], w[3], w[5]};
    double const v[3] = {S[0], S[1], S[2]};
    double const B[3] = {S[3], S[4], S[5]};
    // c = qmg (w_a x v)
    double const c[3] = {qmg * (wa[1] * v[2] - wa[2] * v[1]), qmg * (wa[2] * v[0] - wa[0] * v[2]),
                         qmg * (wa[0] * v[1] - wa[1] * v[0])};
    double const sx = c[0] * S[6] + c[1] * S[7] + c[2] * S[8];
    double ys[6];
    ys[0] = u[0] * sx;
    ys[2] = u[1] * sx;
    ys[4] = u[2] * sx;
    ys[1] = w[0] + qmg * (B[1] * wa[2] - B[2] * wa[1]);
    ys[3] = w[2] + qmg * (B[2] * wa[0] - B[0] * wa[2]);
    ys[5] = w[4] + qmg * (B[0] * wa[1] - B[1] * wa[0]);
    int const p = (int)S[10];
    if (p >= 0) pa.add(p, S[9] * (c[0] * fpole[0] + c[1] * fpole[1] + c[2] * fpole[2]));
#pragma unroll
    for (int q = 0; q < 6; ++q) ybar[q] += ys[q];
    if (s > 0) {
#pragma unroll
      for (int q = 0; q < 6; ++q) kb[s - 1][q] += back[s] * ys[q];
    }
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) lam[c] = ybar[c];
}

// Recompute checkpoint segments [sa, sb) of one RK4 leg for every realisation in
// parallel, one lane per (segment, realisation), writing each step's stage records to
// rec[((i - sa*C) * kStepRec + f) * np + r] (consecutive lanes = consecutive
// realisations, so the stores coalesce).  The recomputed steps are bitwise those of
// the forward pass (same arithmetic from the same checkpointed state).
template <int NU>
__global__ __launch_bounds__(64) void rk4_record_kernel(const DevUndulator* __restrict__ U, int nu,
                                                         const double* __restrict__ poles, int npl,
                                                         const BatchParticle* __restrict__ Ps, int np, Checkpoints ck,
                                                         int backward_leg, int64_t sa, int64_t sb,
                                                         const double* __restrict__ gy,
                                                         const double* __restrict__ ga, int64_t total,
                                                         double* __restrict__ rec) {
  extern __shared__ double spole[];  // npl x 64 (launch_rk4_adjoint)
  int64_t const lane = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (lane >= (sb - sa) * np) return;
  int const r = (int)(lane % np);
  int64_t const seg = sa + lane / np;
  BatchParticle const q = Ps[r];
  int ps;
  const double* const pt = stage_poles(poles + q.pole_base, spole, npl, &ps);
  RK4Ctx<NU> const R{Fields<NU>(U, nu, pt, ps), q.qmg};
  int64_t const C = ck.every;
  int64_t const nleg = backward_leg ? q.nb : q.nf;
  int64_t const slot = backward_leg ? (q.nf + C - 1) / C + seg : seg;
  double const h = backward_leg ? -q.dt : q.dt, hh = h * 0.5, h6 = div_rn(h, 6.0, kInvSix);
  double y[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) y[c] = ck.state[(r * ck.stride + slot) * 6 + c];
  int64_t const i0 = seg * C, i1 = (i0 + C < nleg) ? i0 + C : nleg;
  for (int64_t i = i0; i < i1; ++i) {
    double* const dst = rec + (size_t)(i - sa * C) * kStepRec * np + r;
    step_record(R, y, h, hh, h6, dst, np);
    // forward leg: step i leads from the state of point nb + i; backward leg: to that of point nb - 1 - i
    int64_t const pt = q.off + (backward_leg ? q.nb - 1 - i : q.nb + i);
#pragma unroll
    for (int c = 0; c < 6; ++c) dst[(size_t)(4 * kStageRec + c) * np] = gy[c * total + pt];
    // the acceleration k1/c of step i is stored with point nb + i + 1 (forward leg, none
    // for the last step) or with the point the step leads to (backward leg)
    bool const has_a = ga && (backward_leg || i + 1 < q.nf);
    int64_t const pa = backward_leg ? pt : pt + 1;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      dst[(size_t)(4 * kStageRec + 6 + c) * np] = has_a ? ga[c * total + pa] * kInvC : 0.0;
  }
}

// The reverse sweep is linear in the cotangent it starts from, so every segment is
// swept independently and in parallel: sweep v = 0 starts from 0 and adds the points'
// direct gradients (particular part), sweeps v = 1..6 start from the unit cotangent
// e_{v-1} with no direct terms (the segment's transposed transfer matrix, and its pole
// sensitivities).  Lane (segment, v, realisation); consecutive lanes = consecutive
// realisations, so the record loads coalesce.  Outputs per lane: the cotangent at the
// segment start, lam_out[lane*6 + c], and dL/d delta, del_out[((lane/np)*npoles + p)*np + r].
__global__ __launch_bounds__(64) void rk4_segment_sweep_kernel(const DevUndulator* __restrict__ U, int poled,
                                                                const BatchParticle* __restrict__ Ps, int np, int every,
                                                                int backward_leg, int64_t sa, int64_t sb,
                                                                const double* __restrict__ rec,
                                                                double* __restrict__ lam_out,
                                                                double* __restrict__ del_out, int npoles) {
  int64_t const lane = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (lane >= (sb - sa) * kSweeps * np) return;
  int const r = (int)(lane % np);
  int const v = (int)((lane / np) % kSweeps);
  int64_t const seg = sa + lane / ((int64_t)np * kSweeps);
  BatchParticle const q = Ps[r];
  double const u[3] = {U[0].u[0], U[0].u[1], U[0].u[2]};
  double const fpole[3] = {U[poled].field[0], U[poled].field[1], U[poled].field[2]};
  PoleAcc pa{del_out + (size_t)(lane / np) * npoles * np + r, 0.0, -1, np};
  int64_t const C = every;
  int64_t const nleg = backward_leg ? q.nb : q.nf;
  double const h = backward_leg ? -q.dt : q.dt;
  int64_t const i0 = seg * C, i1 = (i0 + C < nleg) ? i0 + C : nleg;
  double lam[6] = {0, 0, 0, 0, 0, 0};
  if (v > 0) lam[v - 1] = 1.0;
  double const gdir = v == 0 ? 1.0 : 0.0;
  const double* const base = rec + (size_t)(i0 - sa * C) * kStepRec * np + r;
  double buf[2][kStepRec];
  auto load = [&](int64_t i, double* st) {
    const double* src = base + (size_t)(i - i0) * kStepRec * np;
#pragma unroll
    for (int f = 0; f < kStepRec; ++f) st[f] = src[(size_t)f * np];
  };
  load(i1 - 1, buf[0]);
  if (i1 - 2 >= i0) load(i1 - 2, buf[1]);
  for (int64_t i = i1 - 1; i >= i0; i -= 2) {
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      int64_t const ii = i - d;
      if (ii >= i0) {
        const double* g = buf[d] + 4 * kStageRec;
        double const gk1[3] = {gdir * g[6], gdir * g[7], gdir * g[8]};
        if (backward_leg) {
#pragma unroll
          for (int c = 0; c < 6; ++c) lam[c] = __builtin_fma(gdir, g[c], lam[c]);
          step_adjoint(buf[d], lam, h, q.qmg, u, fpole, gk1, pa);
        } else {
          step_adjoint(buf[d], lam, h, q.qmg, u, fpole, gk1, pa);
#pragma unroll
          for (int c = 0; c < 6; ++c) lam[c] = __builtin_fma(gdir, g[c], lam[c]);
        }
        if (ii - 2 >= i0) load(ii - 2, buf[d]);
      }
    }
  }
  pa.flush();
#pragma unroll
  for (int c = 0; c < 6; ++c) lam_out[(size_t)lane * 6 + c] = lam[c];
}

// Chains the segments of one pass for each realisation, last segment first:
// lam_start = L0 + sum_v lam_end[v] L_{v+1},  grad += D0 + sum_v lam_end[v] D_{v+1},
// lam_end of the previous segment = lam_start (lam_buf carries it between passes).
__global__ __launch_bounds__(64) void rk4_combine_kernel(int np, int64_t sa, int64_t sb,
                                                          const double* __restrict__ lam_out,
                                                          const double* __restrict__ del_out,
                                                          double* __restrict__ lam_buf, double* __restrict__ grad,
                                                          int npoles) {
  int const r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= np) return;
  double le[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) le[c] = lam_buf[(size_t)r * 6 + c];
  for (int64_t seg = sb - 1; seg >= sa; --seg) {
    auto row = [&](int v) { return ((seg - sa) * kSweeps + v) * (int64_t)np + r; };
    double ls[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) ls[c] = lam_out[(size_t)row(0) * 6 + c];
#pragma unroll
    for (int v = 1; v < kSweeps; ++v)
#pragma unroll
      for (int c = 0; c < 6; ++c) ls[c] = __builtin_fma(le[v - 1], lam_out[(size_t)row(v) * 6 + c], ls[c]);
    auto drow = [&](int v, int p) { return (((seg - sa) * kSweeps + v) * (int64_t)npoles + p) * np + r; };
    for (int p = 0; p < npoles; ++p) {
      double gp = del_out[drow(0, p)];
#pragma unroll
      for (int v = 1; v < kSweeps; ++v) gp = __builtin_fma(le[v - 1], del_out[drow(v, p)], gp);
      grad[(size_t)r * npoles + p] += gp;
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) le[c] = ls[c];
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) lam_buf[(size_t)r * 6 + c] = le[c];
}

}  // namespace

hipError_t launch_zero(double* a, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  zero_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(a, n);
  return hipGetLastError();
}

hipError_t launch_loss(const double* flux_pp, const double* target, int ns, int np, double* loss, hipStream_t s) {
  loss_kernel<<<np, 256, 0, s>>>(flux_pp, target, ns, loss);
  return hipGetLastError();
}

hipError_t launch_flux_adjoint(BatchTraj const& bt, const BatchParticle* P, int np, int64_t nmax, const double* obs,
                               int ns, double omega, double cw, double c0w, double c2, const double* field_pp,
                               const double* flux_pp, const double* target, const double* dloss, double* gy,
                               int64_t total, bool mixed, AdjScreen const& sc, hipStream_t s) {
  dim3 const grid((unsigned)((nmax + kBlock - 1) / kBlock), (unsigned)np);
  if (mixed)
    flux_adjoint_mixed_kernel<<<grid, kBlock, 0, s>>>(bt, P, obs, ns, omega, cw, c0w, c2, field_pp, flux_pp, target,
                                                      dloss, gy, total, sc);
  else
    flux_adjoint_kernel<<<grid, kBlock, 0, s>>>(bt, P, obs, ns, omega, cw, c0w, c2, field_pp, flux_pp, target, dloss,
                                                gy, total);
  return hipGetLastError();
}

hipError_t launch_pd_adjoint(BatchTraj const& bt, const BatchParticle* P, int np, int64_t nmax, const double* obs,
                             const double* nrm, int ns, int directional, double scale, const double* raw,
                             const double* pd, const double* target, const double* dloss, double* gy, double* ga,
                             int64_t total, hipStream_t s) {
  dim3 const grid((unsigned)((nmax + kBlock - 1) / kBlock), (unsigned)np);
  pd_adjoint_kernel<<<grid, kBlock, 0, s>>>(bt, P, obs, nrm, ns, directional, scale, raw, pd, target, dloss, gy, ga,
                                            total);
  return hipGetLastError();
}

size_t adjoint_record_doubles(int64_t steps, int np) { return (size_t)steps * kStepRec * np; }
size_t adjoint_sweep_doubles(int64_t segments, int np, int npoles) {
  return (size_t)segments * kSweeps * np * (6 + npoles);
}

hipError_t launch_rk4_adjoint(const DevUndulator* U, int nu, int poled, const double* poles, const BatchParticle* P,
                              int np, int64_t nf, int64_t nb, Checkpoints const& ck, int64_t seg_per_pass,
                              const double* gy, const double* ga, int64_t total, double* rec, double* sweep,
                              double* lam_buf, double* grad, int npoles, hipStream_t s) {
  int64_t const C = ck.every;
  double* const lam_out = sweep;
  double* const del_out = sweep + (size_t)seg_per_pass * kSweeps * np * 6;
  for (int leg = 0; leg < 2; ++leg) {
    int64_t const nleg = leg ? nb : nf;
    if (nleg <= 0) continue;
    int64_t const nseg = (nleg + C - 1) / C;
    hipError_t e = launch_zero(lam_buf, (int64_t)np * 6, s);
    if (e != hipSuccess) return e;
    for (int64_t sb = nseg; sb > 0; sb -= seg_per_pass) {
      int64_t const sa = sb - seg_per_pass > 0 ? sb - seg_per_pass : 0;
      int64_t const lanes = (sb - sa) * np;
      dim3 const grid((unsigned)((lanes + 63) / 64));
      int const npl = pole_lds_count(npoles);
      size_t const lds = (size_t)npl * kPoleLane * sizeof(double);
      if (nu == 1)
        rk4_record_kernel<1><<<grid, 64, lds, s>>>(U, nu, poles, npl, P, np, ck, leg, sa, sb, gy, ga, total, rec);
      else if (nu == 2)
        rk4_record_kernel<2><<<grid, 64, lds, s>>>(U, nu, poles, npl, P, np, ck, leg, sa, sb, gy, ga, total, rec);
      else
        rk4_record_kernel<0><<<grid, 64, lds, s>>>(U, nu, poles, npl, P, np, ck, leg, sa, sb, gy, ga, total, rec);
      e = hipGetLastError();
      if (e != hipSuccess) return e;
      e = launch_zero(del_out, (sb - sa) * kSweeps * np * npoles, s);
      if (e != hipSuccess) return e;
      int64_t const slanes = lanes * kSweeps;
      rk4_segment_sweep_kernel<<<(unsigned)((slanes + 63) / 64), 64, 0, s>>>(U, poled, P, np, ck.every, leg, sa, sb,
                                                                             rec, lam_out, del_out, npoles);
      e = hipGetLastError();
      if (e != hipSuccess) return e;
      rk4_combine_kernel<<<(np + 63) / 64, 64, 0, s>>>(np, sa, sb, lam_out, del_out, lam_buf, grad, npoles);
      e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

}  // namespace srs
