// Batched multi-particle radiation on gfx950 (the reference's nparticles loop,
// OSCARSSR.cc:2465-2479, as one device pipeline):
//
//   rk4_batch_kernel      one lane per particle integrates its RK4 trajectory
//                         (ragged lengths, packed SoA) in FP64 and also writes the
//                         FP32 copies the mixed form reads;
//   zero_counter_kernel   resets the work-queue head (a kernel, so the pipeline can
//                         be captured in a hipGraph without memset nodes);
//   batch_flux_kernel     persistent: each block pops (particle, observation tile,
//                         energy group) items from the queue until it is empty and
//                         writes that particle's flux for 256 points;
//   particle_reduce       Kahan-sums the per-particle maps in particle order with
//                         the weight 1/NP (T3DScalarContainer::AddToPoint order).
#include "srs_batch.h"
#include "srs_pd.h"
#include "srs_rk4.h"
#include "srs_spectrum.h"
#include "srs_terms.h"

namespace srs {
namespace {

// Same point layout as the host trajectory (OSCARSSR.cc:562-673): forward pass from
// T0 stores the state before each step with the previous derivative (A/c lags one
// point; 0 at the first point), the backward pass fills the nb points before T0.
template <int NU, Force FORCE>
__global__ __launch_bounds__(64) void rk4_batch_kernel(const DevUndulator* __restrict__ U, int nu,
                                                        const double* __restrict__ poles, int npl,
                                                        const BatchParticle* __restrict__ Ps, int np, BatchTraj bt,
                                                        Checkpoints ck, RK4Seg sg) {
  extern __shared__ double spole[];  // npl x 64 (launch_rk4_batch)
  int const p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  BatchParticle const q = Ps[p];
  int ps;
  const double* const pt = stage_poles(poles + q.pole_base, spole, npl, &ps);
  RK4Ctx<NU, FORCE> const R{Fields<NU>(U, nu, pt, ps), q.qmg};
  double x[6], d[6] = {0, 0, 0, 0, 0, 0};
  auto init = [&]() {
    x[0] = q.x0[0];
    x[1] = q.beta0[0] * kC;
    x[2] = q.x0[1];
    x[3] = q.beta0[1] * kC;
    x[4] = q.x0[2];
    x[5] = q.beta0[2] * kC;
#pragma unroll
    for (int c = 0; c < 6; ++c) d[c] = 0.0;
  };
  auto load = [&](int leg) {
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      x[c] = sg.state[((size_t)p * 2 + leg) * 12 + c];
      d[c] = sg.state[((size_t)p * 2 + leg) * 12 + 6 + c];
    }
  };
  auto keep = [&](int leg) {
    if (!sg.state) return;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      sg.state[((size_t)p * 2 + leg) * 12 + c] = x[c];
      sg.state[((size_t)p * 2 + leg) * 12 + 6 + c] = d[c];
    }
  };
  auto save = [&](int64_t slot) {
#pragma unroll
    for (int c = 0; c < 6; ++c) ck.state[(p * ck.stride + slot) * 6 + c] = x[c];
  };
  int64_t const nf_slots = ck.state ? (q.nf + ck.every - 1) / ck.every : 0;
  bool tr = false;  // a point of this segment does not travel along z
  if (sg.lo < q.nf) {
    if (sg.lo == 0) init();
    else load(0);
    double const h = q.dt, hh = h * 0.5, h6 = div_rn(h, 6.0, kInvSix);
    int64_t const i1 = sg.hi < q.nf ? sg.hi : q.nf;
    for (int64_t i = sg.lo; i < i1; ++i) {
      if (ck.state && i % ck.every == 0) save(i / ck.every);
      tr |= store_point(bt, q.off + q.nb + i, x, d);
      R.derivs(x, d);
      R.step(x, d, h, hh, h6);
    }
    keep(0);
  }
  if (sg.lo < q.nb) {
    if (sg.lo == 0) init();
    else load(1);
    double const h = -q.dt, hh = h * 0.5, h6 = div_rn(h, 6.0, kInvSix);
    int64_t const i1 = sg.hi < q.nb ? sg.hi : q.nb;
    for (int64_t i = sg.lo; i < i1; ++i) {
      if (ck.state && i % ck.every == 0) save(nf_slots + i / ck.every);
      R.derivs(x, d);
      R.step(x, d, h, hh, h6);
      tr |= store_point(bt, q.off + q.nb - 1 - i, x, d);
    }
    keep(1);
  }
  // segments run in order, each after the previous one's flux: the flag covers every
  // point a later kernel of this segment reads
  if (bt.tflag) bt.tflag[p] = (sg.lo == 0 ? 0 : bt.tflag[p]) | (tr ? 1 : 0);
}

__global__ void zero_counter_kernel(unsigned* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0u;
}

// LDS of the batched kernel: the FP64 tile, or the mixed form's union.
template <int TILE_, int NE, bool MIXED>
struct BatchTile {
  FluxTile<TILE_> raw;
};
template <int TILE_, int NE>
struct BatchTile<TILE_, NE, true> {
  union {
    FluxTile<TILE_> raw;
    ParaTile<TILE_, NE> para;
  };
};

// Persistent flux kernel.  Item = (particle, observation tile of 256 points, group of
// NE energies), popped from a global counter by lane 0 and broadcast through LDS.
// Every block exits when the counter passes the item count.  The mixed form takes
// the paraxial terms for every trajectory tile that passes their bounds.
// Trajectory points of one item within steps [lo, hi) of both RK4 legs, as local
// index ranges: backward-leg steps i hold points nb - 1 - i, forward-leg steps i hold
// points nb + i.
struct ItemRanges {
  int64_t r0[2], r1[2];
  __device__ __forceinline__ ItemRanges(BatchParticle const& q, FluxSeg const& sg) {
    int64_t const bl = q.nb < sg.lo ? q.nb : sg.lo, bh = q.nb < sg.hi ? q.nb : sg.hi;
    int64_t const fl = q.nf < sg.lo ? q.nf : sg.lo, fh = q.nf < sg.hi ? q.nf : sg.hi;
    r0[0] = q.nb - bh;
    r1[0] = q.nb - bl;
    r0[1] = q.nb + fl;
    r1[1] = q.nb + fh;
  }
};

template <int NE, int TILE_, bool MIXED>
__global__ __launch_bounds__(kBlock) void batch_flux_kernel(BatchTraj bt, const BatchParticle* __restrict__ P,
                                                            const double* __restrict__ obs, int ns, BatchEnergies be,
                                                            unsigned nitems, unsigned* __restrict__ counter,
                                                            double* __restrict__ flux_pp, FluxSeg sg) {
  __shared__ BatchTile<TILE_, NE, MIXED> T;
  __shared__ double red[4 * (kBlock / 64)];
  __shared__ unsigned s_item;
  int const tid = threadIdx.x;
  int const ntiles = (ns + kBlock - 1) / kBlock;
  int const ngroups = (be.ne + NE - 1) / NE;
  for (;;) {
    __syncthreads();  // every lane has finished with the previous item and s_item
    if (tid == 0) s_item = atomicAdd(counter, 1u);
    __syncthreads();
    unsigned const item = s_item;
    if (item >= nitems) break;
    int const g = (int)(item % (unsigned)ngroups);
    unsigned const rest = item / (unsigned)ngroups;
    int const tile = (int)(rest % (unsigned)ntiles);
    int const p = (int)(rest / (unsigned)ntiles);
    BatchParticle const q = P[p];
    ItemRanges const rg(q, sg);
    EnergySet es;
    es.ne = NE;
    float cwf[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      int const ee = (g * NE + e < be.ne) ? g * NE + e : be.ne - 1;
      es.omega[e] = be.omega[ee];
      es.cw[e] = be.cw[ee];
      es.kw[e] = (q.c0 * be.omega[ee]) * q.dt;  // (C0*omega)*dt with the particle's C0, OSCARSSR.cc:2080
      es.fr[e] = be.fr[ee];
      es.fw[e] = be.fw[ee];
      cwf[e] = (float)be.cw[ee];
    }
    int const j0 = tile * kBlock;
    int const j = j0 + tid;
    int const jj = j < ns ? j : j0;  // idle lanes repeat the tile's first point
    double const ox = obs[jj], oy = obs[ns + jj], oz = obs[2 * ns + jj];
    // this segment's six field sums S(e, q6) (+ the earlier segments', in segment order);
    // turn: paraxial sums, which take the phase omega Z/c of their plane z = turn_z
    bool turn = false;
    double turn_z = 0.0;
    auto finish = [&](auto sum_of) {
      if (j >= ns) return;
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        int const ee = g * NE + e;
        if (ee >= be.ne) continue;
        double S[6];
#pragma unroll
        for (int q6 = 0; q6 < 6; ++q6) S[q6] = sum_of(e, q6);
        if (turn) {
          double tc, ts;
          para_plane_turn(es.omega[e], turn_z, &tc, &ts);
#pragma unroll
          for (int ax = 0; ax < 3; ++ax) {
            double const re = S[2 * ax], im = S[2 * ax + 1];
            S[2 * ax] = __builtin_fma(re, tc, -(im * ts));
            S[2 * ax + 1] = __builtin_fma(re, ts, im * tc);
          }
        }
        size_t const b = ((size_t)p * be.ne + ee) * 6 * ns + j;
        if (!sg.first) {
#pragma unroll
          for (int q6 = 0; q6 < 6; ++q6) S[q6] = sg.field_acc[b + (size_t)q6 * ns] + S[q6];
        }
        if (sg.field_acc && !(sg.first && sg.final && !sg.keep_field)) {
#pragma unroll
          for (int q6 = 0; q6 < 6; ++q6) sg.field_acc[b + (size_t)q6 * ns] = S[q6];
        }
        if (sg.final) flux_pp[((size_t)p * be.ne + ee) * ns + j] = flux_from_sums(S[0], S[1], S[2], S[3], S[4], S[5],
                                                                                    es.kw[e], q.c2);
      }
    };
    if (MIXED && bt.tflag && bt.tflag[p]) {
      // a point of this particle does not travel along z, which the mixed terms assume
      // (they stage 1 - beta_z): the item sums the FP64 terms instead
      AccFP64<NE> d;
      d.zero();
      for (int part = 0; part < 2; ++part) {
        int64_t const lo = rg.r0[part], hi = rg.r1[part];
        for (int64_t base = lo; base < hi; base += TILE_) {
          int const cnt = (hi - base) < TILE_ ? (int)(hi - base) : TILE_;
          __syncthreads();
          for (int k = tid; k < cnt; k += kBlock) {
            int64_t const i = q.off + base + k;
            T.raw.t[k] = q.dt * (double)(base + k);
            T.raw.x[k] = bt.x[i];
            T.raw.y[k] = bt.y[i];
            T.raw.z[k] = bt.z[i];
            T.raw.bx[k] = bt.bx[i];
            T.raw.by[k] = bt.by[i];
            T.raw.bz[k] = bt.bz[i];
          }
          __syncthreads();
          flux_terms_fp64<NE, TILE_>(T.raw, cnt, ox, oy, oz, es, d);
        }
      }
      finish([&](int e, int q6) { return (q6 & 1) ? d.im[e][q6 >> 1] : d.re[e][q6 >> 1]; });
      continue;
    }
    AccFP64<NE> a64;
    AccMixed<NE> a32;
    bool para = false;
    ParaBlock pb{};
    if constexpr (MIXED) {
      a32.zero();
      pb = para_block(ox, oy, oz, obs[2 * ns + j0], red);
      bool const p0 = para_check(pb, bt.x + q.off, bt.y + q.off, bt.z + q.off, rg.r0[0], rg.r1[0], es);
      bool const p1 = para_check(pb, bt.x + q.off, bt.y + q.off, bt.z + q.off, rg.r0[1], rg.r1[1], es);
      para = p0 && p1;
    } else {
      a64.zero();
    }
    for (int part = 0; part < 2; ++part) {
      int64_t const lo = rg.r0[part], hi = rg.r1[part];
      if (para) {
        if constexpr (MIXED) {
          for (int64_t base = lo; base < hi; base += TILE_) {
            int const cnt = (hi - base) < TILE_ ? (int)(hi - base) : TILE_;
            __syncthreads();
            int const cntp = (cnt + kSub - 1) / kSub * kSub;
            for (int k = tid; k < cntp; k += kBlock) {
              int64_t const i = q.off + base + k;
              if (k < cnt)
                para_stage<TILE_, NE>(T.para, k, bt.x[i], bt.y[i], bt.z[i], bt.fbx[i], bt.fby[i], bt.fomb[i],
                                      base + k, q.dt, pb, es);
              else
                para_pad<TILE_, NE>(T.para, k);
            }
            __syncthreads();
            flux_terms_para<NE, TILE_>(T.para, cntp, ox, oy, es, cwf, a32);
          }
        }
      } else {
        for (int64_t base = lo; base < hi; base += TILE_) {
          int const cnt = (hi - base) < TILE_ ? (int)(hi - base) : TILE_;
          __syncthreads();
          for (int k = tid; k < cnt; k += kBlock) {
            int64_t const i = q.off + base + k;
            T.raw.t[k] = q.dt * (double)(base + k);
            T.raw.x[k] = bt.x[i];
            T.raw.y[k] = bt.y[i];
            T.raw.z[k] = bt.z[i];
            if (MIXED) {
              T.raw.fbx[k] = bt.fbx[i];
              T.raw.fby[k] = bt.fby[i];
              T.raw.fomb[k] = bt.fomb[i];
            } else {
              T.raw.bx[k] = bt.bx[i];
              T.raw.by[k] = bt.by[i];
              T.raw.bz[k] = bt.bz[i];
            }
          }
          __syncthreads();
          if (MIXED) flux_terms_mixed<NE, TILE_>(T.raw, cnt, ox, oy, oz, es, cwf, a32);
          else flux_terms_fp64<NE, TILE_>(T.raw, cnt, ox, oy, oz, es, a64);
        }
      }
    }
    if constexpr (MIXED) {
      turn = para;  // paraxial sums take the plane's phase omega Z/c
      turn_z = pb.z;
      finish([&](int e, int q6) { return a32.value(e, q6); });
    } else {
      finish([&](int e, int q6) { return (q6 & 1) ? a64.im[e][q6 >> 1] : a64.re[e][q6 >> 1]; });
    }
  }
}

// The mixed persistent kernel at one energy per item with two observation points per
// lane: item = (particle, observation tile of 512 points, energy), lane tid owning points
// j0 + tid and j0 + 256 + tid, so each staged paraxial record serves both
// (flux_terms_para2).  Particles flagged as not travelling along z take the FP64 terms
// and items that fail the paraxial bounds the general mixed terms, one point after the
// other.  Same queue, segments and outputs as batch_flux_kernel.  Held to 4 waves per
// SIMD (128 VGPRs): the fallback paths' extra registers spill, but only outside the
// integration loops (per tile and per item), which the compiler left without spills.
template <int TILE_>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void batch_flux2_kernel(
    BatchTraj bt, const BatchParticle* __restrict__ P, const double* __restrict__ obs, int ns, BatchEnergies be,
    unsigned nitems, unsigned* __restrict__ counter, double* __restrict__ flux_pp, FluxSeg sg) {
  __shared__ BatchTile<TILE_, 1, true> T;
  __shared__ double red[4 * (kBlock / 64)];
  __shared__ unsigned s_item;
  int const tid = threadIdx.x;
  int const ntiles = (ns + 2 * kBlock - 1) / (2 * kBlock);
  for (;;) {
    __syncthreads();  // every lane has finished with the previous item and s_item
    if (tid == 0) s_item = atomicAdd(counter, 1u);
    __syncthreads();
    unsigned const item = s_item;
    if (item >= nitems) break;
    int const ee = (int)(item % (unsigned)be.ne);
    unsigned const rest = item / (unsigned)be.ne;
    int const tile = (int)(rest % (unsigned)ntiles);
    int const p = (int)(rest / (unsigned)ntiles);
    BatchParticle const q = P[p];
    ItemRanges const rg(q, sg);
    EnergySet es;
    es.ne = 1;
    es.omega[0] = be.omega[ee];
    es.cw[0] = be.cw[ee];
    es.kw[0] = (q.c0 * be.omega[ee]) * q.dt;  // (C0*omega)*dt with the particle's C0, OSCARSSR.cc:2080
    es.fr[0] = be.fr[ee];
    es.fw[0] = be.fw[ee];
    float cwf = (float)be.cw[ee];
    int const j0 = tile * (2 * kBlock);
    int const jp[2] = {j0 + tid, j0 + kBlock + tid};
    double ox[2], oy[2], oz[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      int const jj = jp[k] < ns ? jp[k] : j0;  // idle lanes repeat the tile's first point
      ox[k] = obs[jj];
      oy[k] = obs[ns + jj];
      oz[k] = obs[2 * ns + jj];
    }
    // this segment's six field sums S(q6) of point j (+ the earlier segments', in segment order)
    auto finish = [&](int j, auto sum_of) {
      if (j >= ns) return;
      double S[6];
#pragma unroll
      for (int q6 = 0; q6 < 6; ++q6) S[q6] = sum_of(q6);
      size_t const b = ((size_t)p * be.ne + ee) * 6 * ns + j;
      if (!sg.first) {
#pragma unroll
        for (int q6 = 0; q6 < 6; ++q6) S[q6] = sg.field_acc[b + (size_t)q6 * ns] + S[q6];
      }
      if (sg.field_acc && !(sg.first && sg.final && !sg.keep_field)) {
#pragma unroll
        for (int q6 = 0; q6 < 6; ++q6) sg.field_acc[b + (size_t)q6 * ns] = S[q6];
      }
      if (sg.final)
        flux_pp[((size_t)p * be.ne + ee) * ns + j] = flux_from_sums(S[0], S[1], S[2], S[3], S[4], S[5], es.kw[0], q.c2);
    };
    // a point of this particle does not travel along z, which the mixed terms assume
    // (they stage 1 - beta_z): such items take the FP64 terms (block-uniform)
    bool const flagged = bt.tflag && bt.tflag[p];
    bool para = false;
    if (!flagged) {
      double const z0 = obs[2 * ns + j0];
      ParaBlock const pb = para_block_range(__builtin_fmin(ox[0], ox[1]), __builtin_fmax(ox[0], ox[1]),
                                            __builtin_fmin(oy[0], oy[1]), __builtin_fmax(oy[0], oy[1]),
                                            oz[0] == z0 && oz[1] == z0, z0, red);
      bool const p0 = para_check(pb, bt.x + q.off, bt.y + q.off, bt.z + q.off, rg.r0[0], rg.r1[0], es);
      bool const p1 = para_check(pb, bt.x + q.off, bt.y + q.off, bt.z + q.off, rg.r0[1], rg.r1[1], es);
      para = p0 && p1;
      if (para) {
        AccMixed<1> a[2];
        a[0].zero();
        a[1].zero();
        for (int part = 0; part < 2; ++part) {
          int64_t const lo = rg.r0[part], hi = rg.r1[part];
          for (int64_t base = lo; base < hi; base += TILE_) {
            int const cnt = (hi - base) < TILE_ ? (int)(hi - base) : TILE_;
            int const cntp = (cnt + kSub - 1) / kSub * kSub;
            __syncthreads();
            for (int kk = tid; kk < cntp; kk += kBlock) {
              int64_t const i = q.off + base + kk;
              if (kk < cnt)
                para_stage<TILE_, 1>(T.para, kk, bt.x[i], bt.y[i], bt.z[i], bt.fbx[i], bt.fby[i], bt.fomb[i],
                                     base + kk, q.dt, pb, es);
              else
                para_pad<TILE_, 1>(T.para, kk);
            }
            __syncthreads();
            flux_terms_para2<TILE_>(T.para, cntp, ox, oy, cwf, a);
          }
        }
        double tc, ts;  // paraxial sums take the plane's phase omega Z/c
        para_plane_turn(es.omega[0], pb.z, &tc, &ts);
        finish(jp[0], [&](int q6) { return turned([&](int k) { return a[0].value(0, k); }, q6, tc, ts); });
        finish(jp[1], [&](int q6) { return turned([&](int k) { return a[1].value(0, k); }, q6, tc, ts); });
        continue;
      }
    }
    // otherwise one observation point after the other: the FP64 terms (flagged) or the
    // general mixed terms (outside the paraxial bounds), the tile staged for each point;
    // coordinates re-read per point (fewer live registers in the kernel)
#pragma unroll 1
    for (int k = 0; k < 2; ++k) {
      int const j = k ? jp[1] : jp[0];
      int const jj = j < ns ? j : j0;
      double const px = obs[jj], py = obs[ns + jj], pz = obs[2 * ns + jj];
      if (flagged) {
        AccFP64<1> d;
        d.zero();
        for (int part = 0; part < 2; ++part) {
          int64_t const lo = rg.r0[part], hi = rg.r1[part];
          for (int64_t base = lo; base < hi; base += TILE_) {
            int const cnt = (hi - base) < TILE_ ? (int)(hi - base) : TILE_;
            __syncthreads();
            for (int kk = tid; kk < cnt; kk += kBlock) {
              int64_t const i = q.off + base + kk;
              T.raw.t[kk] = q.dt * (double)(base + kk);
              T.raw.x[kk] = bt.x[i];
              T.raw.y[kk] = bt.y[i];
              T.raw.z[kk] = bt.z[i];
              T.raw.bx[kk] = bt.bx[i];
              T.raw.by[kk] = bt.by[i];
              T.raw.bz[kk] = bt.bz[i];
            }
            __syncthreads();
            flux_terms_fp64<1, TILE_>(T.raw, cnt, px, py, pz, es, d);
          }
        }
        finish(j, [&](int q6) { return (q6 & 1) ? d.im[0][q6 >> 1] : d.re[0][q6 >> 1]; });
      } else {
        AccMixed<1> g;
        g.zero();
        for (int part = 0; part < 2; ++part) {
          int64_t const lo = rg.r0[part], hi = rg.r1[part];
          for (int64_t base = lo; base < hi; base += TILE_) {
            int const cnt = (hi - base) < TILE_ ? (int)(hi - base) : TILE_;
            __syncthreads();
            for (int kk = tid; kk < cnt; kk += kBlock) {
              int64_t const i = q.off + base + kk;
              T.raw.t[kk] = q.dt * (double)(base + kk);
              T.raw.x[kk] = bt.x[i];
              T.raw.y[kk] = bt.y[i];
              T.raw.z[kk] = bt.z[i];
              T.raw.fbx[kk] = bt.fbx[i];
              T.raw.fby[kk] = bt.fby[i];
              T.raw.fomb[kk] = bt.fomb[i];
            }
            __syncthreads();
            flux_terms_mixed<1, TILE_>(T.raw, cnt, px, py, pz, es, &cwf, g);
          }
        }
        finish(j, [&](int q6) { return g.value(0, q6); });
      }
    }
  }
}

__global__ void particle_reduce_kernel(const double* __restrict__ flux_pp, int np, int nout, double weight,
                                       double* __restrict__ out) {
  int const o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nout) return;
  double s = 0, c = 0;
  for (int p = 0; p < np; ++p) {
    double const y = flux_pp[(size_t)p * nout + o] * weight - c;
    double const t = s + y;
    c = (t - s) - y;
    s = t;
  }
  out[o] = s;
}

template <int NE, int TILE_, bool MIXED>
hipError_t launch_batch_t(BatchTraj const& bt, const BatchParticle* P, const double* obs, int ns,
                          BatchEnergies const& be, unsigned nitems, unsigned* counter, double* flux_pp,
                          FluxSeg const& sg, int nblocks, hipStream_t s) {
  batch_flux_kernel<NE, TILE_, MIXED><<<nblocks, kBlock, 0, s>>>(bt, P, obs, ns, be, nitems, counter, flux_pp, sg);
  return hipGetLastError();
}

template <bool MIXED>
hipError_t launch_batch_m(BatchConfig const& cfg, BatchTraj const& bt, const BatchParticle* P, const double* obs, int ns,
                          BatchEnergies const& be, unsigned nitems, unsigned* counter, double* flux_pp,
                          FluxSeg const& sg, int nblocks, hipStream_t s) {
  if (cfg.ppl == 2) {
    if (!MIXED || cfg.ne_per_item != 1) return hipErrorInvalidValue;
    if (cfg.tile == 128) {
      batch_flux2_kernel<128><<<nblocks, kBlock, 0, s>>>(bt, P, obs, ns, be, nitems, counter, flux_pp, sg);
      return hipGetLastError();
    }
    if (cfg.tile == 256) {
      batch_flux2_kernel<256><<<nblocks, kBlock, 0, s>>>(bt, P, obs, ns, be, nitems, counter, flux_pp, sg);
      return hipGetLastError();
    }
    return hipErrorInvalidValue;
  }
  if (cfg.ppl != 1) return hipErrorInvalidValue;
#define SRS_BATCH_CASE(NE_, TL_)                 \
  if (cfg.ne_per_item == NE_ && cfg.tile == TL_) \
    return launch_batch_t<NE_, TL_, MIXED>(bt, P, obs, ns, be, nitems, counter, flux_pp, sg, nblocks, s);
  SRS_BATCH_CASE(1, 64)
  SRS_BATCH_CASE(1, 128)
  SRS_BATCH_CASE(1, 256)
  SRS_BATCH_CASE(2, 128)
  SRS_BATCH_CASE(2, 256)
  SRS_BATCH_CASE(3, 64)
  SRS_BATCH_CASE(3, 128)
  SRS_BATCH_CASE(3, 256)
  SRS_BATCH_CASE(4, 128)
  SRS_BATCH_CASE(4, 256)
#undef SRS_BATCH_CASE
  return hipErrorInvalidValue;
}

// Power density of every realisation over the points of RK4 steps [lo, hi) of both
// legs (backward-leg points [nb - min(hi, nb), nb - min(lo, nb)) then forward-leg
// points [nb + min(lo, nf), nb + min(hi, nf)), taken as one index range split in nch
// chunks): block (observation tile, realisation r, chunk ch), flattened into
// blockIdx.x with the tile fastest; one lane per observation point, trajectory
// points staged in LDS; part[(r * nch + ch) * ns + j] = the chunk's sum of pd_term,
// or += it when add != 0 (time segments after the first, which run in order on one
// stream, so every slot adds its segments in segment order).
// MIXED: the FP32 term (pd_mixed_term), its records staged from the FP64 trajectory per
// tile (1 - |beta|^2 in FP64), 32-term FP32 sub-sums in a Kahan FP32 accumulator; the
// chunk's sum enters the FP64 slot as in the FP64 form.
template <bool MIXED, bool DIR>
__device__ __forceinline__ void pd_batch_mixed_chunk(BatchTraj const& bt, int64_t i0, int64_t i1, int64_t nbs,
                                                     int64_t first_b, int64_t first_f, double ox, double oy, double oz,
                                                     double qx, double qy, double qz, float4* sP, float4* sB,
                                                     float4* sA, double& sum) {
  int const tid = threadIdx.x;
  float const fx = (float)ox, fy = (float)oy, fz = (float)oz, gx = (float)qx, gy = (float)qy, gz = (float)qz;
  KahanF acc{0.f, 0.f};
  for (int64_t base = i0; base < i1; base += kBlock) {
    int const cnt = (i1 - base) < kBlock ? (int)(i1 - base) : kBlock;
    __syncthreads();
    if (tid < cnt) {
      int64_t const v = base + tid;
      int64_t const i = v < nbs ? first_b + v : first_f + v;
      double const bx = bt.bx[i], by = bt.by[i], bz = bt.bz[i];
      double const omb = 1.0 - bz;
      float const g = (float)(omb * (1.0 + bz) - bx * bx - by * by);
      sP[tid] = make_float4((float)bt.x[i], (float)bt.y[i], (float)bt.z[i], g * (kPdS * kPdS));
      sB[tid] = make_float4((float)bx * kPdS, (float)by * kPdS, (float)omb * kPdS, 0.f);
      sA[tid] = make_float4((float)bt.ax[i], (float)bt.ay[i], (float)bt.az[i], 0.f);
    }
    __syncthreads();
    for (int k0 = 0; k0 < cnt; k0 += kSub) {
      int const k1 = (k0 + kSub < cnt) ? k0 + kSub : cnt;
      float sub = 0.f;
      for (int k = k0; k < k1; ++k) sub += pd_mixed_term<DIR>(fx, fy, fz, gx, gy, gz, sP[k], sB[k], sA[k]);
      acc.add(sub);
    }
  }
  sum = ((double)acc.s - (double)acc.c) * 0x1p55;
}

template <bool MIXED>
__global__ __launch_bounds__(kBlock) void pd_batch_kernel(BatchTraj bt, const BatchParticle* __restrict__ P,
                                                          const double* __restrict__ obs,
                                                          const double* __restrict__ nrm, int ns, int directional,
                                                          int64_t lo, int64_t hi, int nch, int ntiles, int add,
                                                          double* __restrict__ part) {
  // the FP64 tile (x, y, z, beta, A/c), or the mixed form's FP32 records
  __shared__ union {
    double d[9][kBlock];
    float4 f[3][kBlock];
  } sm;
  double* const sx = sm.d[0];
  double* const sy = sm.d[1];
  double* const sz = sm.d[2];
  double* const sbx = sm.d[3];
  double* const sby = sm.d[4];
  double* const sbz = sm.d[5];
  double* const sax = sm.d[6];
  double* const say = sm.d[7];
  double* const saz = sm.d[8];
  float4* const fP = sm.f[0];
  float4* const fB = sm.f[1];
  float4* const fA = sm.f[2];
  int const tid = threadIdx.x;
  int const tile = (int)(blockIdx.x % (unsigned)ntiles);
  int const rc = (int)(blockIdx.x / (unsigned)ntiles);
  int const r = rc / nch, ch = rc % nch;
  BatchParticle const q = P[r];
  int64_t const bl = lo < q.nb ? lo : q.nb, bh = hi < q.nb ? hi : q.nb;
  int64_t const fl = lo < q.nf ? lo : q.nf, fh = hi < q.nf ? hi : q.nf;
  int64_t const nbs = bh - bl, n = nbs + (fh - fl);
  int64_t const first_b = q.off + q.nb - bh, first_f = q.off + q.nb + fl - nbs;
  int64_t const chunk = (n + nch - 1) / nch;
  int64_t const i0 = ch * chunk;
  int64_t const i1 = (i0 + chunk < n) ? i0 + chunk : n;
  int const j = tile * kBlock + tid;
  int const jj = j < ns ? j : ns - 1;
  double const ox = obs[jj], oy = obs[ns + jj], oz = obs[2 * ns + jj];
  double const qx = nrm[jj], qy = nrm[ns + jj], qz = nrm[2 * ns + jj];
  double sum = 0.0;
  // a point of this particle does not travel along z, which the mixed terms assume (they
  // stage 1 - beta_z): the FP64 terms below instead (block-uniform: one particle per block)
  bool const mixed = MIXED && !(bt.tflag && bt.tflag[r]);
  if (mixed) {
    if (directional)
      pd_batch_mixed_chunk<MIXED, true>(bt, i0, i1, nbs, first_b, first_f, ox, oy, oz, qx, qy, qz, fP, fB, fA, sum);
    else
      pd_batch_mixed_chunk<MIXED, false>(bt, i0, i1, nbs, first_b, first_f, ox, oy, oz, qx, qy, qz, fP, fB, fA, sum);
  } else
  for (int64_t base = i0; base < i1; base += kBlock) {
    int const cnt = (i1 - base) < kBlock ? (int)(i1 - base) : kBlock;
    __syncthreads();
    bool tr = false;
    if (tid < cnt) {
      int64_t const v = base + tid;
      int64_t const i = v < nbs ? first_b + v : first_f + v;
      sx[tid] = bt.x[i];
      sy[tid] = bt.y[i];
      sz[tid] = bt.z[i];
      sbx[tid] = bt.bx[i];
      sby[tid] = bt.by[i];
      sbz[tid] = bt.bz[i];
      sax[tid] = bt.ax[i];
      say[tid] = bt.ay[i];
      saz[tid] = bt.az[i];
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
  if (j < ns) {
    double* const dst = part + ((size_t)r * nch + ch) * ns + j;
    *dst = add ? *dst + sum : sum;
  }
}

// Chunks in fixed order, then the reference's prefactor and |.| (OSCARSSR.cc:1563-1573):
// raw[r*ns + j] = the summed integrand, pd[r*ns + j] = the power density.  The
// prefactor is pdq_r * dt_r (pdq_r = |q I|/(16 pi^2 eps0 c) of particle r, formed on the
// host), which rounds as the reference's |qI|/(16 pi^2 eps0 c) * dt.
__global__ void pd_batch_reduce_kernel(const double* __restrict__ part, int nch, int ns, int np,
                                       const BatchParticle* __restrict__ P, int directional,
                                       double* __restrict__ raw, double* __restrict__ pd) {
  int64_t const t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)np * ns) return;
  int64_t const r = t / ns, j = t % ns;
  double s = 0.0;
  for (int c = 0; c < nch; ++c) s += part[((size_t)r * nch + c) * ns + j];
  raw[t] = s;
  s *= P[r].pdq * P[r].dt;
  s /= 1e6;
  if (!directional && s < 0) s = -s;
  pd[t] = s;
}

// Spectrum of every particle at one observation point over the points of RK4 steps
// [lo, hi) of both legs (the same index range and chunking as pd_batch_kernel), on the
// energy-run recurrence of srs_spectrum.h: block (energy group, particle r, chunk ch)
// flattened into blockIdx.x; a group is 64 lanes x RUN energies; the block's 4 waves
// split the chunk in a fixed order, each staging the geometry of 64 points at a time
// in LDS (tau = dt_r it + D/c, (beta - n)/D, n/D^2 in the reference's operation order)
// and adding their FP64 sums in LDS in wave order at the end:
// part[((r * nch + ch) * 6 + q) * ne + e], written or added (add).  *range as in
// spectrum_kernel (srs_spectrum.hip).
constexpr int kSbWaves = 4;
constexpr int kSbFold = 4;

template <bool MIXED, int RUN>
__global__ __launch_bounds__(64 * kSbWaves) void spectrum_batch_kernel(BatchTraj bt, const BatchParticle* __restrict__ P,
                                                                       double ox, double oy, double oz, SpecTables tab,
                                                                       int ne, int64_t lo, int64_t hi, int nch,
                                                                       int ngroups, int add, double* __restrict__ part,
                                                                       int* __restrict__ range) {
  using T = typename SpecLane<MIXED, RUN>::T;
  constexpr int FR = RUN < kSbFold ? RUN : kSbFold;
  struct Tile {
    double tau[64], psi[64];
    T wc[64], ws[64];
    T a[3][64], g[3][64];
  };
  __shared__ union {
    Tile tile[kSbWaves];
    double fold[kSbWaves][FR][6][64];
  } sm;
  int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int const grp = (int)(blockIdx.x % (unsigned)ngroups);
  int const rc = (int)(blockIdx.x / (unsigned)ngroups);
  int const r = rc / nch, ch = rc % nch;
  BatchParticle const q = P[r];
  int64_t const bl = lo < q.nb ? lo : q.nb, bh = hi < q.nb ? hi : q.nb;
  int64_t const fl = lo < q.nf ? lo : q.nf, fh = hi < q.nf ? hi : q.nf;
  int64_t const nbs = bh - bl, n = nbs + (fh - fl);
  int64_t const first_b = q.nb - bh, first_f = q.nb + fl - nbs;  // point index within the particle
  int64_t const chunk = (n + nch - 1) / nch;
  int64_t const c0 = ch * chunk;
  int64_t const c1 = (c0 + chunk < n) ? c0 + chunk : n;
  int64_t const sub = (chunk + kSbWaves - 1) / kSbWaves;
  int64_t const i0 = c0 + wave * sub;
  int64_t const i1 = (i0 + sub < c1) ? i0 + sub : c1;
  int const ntiles = (int)((sub + 63) / 64);  // the same for every wave of the block
  SpecLane<MIXED, RUN> L;
  L.init(tab.omega, tab.cw, tab.rdw, tab.rtl, ne, (grp * 64 + lane) * RUN);
  Tile& tl = sm.tile[wave];
  double tmax = 0.0;
  double const cwmax = wave_max(L.cw_max());
  // mixed: tiles whose |tau| stays within a tenth of every lane's series range take the
  // first-order correction
  double const tlim1 = kFirstOrderFrac * wave_min(L.tau_lim);
  // the step of every run of this block (one value per block, spectrum_run_tables)
  double const dwb = RUN > 1 ? tab.rdw[(grp * 64 * RUN < ne ? grp * 64 * RUN : 0) / RUN] : 0.0;
  for (int tt = 0; tt < ntiles; ++tt) {
    int64_t const base = i0 + (int64_t)tt * 64;
    int const cnt = base < i1 ? ((i1 - base) < 64 ? (int)(i1 - base) : 64) : 0;
    if (tt > 0) wave_sync();  // the wave's previous tile is consumed
    double dmin = 1e300, tabs = 0.0;
    if (lane < cnt) {
      int64_t const v = base + lane;
      int64_t const it = v < nbs ? first_b + v : first_f + v;  // time index of the point
      int64_t const i = q.off + it;
      SpecPoint G;
      spec_geometry(ox, oy, oz, bt.x[i], bt.y[i], bt.z[i], bt.bx[i], bt.by[i], bt.bz[i], q.dt * (double)it, G);
      tl.tau[lane] = G.tau;
      double psi = 0.0;
      T wc = (T)1, ws = (T)0;
      if constexpr (RUN > 1) spec_step<MIXED>(dwb, G.tau, psi, wc, ws);
      tl.psi[lane] = psi;
      tl.wc[lane] = wc;
      tl.ws[lane] = ws;
      tabs = __builtin_fabs(G.tau);
      tmax = __builtin_fmax(tmax, tabs);
      dmin = G.d;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        tl.a[k][lane] = (T)G.a[k];
        tl.g[k][lane] = (T)G.g[k];
      }
    }
    wave_sync();
    // mixed: this tile keeps the x, y near field unless (c/omega)/D <= kSpecNear throughout
    bool const xy_near = !MIXED || cwmax > kSpecNear * wave_min(dmin);
    bool const second = !MIXED || wave_max(tabs) > tlim1;
    for (int k0 = 0; k0 < cnt; k0 += kSub) {
      int const k1 = (k0 + kSub < cnt) ? k0 + kSub : cnt;
      if constexpr (MIXED) {
        // the next point's values are read from LDS while the current one is integrated
        auto run = [&](auto near) {
          double tn = tl.tau[k0], pn = tl.psi[k0];
          T wcn = tl.wc[k0], wsn = tl.ws[k0];
          T an0 = tl.a[0][k0], an1 = tl.a[1][k0], an2 = tl.a[2][k0];
          T gn0 = tl.g[0][k0], gn1 = tl.g[1][k0], gn2 = tl.g[2][k0];
          for (int k = k0; k < k1; ++k) {
            double const tc = tn, psc = pn;
            T const wcc = wcn, wsc = wsn;
            T const a0 = an0, a1 = an1, a2 = an2, g0 = gn0, g1 = gn1, g2 = gn2;
            int const kn = k + 1 < k1 ? k + 1 : k;
            tn = tl.tau[kn];
            pn = tl.psi[kn];
            wcn = tl.wc[kn];
            wsn = tl.ws[kn];
            an0 = tl.a[0][kn];
            an1 = tl.a[1][kn];
            an2 = tl.a[2][kn];
            gn0 = tl.g[0][kn];
            gn1 = tl.g[1][kn];
            gn2 = tl.g[2][kn];
            L.template add<decltype(near)::value, decltype(near)::value>(tc, psc, wcc, wsc, a0, a1, a2, g0, g1, g2);
          }
        };
        // every term of the tile in full, or the fast form: no x, y near field and the
        // first-order correction
        if (xy_near || second)
          run(std::true_type{});
        else
          run(std::false_type{});
      } else {
        for (int k = k0; k < k1; ++k)
          L.add(tl.tau[k], tl.psi[k], tl.wc[k], tl.ws[k], tl.a[0][k], tl.a[1][k], tl.a[2][k], tl.g[0][k], tl.g[1][k],
                tl.g[2][k]);
      }
      L.flush();
    }
  }
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) tmax = __builtin_fmax(tmax, __shfl_xor(tmax, m, 64));
  if (tmax > L.tau_lim) *range = 1;
#pragma unroll
  for (int r0 = 0; r0 < RUN; r0 += FR) {
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < FR; ++rr)
#pragma unroll
      for (int k = 0; k < 6; ++k) sm.fold[wave][rr][k][lane] = L.acc[r0 + rr][k];
    __syncthreads();
    for (int idx = threadIdx.x; idx < FR * 6 * 64; idx += 64 * kSbWaves) {
      int const rr = idx / 384, k = (idx / 64) % 6, l = idx % 64;
      double v = sm.fold[0][rr][k][l];
#pragma unroll
      for (int w = 1; w < kSbWaves; ++w) v += sm.fold[w][rr][k][l];
      int const e = (grp * 64 + l) * RUN + r0 + rr;
      if (e < ne) {
        double* const dst = part + (((size_t)r * nch + ch) * 6 + k) * ne + e;
        *dst = add ? *dst + v : v;
      }
    }
  }
}

// Chunks in fixed order, then the reference's prefactors (OSCARSSR.cc:1202-1246): the
// field sums (c/omega applied per term) times i (C0_r omega) dt_r, F = C2_r sum |.|^2;
// out[r * ne + e].
__global__ void spectrum_batch_reduce_kernel(const double* __restrict__ part, int nch, int ne, int np,
                                             const BatchParticle* __restrict__ P, const double* __restrict__ omega,
                                             double* __restrict__ out) {
  int64_t const t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)np * ne) return;
  int64_t const r = t / ne, e = t % ne;
  double v[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) v[k] = 0.0;
  for (int c = 0; c < nch; ++c)
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] += part[(((size_t)r * nch + c) * 6 + k) * ne + e];
  double const k2 = (P[r].c0 * omega[e]) * P[r].dt;
  double p[3];
#pragma unroll
  for (int qq = 0; qq < 3; ++qq) {
    double const re = -(v[2 * qq + 1] * k2);
    double const im = v[2 * qq] * k2;
    p[qq] = re * re + im * im;
  }
  out[t] = P[r].c2 * (p[0] + p[1] + p[2]);
}

}  // namespace

bool batch_config_valid(BatchConfig const& c) {
  if (c.ppl == 2) return c.ne_per_item == 1 && (c.tile == 128 || c.tile == 256);
  if (c.ppl != 1) return false;
  switch (c.ne_per_item) {
    case 1: return c.tile == 64 || c.tile == 128 || c.tile == 256;
    case 2: return c.tile == 128 || c.tile == 256;
    case 3: return c.tile == 64 || c.tile == 128 || c.tile == 256;
    case 4: return c.tile == 128 || c.tile == 256;
  }
  return false;
}

int batch_max_blocks(BatchConfig const& c, bool mixed) {
  // resident blocks per CU for this instantiation (sizes the persistent grid)
  int nb = 0;
  hipError_t e = hipErrorInvalidValue;
  if (c.ppl == 2) {
    if (mixed && c.tile == 128) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_flux2_kernel<128>, kBlock, 0);
    if (mixed && c.tile == 256) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_flux2_kernel<256>, kBlock, 0);
    return (e == hipSuccess && nb > 0) ? nb : 1;
  }
#define SRS_OCC(NE_, TL_)                                                                                          \
  if (c.ne_per_item == NE_ && c.tile == TL_)                                                                      \
    e = mixed ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_flux_kernel<NE_, TL_, true>, kBlock, 0)  \
              : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_flux_kernel<NE_, TL_, false>, kBlock, 0);
  SRS_OCC(1, 64)
  SRS_OCC(1, 128)
  SRS_OCC(1, 256)
  SRS_OCC(2, 128)
  SRS_OCC(2, 256)
  SRS_OCC(3, 64)
  SRS_OCC(3, 128)
  SRS_OCC(3, 256)
  SRS_OCC(4, 128)
  SRS_OCC(4, 256)
#undef SRS_OCC
  return (e == hipSuccess && nb > 0) ? nb : 1;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): 10 rounds of two 32x32 -> 64-bit
// products with the published multipliers and Weyl key increments.
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint32_t const h0 = __umulhi(0xD2511F53u, c[0]), l0 = 0xD2511F53u * c[0];
    uint32_t const h1 = __umulhi(0xCD9E8D57u, c[2]), l1 = 0xCD9E8D57u * c[2];
    uint32_t const n0 = h1 ^ c[1] ^ k0, n2 = h0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = l1;
    c[2] = n2;
    c[3] = l0;
  }
}

// Particle idx of the beams under key seed: the beam k with u * W < w_0 + ... + w_k
// (u a fourth Philox block's 53-bit uniform, W the weights' sum; no draw for one beam,
// TParticleBeamContainer::GetRandomBeamIndexByWeight), then three Philox blocks (counter
// (lo, hi, j, 0)) -> 53-bit uniforms u1 in (0, 1], u2 in [0, 1) -> Box-Muller normals,
// and TParticleBeam::GetNewParticle (TParticleBeam.cc:228-265) in the reference's
// operation order with the draws n0..n4 (energy, x, y, x', y'), and TParticleA's gamma
// and q/(m gamma) from the drawn beta (TParticleA.cc:353-382).
__device__ __forceinline__ int draw_particle(const DevBeam* __restrict__ beams, const double* __restrict__ wsum,
                                             int nb, uint64_t seed, uint64_t idx, double X[3], double B[3],
                                             double* qmg_out) {
  int sel = 0;
  if (nb > 1) {
    uint32_t c[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), 3u, 0u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint64_t const a = (((uint64_t)c[0] << 32) | c[1]) >> 11;
    double const r = ((double)a * 1.1102230246251565404e-16) * wsum[nb - 1];
    sel = nb - 1;
    for (int k = 0; k < nb; ++k)
      if (r < wsum[k]) {
        sel = k;
        break;
      }
  }
  DevBeam const& b = beams[sel];
  double n[6];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    uint32_t c[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), (uint32_t)j, 0u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint64_t const a = (((uint64_t)c[0] << 32) | c[1]) >> 11, bb = (((uint64_t)c[2] << 32) | c[3]) >> 11;
    double const u1 = (double)(a + 1) * 1.1102230246251565404e-16, u2 = (double)bb * 1.1102230246251565404e-16;
    double const r = sqrt(-2.0 * log(u1));
    n[2 * j] = r * cos(2.0 * kPi * u2);
    n[2 * j + 1] = r * sin(2.0 * kPi * u2);
  }
  double const enew = b.e0 + b.sigma_e * n[0];
  double const gamma = enew / b.mgev;
  double const beta = sqrt(1.0 - 1.0 / (gamma * gamma));
  double const dmid = (b.at[0] - b.x0[0]) * b.u0[0] + (b.at[1] - b.x0[1]) * b.u0[1] + (b.at[2] - b.x0[2]) * b.u0[2];
#pragma unroll
  for (int k = 0; k < 3; ++k) X[k] = b.x0[k] + ((b.h[k] * b.su[0]) * (1 + dmid)) * n[1];
#pragma unroll
  for (int k = 0; k < 3; ++k) X[k] = X[k] + ((b.v[k] * b.su[1]) * (1 + dmid)) * n[2];
#pragma unroll
  for (int k = 0; k < 3; ++k) B[k] = b.u0[k] * beta;
  {  // TVector3D::RotateSelfY
    double const a = b.sp[0] * n[3], sa = sin(a), ca = cos(a), zz = B[2];
    B[2] = ca * zz - sa * B[0];
    B[0] = sa * zz + ca * B[0];
  }
  {  // TVector3D::RotateSelfX
    double const a = -b.sp[1] * n[4], sa = sin(a), ca = cos(a), yy = B[1];
    B[1] = ca * yy - sa * B[2];
    B[2] = sa * yy + ca * B[2];
  }
  double const b2 = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
  double qmg = 0.0;
  if (b2 != 1) {
    double const g = 1.0 / sqrt(1.0 - b2);
    if (b.m != 0 && g != 0) qmg = b.q / b.m / g;
  }
  *qmg_out = qmg;
  return sel;
}

// One lane per particle: counter first + p, the row of srs_particle.
__global__ __launch_bounds__(256) void sample_beam_kernel(const DevBeam* __restrict__ beams,
                                                          const double* __restrict__ wsum, int nb, uint64_t seed,
                                                          uint64_t first, int64_t np,
                                                          SampledParticle* __restrict__ out) {
  int64_t const p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  double X[3], B[3], qmg;
  int const sel = draw_particle(beams, wsum, nb, seed, first + (uint64_t)p, X, B, &qmg);
  DevBeam const& b = beams[sel];
  SampledParticle& o = out[p];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    o.x0[k] = X[k];
    o.beta0[k] = B[k];
  }
  o.qmg = qmg;
  o.t0 = b.t0;
  o.ct_start = b.ct_start;
  o.ct_stop = b.ct_stop;
  o.npoints = b.npoints;
  o.charge = b.q;
  o.current = b.cur;
}

// The same draws written straight into the batch records: the beam's prefactors and
// time grid, the particle's trajectory at p * stride.
__global__ __launch_bounds__(256) void sample_batch_kernel(const DevBeam* __restrict__ beams,
                                                           const double* __restrict__ wsum, int nb, uint64_t seed,
                                                           uint64_t first, int64_t np, int64_t stride,
                                                           BatchParticle* __restrict__ out) {
  int64_t const p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  double X[3], B[3], qmg;
  int const sel = draw_particle(beams, wsum, nb, seed, first + (uint64_t)p, X, B, &qmg);
  DevBeam const& b = beams[sel];
  BatchParticle& o = out[p];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    o.x0[k] = X[k];
    o.beta0[k] = B[k];
  }
  o.qmg = qmg;
  o.dt = b.dt;
  o.nb = b.gnb;
  o.nf = b.gnf;
  o.off = p * stride;
  o.pole_base = 0;
  o.c0 = b.c0;
  o.c2 = b.c2;
  o.pdq = b.pdq;
}

hipError_t launch_sample_beam(const DevBeam* beams, const double* wsum, int nb, uint64_t seed, uint64_t first,
                              int64_t np, SampledParticle* out, hipStream_t s) {
  sample_beam_kernel<<<(unsigned)((np + 255) / 256), 256, 0, s>>>(beams, wsum, nb, seed, first, np, out);
  return hipGetLastError();
}

hipError_t launch_sample_batch(const DevBeam* beams, const double* wsum, int nb, uint64_t seed, uint64_t first,
                               int64_t np, int64_t stride, BatchParticle* out, hipStream_t s) {
  sample_batch_kernel<<<(unsigned)((np + 255) / 256), 256, 0, s>>>(beams, wsum, nb, seed, first, np, stride, out);
  return hipGetLastError();
}

hipError_t launch_rk4_batch(const DevUndulator* U, int nu, bool local, int force, const double* poles, int npoles,
                            const BatchParticle* P, int np, BatchTraj const& bt, Checkpoints const& ck,
                            RK4Seg const& sg, hipStream_t s) {
  dim3 const grid((np + 63) / 64);
  int const npl = pole_lds_count(npoles);
  size_t const lds = (size_t)npl * kPoleLane * sizeof(double);
  if (force == kForceE) rk4_batch_kernel<0, Force::E><<<grid, 64, lds, s>>>(U, nu, poles, npl, P, np, bt, ck, sg);
  else if (force == kForceEB) rk4_batch_kernel<0, Force::EB><<<grid, 64, lds, s>>>(U, nu, poles, npl, P, np, bt, ck, sg);
  else if (nu == 1 && !local) rk4_batch_kernel<1, Force::B><<<grid, 64, lds, s>>>(U, nu, poles, npl, P, np, bt, ck, sg);
  else if (nu == 2 && !local) rk4_batch_kernel<2, Force::B><<<grid, 64, lds, s>>>(U, nu, poles, npl, P, np, bt, ck, sg);
  else rk4_batch_kernel<0, Force::B><<<grid, 64, lds, s>>>(U, nu, poles, npl, P, np, bt, ck, sg);
  return hipGetLastError();
}

hipError_t launch_zero_counter(unsigned* c, hipStream_t s) {
  zero_counter_kernel<<<1, 64, 0, s>>>(c);
  return hipGetLastError();
}

hipError_t launch_batch_flux(BatchConfig const& cfg, bool mixed, BatchTraj const& bt, const BatchParticle* P,
                             const double* obs, int ns, BatchEnergies const& be, unsigned nitems, unsigned* counter,
                             double* flux_pp, FluxSeg const& sg, int nblocks, hipStream_t s) {
  return mixed ? launch_batch_m<true>(cfg, bt, P, obs, ns, be, nitems, counter, flux_pp, sg, nblocks, s)
               : launch_batch_m<false>(cfg, bt, P, obs, ns, be, nitems, counter, flux_pp, sg, nblocks, s);
}

hipError_t launch_particle_reduce(const double* flux_pp, int np, int nout, double weight, double* out, hipStream_t s) {
  particle_reduce_kernel<<<(nout + 255) / 256, 256, 0, s>>>(flux_pp, np, nout, weight, out);
  return hipGetLastError();
}

int pd_batch_chunks(int64_t npoints, int ns, int np) {
  // enough blocks to fill the device (~2048), chunks of at least 1024 points
  int64_t const tiles = (int64_t)((ns + kBlock - 1) / kBlock) * np;
  int64_t want = (2048 + tiles - 1) / tiles;
  int64_t const most = (npoints + 1023) / 1024;
  if (want > most) want = most;
  return (int)(want < 1 ? 1 : want);
}

hipError_t launch_pd_batch(BatchTraj const& bt, const BatchParticle* P, int np, const double* obs, const double* nrm,
                           int ns, int directional, int64_t lo, int64_t hi, int nch, int add, double* part,
                           bool mixed, hipStream_t s) {
  int const ntiles = (ns + kBlock - 1) / kBlock;
  int64_t const nblocks = (int64_t)ntiles * np * nch;
  if (nblocks > 0x7fffffff) return hipErrorInvalidValue;
  if (mixed)
    pd_batch_kernel<true><<<(unsigned)nblocks, kBlock, 0, s>>>(bt, P, obs, nrm, ns, directional, lo, hi, nch, ntiles,
                                                               add, part);
  else
    pd_batch_kernel<false><<<(unsigned)nblocks, kBlock, 0, s>>>(bt, P, obs, nrm, ns, directional, lo, hi, nch,
                                                                ntiles, add, part);
  return hipGetLastError();
}

hipError_t launch_pd_batch_reduce(const double* part, int nslot, int ns, int np, const BatchParticle* P,
                                  int directional, double* raw, double* pd, hipStream_t s) {
  int64_t const n = (int64_t)np * ns;
  pd_batch_reduce_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(part, nslot, ns, np, P, directional, raw, pd);
  return hipGetLastError();
}

int spectrum_batch_groups(int ne, int run) { return (ne + 64 * run - 1) / (64 * run); }

hipError_t launch_spectrum_batch(BatchTraj const& bt, const BatchParticle* P, int np, const double obs[3],
                                 SpecTables const& tab, int ne, int run, int64_t lo, int64_t hi, int nch, int add,
                                 double* part, int* range, bool mixed, hipStream_t s) {
  int const ng = spectrum_batch_groups(ne, run);
  int64_t const nblocks = (int64_t)ng * np * nch;
  if (nblocks > 0x7fffffff) return hipErrorInvalidValue;
  dim3 const grid((unsigned)nblocks), block(64 * kSbWaves);
  if (mixed) {
    if (run > 1)
      spectrum_batch_kernel<true, kSpecRun><<<grid, block, 0, s>>>(bt, P, obs[0], obs[1], obs[2], tab, ne, lo, hi, nch,
                                                                    ng, add, part, range);
    else
      spectrum_batch_kernel<true, 1><<<grid, block, 0, s>>>(bt, P, obs[0], obs[1], obs[2], tab, ne, lo, hi, nch, ng, add,
                                                             part, range);
  } else {
    if (run > 1)
      spectrum_batch_kernel<false, kSpecRun><<<grid, block, 0, s>>>(bt, P, obs[0], obs[1], obs[2], tab, ne, lo, hi,
                                                                     nch, ng, add, part, range);
    else
      spectrum_batch_kernel<false, 1><<<grid, block, 0, s>>>(bt, P, obs[0], obs[1], obs[2], tab, ne, lo, hi, nch, ng,
                                                              add, part, range);
  }
  return hipGetLastError();
}

hipError_t launch_spectrum_batch_reduce(const double* part, int nch, int ne, int np, const BatchParticle* P,
                                        const double* omega, double* out, hipStream_t s) {
  int64_t const n = (int64_t)np * ne;
  spectrum_batch_reduce_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(part, nch, ne, np, P, omega, out);
  return hipGetLastError();
}

}  // namespace srs
