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
#include "srs_rk4.h"
#include "srs_terms.h"

namespace srs {
namespace {

// Same point layout as the host trajectory (OSCARSSR.cc:562-673): forward pass from
// T0 stores the state before each step with the previous derivative (A/c lags one
// point; 0 at the first point), the backward pass fills the nb points before T0.
template <int NU>
__global__ __launch_bounds__(64) void rk4_batch_kernel(const DevUndulator* __restrict__ U, int nu,
                                                        const double* __restrict__ poles,
                                                        const BatchParticle* __restrict__ Ps, int np, BatchTraj bt,
                                                        Checkpoints ck, RK4Seg sg) {
  int const p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  BatchParticle const q = Ps[p];
  RK4Ctx<NU> const R{Fields<NU>(U, nu, poles + q.pole_base), q.qmg};
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
  if (sg.lo < q.nf) {
    if (sg.lo == 0) init();
    else load(0);
    double const h = q.dt, hh = h * 0.5, h6 = div_rn(h, 6.0, kInvSix);
    int64_t const i1 = sg.hi < q.nf ? sg.hi : q.nf;
    for (int64_t i = sg.lo; i < i1; ++i) {
      if (ck.state && i % ck.every == 0) save(i / ck.every);
      store_point(bt, q.off + q.nb + i, x, d);
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
      store_point(bt, q.off + q.nb - 1 - i, x, d);
    }
    keep(1);
  }
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
      es.kw[e] = be.c0w[ee] * q.dt;  // (C0*omega)*dt, OSCARSSR.cc:2080
      es.fr[e] = be.fr[ee];
      es.fw[e] = be.fw[ee];
      cwf[e] = (float)be.cw[ee];
    }
    int const j0 = tile * kBlock;
    int const j = j0 + tid;
    int const jj = j < ns ? j : j0;  // idle lanes repeat the tile's first point
    double const ox = obs[jj], oy = obs[ns + jj], oz = obs[2 * ns + jj];
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
          if (MIXED) flux_terms_mixed<NE, TILE_>(T.raw, cnt, base, q.dt, ox, oy, oz, es, cwf, a32);
          else flux_terms_fp64<NE, TILE_>(T.raw, cnt, base, q.dt, ox, oy, oz, es, a64);
        }
      }
    }
    if (j < ns) {
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        int const ee = g * NE + e;
        if (ee >= be.ne) continue;
        // this segment's six field sums (+ the earlier segments', in segment order)
        double S[6];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          S[2 * c] = MIXED ? a32.value(e, 2 * c) : a64.re[e][c];
          S[2 * c + 1] = MIXED ? a32.value(e, 2 * c + 1) : a64.im[e][c];
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
                                                                                    es.kw[e], be.c2);
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

}  // namespace

bool batch_config_valid(BatchConfig const& c) {
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

hipError_t launch_rk4_batch(const DevUndulator* U, int nu, const double* poles, const BatchParticle* P, int np,
                            BatchTraj const& bt, Checkpoints const& ck, RK4Seg const& sg, hipStream_t s) {
  dim3 const grid((np + 63) / 64);
  if (nu == 1) rk4_batch_kernel<1><<<grid, 64, 0, s>>>(U, nu, poles, P, np, bt, ck, sg);
  else if (nu == 2) rk4_batch_kernel<2><<<grid, 64, 0, s>>>(U, nu, poles, P, np, bt, ck, sg);
  else rk4_batch_kernel<0><<<grid, 64, 0, s>>>(U, nu, poles, P, np, bt, ck, sg);
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

}  // namespace srs
