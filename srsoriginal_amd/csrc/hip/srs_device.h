// Device-side constants and math shared by the radiation kernels (gfx950).
//
// Physical constants are the reference's dated CODATA values
// (TOSCARSSR.h:20-69), so that host and device derive identical omegas and
// prefactors.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace srs {

constexpr double kPi = 3.14159265358979323846;
constexpr double kC = 2.99792458e8;
constexpr double kQe = 1.602176462e-19;
constexpr double kHbar = 1.054571596e-34;

// Transverse velocity above which a trajectory point does not travel along z.  The
// kernels' fast forms take z as the beam axis: the FP64 terms round only n_z as the
// reference's R/|R| does (n_x, n_y by the reciprocal: with |beta_x|, |beta_y| <= 1e-3
// their last bit moves 1 - beta.n by <= 1e-3 ulp), and the mixed terms stage 1 - beta_z.
// Tiles with a point beyond it take the exact FP64 terms (every component of n
// rounded as the reference rounds it).
constexpr double kTransverseBeta = 1e-3;

__device__ __forceinline__ bool transverse_beam(double bx, double by) {
  return __builtin_fabs(bx) > kTransverseBeta || __builtin_fabs(by) > kTransverseBeta;
}

// Trajectory as SoA device arrays (x, beta, A/c), n points spaced dt in time.
struct TrajView {
  const double* x;
  const double* y;
  const double* z;
  const double* bx;
  const double* by;
  const double* bz;
  const double* ax;
  const double* ay;
  const double* az;
  int64_t n;
  double dt;
};

}  // namespace srs

#define SRS_FN __device__ __forceinline__
#include "srs_math.h"
