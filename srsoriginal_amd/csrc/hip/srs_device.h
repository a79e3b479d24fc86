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
