// Host build of the kernels' scalar math (srsoriginal_amd/csrc/hip/srs_math.h) for
// tests/test_device_math.py: the same source the gfx950 kernels compile.
#define SRS_FN inline
#include "../srsoriginal_amd/csrc/hip/srs_math.h"

extern "C" {
void dm_sincos(const double* x, long n, double* s, double* c) {
  for (long i = 0; i < n; ++i) srs::sincos_fp64(x[i], s + i, c + i);
}
void dm_div(const double* a, const double* b, long n, double* q) {
  for (long i = 0; i < n; ++i) q[i] = srs::div_rn(a[i], b[i], 1.0 / b[i]);
}
void dm_reduce(const double* x, long n, double* r) {
  for (long i = 0; i < n; ++i) r[i] = srs::reduce_two_pi(x[i]);
}
}

extern "C" void dm_sqrt(const double* x, const double* y0, long n, double* out) {
  for (long i = 0; i < n; ++i) out[i] = srs::sqrt_rn(x[i], y0[i]);
}
