// VALU issue rates of gfx950 by instruction class, for the roofline of the
// integral kernels.  Each kernel runs 8 independent dependency chains per lane
// (2048 blocks x 256 lanes: 8 waves per SIMD) and writes one value per lane; the
// rate is wave instructions per second, divided by SIMD cycles (GRBM_GUI_ACTIVE
// under rocprofv3 gives the clock).  The operand class matters: an FP32
// instruction whose operands are all VGPRs (or inline/literal constants) and one
// with an SGPR operand are timed separately.
#include <hip/hip_runtime.h>

#include <cstdio>

#define CK(x)                                                                       \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);   \
      return 1;                                                                     \
    }                                                                               \
  } while (0)

enum Op { kFmaS, kMulS, kFmaV, kMulV, kAddV, kFmaInline, kRcp, kRsq, kSin, kFmaRcp, kNumF32 };
static const char* kName[] = {"fma_f32 sgpr operand",  "mul_f32 sgpr operand", "fma_f32 vgpr",  "mul_f32 vgpr",
                              "add_f32 vgpr",          "fma_f32 inline const", "rcp_f32",       "rsq_f32",
                              "sin_f32",               "fma_f32 + rcp_f32 (2:1)"};
static const double kPerElem[] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1.5};

constexpr int kChains = 8, kInner = 16;

template <int OP>
__global__ __launch_bounds__(256) void f32_kernel(float* out, float a0, float b0, int iters) {
  int const g = blockIdx.x * 256 + threadIdx.x;
  // out starts zeroed: a, b are per-lane values (VGPRs) equal to a0, b0
  float a = a0, b = b0;
  if (OP == kFmaV || OP == kMulV || OP == kAddV || OP == kFmaInline) {
    a = a0 + out[g] * 1e-30f;
    b = b0 - out[g] * 1e-30f;
  }
  float acc[kChains];
#pragma unroll
  for (int i = 0; i < kChains; ++i) acc[i] = threadIdx.x * 1e-3f + i;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int r = 0; r < kInner; ++r) {
#pragma unroll
      for (int i = 0; i < kChains; ++i) {
        if (OP == kFmaS || OP == kFmaV) acc[i] = __builtin_fmaf(acc[i], a, b);
        if (OP == kMulS || OP == kMulV) acc[i] = acc[i] * a;
        if (OP == kAddV) acc[i] = acc[i] + a;
        if (OP == kFmaInline) acc[i] = __builtin_fmaf(acc[i], a, 1.0f);
        if (OP == kRcp) acc[i] = __builtin_amdgcn_rcpf(acc[i]);
        if (OP == kRsq) acc[i] = __builtin_amdgcn_rsqf(acc[i]);
        if (OP == kSin) acc[i] = __builtin_amdgcn_sinf(acc[i]);
        if (OP == kFmaRcp) {
          acc[i] = __builtin_fmaf(acc[i], a, b);
          if (i & 1) acc[i] = __builtin_amdgcn_rcpf(acc[i]);
        }
      }
    }
  }
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < kChains; ++i) s += acc[i];
  out[g] = s;
}

// FP64-side classes: 0 add_f64, 1 add_f64 + rndne_f64, 2 add_f64 + fract_f64,
// 3 cvt_f32_f64 + cvt_f64_f32 + add_f64, 4 cos_f32 (on a float chain)
template <int OP>
__global__ __launch_bounds__(256) void f64c_kernel(double* out, double b0, int iters) {
  int const g = blockIdx.x * 256 + threadIdx.x;
  double const b = b0 - out[g] * 1e-300;
  double acc[kChains];
  float accf[kChains];
#pragma unroll
  for (int i = 0; i < kChains; ++i) {
    acc[i] = threadIdx.x * 1e-3 + i;
    accf[i] = threadIdx.x * 1e-3f + i;
  }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int r = 0; r < kInner; ++r) {
#pragma unroll
      for (int i = 0; i < kChains; ++i) {
        if (OP == 0) acc[i] = acc[i] + b;
        if (OP == 1) acc[i] = __builtin_rint(acc[i] + b);
        if (OP == 2) acc[i] = __builtin_amdgcn_fract(acc[i] + b);
        if (OP == 3) acc[i] = (double)(float)acc[i] + b;
        if (OP == 4) accf[i] = __builtin_amdgcn_cosf(accf[i]);
      }
    }
  }
  double s = 0.;
#pragma unroll
  for (int i = 0; i < kChains; ++i) s += acc[i] + accf[i];
  out[g] = s;
}

template <bool VGPR>
__global__ __launch_bounds__(256) void f64_kernel(double* out, double a0, double b0, int iters) {
  int const g = blockIdx.x * 256 + threadIdx.x;
  double a = a0, b = b0;
  if (VGPR) {
    a = a0 + out[g] * 1e-300;
    b = b0 - out[g] * 1e-300;
  }
  double acc[kChains];
#pragma unroll
  for (int i = 0; i < kChains; ++i) acc[i] = threadIdx.x * 1e-3 + i;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int r = 0; r < kInner; ++r) {
#pragma unroll
      for (int i = 0; i < kChains; ++i) acc[i] = __builtin_fma(acc[i], a, b);
    }
  }
  double s = 0.;
#pragma unroll
  for (int i = 0; i < kChains; ++i) s += acc[i];
  out[g] = s;
}

template <int OP>
static void launch_f32(int blocks, float* out, int iters) {
  float const a = (OP == kAddV) ? 1e-7f : 0.999f;
  f32_kernel<OP><<<blocks, 256>>>(out, a, 1e-3f, iters);
}

int main() {
  int const blocks = 2048, iters = 2000;
  float* out = nullptr;
  double* outd = nullptr;
  CK(hipMalloc(&out, (size_t)blocks * 256 * sizeof(float)));
  CK(hipMalloc(&outd, (size_t)blocks * 256 * sizeof(double)));
  CK(hipMemset(out, 0, (size_t)blocks * 256 * sizeof(float)));
  CK(hipMemset(outd, 0, (size_t)blocks * 256 * sizeof(double)));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  double const waves = blocks * 4.0, per_lane = (double)kChains * kInner * iters;
  static const char* kName64[] = {"add_f64", "add_f64 + rndne_f64", "add_f64 + fract_f64",
                                  "cvt_f32_f64 + cvt_f64_f32 + add_f64", "cos_f32"};
  static const double kPer64[] = {1, 2, 2, 3, 1};
  for (int op = 0; op < kNumF32 + 2 + 5; ++op) {
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      CK(hipEventRecord(e0));
      switch (op) {
        case kFmaS: launch_f32<kFmaS>(blocks, out, iters); break;
        case kMulS: launch_f32<kMulS>(blocks, out, iters); break;
        case kFmaV: launch_f32<kFmaV>(blocks, out, iters); break;
        case kMulV: launch_f32<kMulV>(blocks, out, iters); break;
        case kAddV: launch_f32<kAddV>(blocks, out, iters); break;
        case kFmaInline: launch_f32<kFmaInline>(blocks, out, iters); break;
        case kRcp: launch_f32<kRcp>(blocks, out, iters); break;
        case kRsq: launch_f32<kRsq>(blocks, out, iters); break;
        case kSin: launch_f32<kSin>(blocks, out, iters); break;
        case kFmaRcp: launch_f32<kFmaRcp>(blocks, out, iters); break;
        case kNumF32: f64_kernel<false><<<blocks, 256>>>(outd, 0.999, 1e-3, iters); break;
        case kNumF32 + 1: f64_kernel<true><<<blocks, 256>>>(outd, 0.999, 1e-3, iters); break;
        case kNumF32 + 2: f64c_kernel<0><<<blocks, 256>>>(outd, 1e-3, iters); break;
        case kNumF32 + 3: f64c_kernel<1><<<blocks, 256>>>(outd, 1e-3, iters); break;
        case kNumF32 + 4: f64c_kernel<2><<<blocks, 256>>>(outd, 1e-3, iters); break;
        case kNumF32 + 5: f64c_kernel<3><<<blocks, 256>>>(outd, 1e-3, iters); break;
        default: f64c_kernel<4><<<blocks, 256>>>(outd, 1e-3, iters); break;
      }
      CK(hipGetLastError());
      CK(hipEventRecord(e1));
      CK(hipEventSynchronize(e1));
      float ms = 0.f;
      CK(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) best = ms;
    }
    char const* name = op < kNumF32 ? kName[op]
                       : op == kNumF32 ? "fma_f64 sgpr operand"
                       : op == kNumF32 + 1 ? "fma_f64 vgpr"
                                           : kName64[op - kNumF32 - 2];
    double const n = waves * per_lane * (op < kNumF32 ? kPerElem[op] : op < kNumF32 + 2 ? 1.0 : kPer64[op - kNumF32 - 2]);
    std::printf("%-26s %8.3f ms  %.3e wave instructions/s  %.3f per SIMD-cycle at 2.4 GHz\n", name, best,
                n / (best * 1e-3), n / (best * 1e-3) / (1024 * 2.4e9));
  }
  CK(hipFree(out));
  CK(hipFree(outd));
  return 0;
}
