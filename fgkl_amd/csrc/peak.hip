// Vector ALU issue-rate microbenchmark (fgkl_measure_valu_rate): the rate at which a full grid of
// waves issues unpacked FP32 FMAs, FP64 FMAs and 32-bit integer adds, the instructions the sweep,
// rescue, adjoint and Smith-Waterman kernels are built from.  Each lane runs 16 independent
// dependency chains, so the loop is bound by issue, not by latency; the grid keeps eight waves
// on every SIMD.  Compiled like every kernel of the library (packed FP32 disabled), so the FP32
// figure is the unpacked rate the library's kernels can reach.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/fgkl.h"
#include "fgkl_internal.hpp"

namespace fgkl {
namespace {

constexpr int kChains = 16;
constexpr int kBlock = 256;

__global__ void __launch_bounds__(kBlock) rate_f32(float* out, int iters, float b, float c) {
  float a[kChains];
#pragma unroll
  for (int k = 0; k < kChains; ++k) a[k] = float(threadIdx.x + k) * 1e-3f;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < kChains; ++k) a[k] = fmaf(a[k], b, c);
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < kChains; ++k) s += a[k];
  if (s == 12345.678f) out[threadIdx.x] = s;  // keeps the loop; never true for b < 1
}

__global__ void __launch_bounds__(kBlock) rate_f64(double* out, int iters, double b, double c) {
  double a[kChains];
#pragma unroll
  for (int k = 0; k < kChains; ++k) a[k] = double(threadIdx.x + k) * 1e-3;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < kChains; ++k) a[k] = fma(a[k], b, c);
  }
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kChains; ++k) s += a[k];
  if (s == 12345.678) out[threadIdx.x] = s;
}

// Two integer instructions per chain step (an add, then an xor), so the chains stay
// non-linear and the compiler cannot close the loop into a formula.
__global__ void __launch_bounds__(kBlock) rate_u32(uint32_t* out, int iters, uint32_t v, uint32_t w) {
  uint32_t a[kChains];
#pragma unroll
  for (int k = 0; k < kChains; ++k) a[k] = threadIdx.x * 2654435761u + k;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < kChains; ++k) a[k] = (a[k] + v) ^ w;
  }
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kChains; ++k) s ^= a[k];
  if (s == 0x12345678u) out[threadIdx.x] = s;
}

#define PK_TRY(expr)                                                                          \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess) {                                                                   \
      set_error(kInternal, std::string("HIP error: ") + hipGetErrorString(_e) + " at " + #expr); \
      return kInternal;                                                                       \
    }                                                                                         \
  } while (0)

}  // namespace
}  // namespace fgkl

using namespace fgkl;

extern "C" int fgkl_measure_valu_rate(int kind, double* lane_ops_per_second, double* best_ms) {
  if (kind < 0 || kind > 2 || !lane_ops_per_second)
    return (set_error(kInvalidArgument, "kind must be 0 (FP32 FMA), 1 (FP64 FMA) or 2 (u32 ops)"),
            kInvalidArgument);
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return (set_error(kInternal, "no HIP device is available: the fgkl GPU engine needs an MI355X (gfx950)"),
            kInternal);
  int dev = 0, cus = 0;
  PK_TRY(hipGetDevice(&dev));
  PK_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int blocks = cus * 8;  // eight waves on each of a CU's four SIMDs
  const int iters = kind == 1 ? 8000 : 20000;
  void* out = nullptr;
  PK_TRY(hipMalloc(&out, kBlock * sizeof(double)));
  hipEvent_t e0, e1;
  PK_TRY(hipEventCreate(&e0));
  PK_TRY(hipEventCreate(&e1));
  double best = 1e30;
  for (int rep = 0; rep < 6; ++rep) {  // the first is a warm-up
    PK_TRY(hipEventRecord(e0, nullptr));
    if (kind == 0)
      hipLaunchKernelGGL(rate_f32, dim3(blocks), dim3(kBlock), 0, nullptr,
                         static_cast<float*>(out), iters, 0.999f, 1e-3f);
    else if (kind == 1)
      hipLaunchKernelGGL(rate_f64, dim3(blocks), dim3(kBlock), 0, nullptr,
                         static_cast<double*>(out), iters, 0.999, 1e-3);
    else
      hipLaunchKernelGGL(rate_u32, dim3(blocks), dim3(kBlock), 0, nullptr,
                         static_cast<uint32_t*>(out), iters, 0x9E3779B9u, 0x85EBCA6Bu);
    PK_TRY(hipGetLastError());
    PK_TRY(hipEventRecord(e1, nullptr));
    PK_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    PK_TRY(hipEventElapsedTime(&ms, e0, e1));
    if (rep > 0 && ms < best) best = ms;
  }
  PK_TRY(hipEventDestroy(e0));
  PK_TRY(hipEventDestroy(e1));
  PK_TRY(hipFree(out));
  const double ops_per_step = kind == 2 ? 2.0 * kChains : double(kChains);
  const double lane_ops = double(blocks) * kBlock * iters * ops_per_step;
  *lane_ops_per_second = lane_ops / (best * 1e-3);
  if (best_ms) *best_ms = best;
  return kOk;
}
