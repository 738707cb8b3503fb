// Measured ceilings for the roofline: FP64 FMA throughput of this GPU.  Each lane runs eight
// independent FMA chains (so issue, not latency, limits it) for `iters` iterations; one
// launch of `blocks` 256-thread workgroups performs 2 * 8 * iters * 256 * blocks flops.
#include "chemgen_amd.h"

#include <hip/hip_runtime.h>

namespace {

__global__ void __launch_bounds__(256) fma_chains_k(double* __restrict__ out, double a, double b, int iters) {
  double x0 = threadIdx.x * 1e-3, x1 = x0 + 1.0, x2 = x0 + 2.0, x3 = x0 + 3.0;
  double x4 = x0 + 4.0, x5 = x0 + 5.0, x6 = x0 + 6.0, x7 = x0 + 7.0;
  for (int i = 0; i < iters; ++i) {
    x0 = __builtin_fma(x0, a, b); x1 = __builtin_fma(x1, a, b);
    x2 = __builtin_fma(x2, a, b); x3 = __builtin_fma(x3, a, b);
    x4 = __builtin_fma(x4, a, b); x5 = __builtin_fma(x5, a, b);
    x6 = __builtin_fma(x6, a, b); x7 = __builtin_fma(x7, a, b);
  }
  out[blockIdx.x * 256 + threadIdx.x] = ((x0 + x1) + (x2 + x3)) + ((x4 + x5) + (x6 + x7));
}

}  // namespace

extern "C" int cg_probe_fp64_fma(double* out, int32_t blocks, int32_t iters, void* stream) {
  if (!out || blocks <= 0 || iters <= 0) return CG_EINVAL;
  hipLaunchKernelGGL(fma_chains_k, dim3(blocks), dim3(256), 0, (hipStream_t)stream, out, 0.999999, 1e-7, iters);
  return hipGetLastError() == hipSuccess ? CG_OK : CG_EHIP;
}
