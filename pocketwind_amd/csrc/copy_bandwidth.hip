// copy_bandwidth.hip -- device-to-device float4 copy probe (src/lbm/shaders/copy_bandwidth.wgsl,
// src/app/bench-page.ts:204-242), the measured ceiling the reference reports its roofline
// against; also the reference copy for scaling rocprof HBM byte counters.
#include <hip/hip_runtime.h>

#include "solver_internal.h"

namespace pw {

// One float4 per lane with non-temporal accesses: the fastest plain copy measured on gfx950
// (a grid-stride loop with cached accesses reaches about 10 % less).
typedef float float4v __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void copy_float4_kernel(const float4v* __restrict__ in,
                                                          float4v* __restrict__ out, size_t n) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) __builtin_nontemporal_store(__builtin_nontemporal_load(in + k), out + k);
}

}  // namespace pw

extern "C" pw_status pw_copy_bandwidth(size_t bytes, uint32_t passes, double* gbps) {
  using namespace pw;
  if (!gbps || bytes < 16 || passes == 0 || bytes / 16 > 255ull * 0xffffffffull) {
    set_error("Copy probe needs at least 16 bytes, at most one grid of blocks, and one pass.");
    return PW_EINVAL;
  }
  const size_t n = bytes / 16;
  float4v *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  pw_status status = PW_OK;
  hipError_t e = hipSuccess;
  const size_t block_count = (n + 255) / 256;
  const unsigned blocks = (unsigned)block_count;
  float ms = 0;
  if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) goto done;
  if ((e = hipMalloc(&a, n * 16)) != hipSuccess) goto done;
  if ((e = hipMalloc(&b, n * 16)) != hipSuccess) goto done;
  if ((e = hipMemsetAsync(a, 0, n * 16, st)) != hipSuccess) goto done;
  if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) goto done;
  hipLaunchKernelGGL(copy_float4_kernel, dim3(blocks), dim3(256), 0, st, a, b, n);  // warm-up
  if ((e = hipEventRecord(e0, st)) != hipSuccess) goto done;
  for (uint32_t k = 0; k < passes; ++k)
    hipLaunchKernelGGL(copy_float4_kernel, dim3(blocks), dim3(256), 0, st, (k & 1) ? b : a, (k & 1) ? a : b, n);
  if ((e = hipEventRecord(e1, st)) != hipSuccess) goto done;
  if ((e = hipEventSynchronize(e1)) != hipSuccess) goto done;
  if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) goto done;
  *gbps = 2.0 * (double)(n * 16) * passes / (ms * 1e6);
done:
  if (e != hipSuccess) status = device_error(e, "copy probe");
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (a) hipFree(a);
  if (b) hipFree(b);
  if (st) hipStreamDestroy(st);
  return status;
}
