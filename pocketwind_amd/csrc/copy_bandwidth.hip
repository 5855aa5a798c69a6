// copy_bandwidth.hip -- device-to-device float4 copy probe (src/lbm/shaders/copy_bandwidth.wgsl,
// src/app/bench-page.ts:204-242), the measured ceiling the reference reports its roofline
// against; also the reference copy for scaling rocprof HBM byte counters.
#include <hip/hip_runtime.h>

#include "solver_internal.h"

namespace pw {

__global__ __launch_bounds__(256) void copy_float4_kernel(const float4* __restrict__ in,
                                                          float4* __restrict__ out, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) out[k] = in[k];
}

}  // namespace pw

extern "C" pw_status pw_copy_bandwidth(size_t bytes, uint32_t passes, double* gbps) {
  using namespace pw;
  if (!gbps || bytes < 16 || passes == 0) {
    set_error("Copy probe needs at least 16 bytes and one pass.");
    return PW_EINVAL;
  }
  const size_t n = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  pw_status status = PW_OK;
  hipError_t e = hipSuccess;
  int cus = 256;
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0);
  const unsigned blocks = (unsigned)(cus * 8);
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
