// field_export.hip -- the two binary blocks of the legacy VTK field file (src/io/vtk.ts:27-80,
// src/io/field-export.ts:22-34), produced on the device from the macro buffer in one pass:
// densities as big-endian float32, then velocities scaled to metres per second as big-endian
// float32 triples. The host adds the text lines around them.
//
// The scaling follows Float32Array.map in latticeToField: the float32 lattice value times the
// float64 scale, rounded once to float32.
#include <hip/hip_runtime.h>

#include <cmath>

#include "solver_internal.h"

namespace pw {
namespace {

__device__ __forceinline__ uint32_t be_bits(float v) { return __builtin_bswap32(__float_as_uint(v)); }

__global__ __launch_bounds__(256) void vtk_blocks_kernel(const float4* __restrict__ macro, uint32_t cells,
                                                         double scale, uint32_t* __restrict__ rho_be,
                                                         uint32_t* __restrict__ velocity_be) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= cells) return;
  const float4 m = macro[n];
  rho_be[n] = be_bits(m.x);
  velocity_be[3 * (size_t)n] = be_bits((float)((double)m.y * scale));
  velocity_be[3 * (size_t)n + 1] = be_bits((float)((double)m.z * scale));
  velocity_be[3 * (size_t)n + 2] = be_bits((float)((double)m.w * scale));
}

}  // namespace
}  // namespace pw

using namespace pw;

extern "C" {

pw_status pw_solver_field_vtk_blocks(pw_solver* s, double velocity_scale, void* rho_be, size_t rho_bytes,
                                     void* velocity_be, size_t velocity_bytes) {
  if (!solver_live(s)) {
    set_error("Null or destroyed solver handle.");
    return PW_EINVAL;
  }
  const size_t cells = s->cells;
  if (!rho_be || !velocity_be || rho_bytes != 4 * cells || velocity_bytes != 12 * cells) {
    set_error("The VTK blocks need 4 bytes of density and 12 bytes of velocity per cell.");
    return PW_EINVAL;
  }
  if (!std::isfinite(velocity_scale) || !(velocity_scale > 0.0)) {
    set_error("The velocity scale must be a positive finite number.");
    return PW_EINVAL;
  }
  uint32_t* dev = nullptr;
  hipError_t e = hipMallocAsync((void**)&dev, 16 * cells, s->stream);
  if (e != hipSuccess) return device_error(e, "VTK block buffer");
  hipLaunchKernelGGL(vtk_blocks_kernel, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s->stream, s->macro,
                     (uint32_t)cells, velocity_scale, dev, dev + cells);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(rho_be, dev, 4 * cells, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(velocity_be, dev + cells, 12 * cells, hipMemcpyDeviceToHost, s->stream);
  (void)hipFreeAsync(dev, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  return e == hipSuccess ? PW_OK : device_error(e, "VTK blocks");
}

}  // extern "C"
