// mlp.hip -- training and batched evaluation of the learned wall closure's network.
//
// A training step is two launches: the gradient kernel (one wavefront per block, tiles of 32
// samples, forward and backward on bf16 MFMA with f32 accumulation, per-block partial
// gradients) and the update kernel (one thread per parameter: fixed-order sum of the
// partials, Adam on the f32 master weights, bf16 repacking for the matrix cores). Summation
// order never depends on timing, so a step is bit-reproducible.
//
// The network has no reference counterpart; see mlp.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <new>
#include <vector>

#include "mlp.h"
#include "solver_internal.h"

namespace pw {
namespace mlp {

constexpr int kTile = 32;              // samples per pass of the gradient kernel
constexpr int kPartial = kParams + 1;  // partial gradient plus the partial squared error
constexpr unsigned kMaxGradBlocks = 256;

// Sample `k` of minibatch `step`: uniform over the data set (with replacement).
__device__ __forceinline__ uint32_t sample_index(uint32_t seed, uint32_t step, uint32_t k, uint32_t n) {
  uint32_t h = seed * 0x9e3779b9u ^ step * 0x85ebca6bu ^ k * 0xc2b2ae35u;
  h = (h ^ (h >> 16)) * 0x7feb352du;
  h = (h ^ (h >> 15)) * 0x846ca68bu;
  h ^= h >> 16;
  return (uint32_t)(((uint64_t)h * n) >> 32);
}

__global__ __launch_bounds__(64) void forward_kernel(const Packed* __restrict__ P, const float* __restrict__ x,
                                                     float* __restrict__ y, uint32_t rows) {
  __shared__ __attribute__((aligned(16))) uint16_t xs[16 * kInPad];
  __shared__ __attribute__((aligned(16))) uint16_t h1[16 * kHidden];
  __shared__ __attribute__((aligned(16))) uint16_t h2[16 * kHidden];
  const int l = threadIdx.x, r = l & 15, g = l >> 4;
  const uint32_t row0 = blockIdx.x * 16;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k = 8 * g + j;
    const uint32_t row = row0 + r;
    xs[r * kInPad + k] = (k < kIn && row < rows) ? bf16_bits(x[(size_t)row * kIn + k]) : 0;
  }
  wave_sync();
  const f32x4_t out = forward16(P, xs, h1, h2);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t row = row0 + 4 * g + j;
    if (r < kOut && row < rows) y[(size_t)row * kOut + r] = out[j];
  }
}

struct GradLds {
  uint16_t xs[kTile * kInPad];     // X [s][in]
  uint16_t xt[kInPad * kTile];     // X^T [in][s]
  uint16_t h1[kTile * kHidden];    // H1 [s][h]
  uint16_t h1t[kHidden * kTile];   // H1^T [h][s]
  uint16_t h2[kTile * kHidden];    // H2 [s][h], then dZ2 [s][h]
  uint16_t h2t[kHidden * kTile];   // H2^T [h][s], then dZ2^T [h][s]
  uint16_t dz1t[kHidden * kTile];  // dZ1^T [h][s]
  uint16_t dy[kTile * kInPad];     // dY [s][out, padded to 32]
  uint16_t dyt[kOutPad * kTile];   // dY^T [out, padded to 16][s]
  float z1[kTile * kHidden];       // Z1 [s][h], then dZ1
  float z2[kTile * kHidden];       // Z2 [s][h], then dZ2
  float t[kTile * 4];              // targets and the sample mask
};

// Partial gradient of the minibatch over the tiles blockIdx.x, blockIdx.x + gridDim.x, ...
// Minibatch sample k is data row k (sampled == 0) or sample_index(seed, step, k, n_data).
// dy_scale = 2 / (3 batch) makes the gradient that of the mean squared error.
__global__ __launch_bounds__(64) void grad_kernel(const Packed* __restrict__ P, const float* __restrict__ x,
                                                  const float* __restrict__ t, uint32_t n_data, uint32_t batch,
                                                  uint32_t sampled, uint32_t seed, uint32_t step, float dy_scale,
                                                  float* __restrict__ partials) {
  __shared__ __attribute__((aligned(16))) GradLds L;
  const int l = threadIdx.x, r = l & 15, g = l >> 4;
  const f32x4_t zero = {0.0f, 0.0f, 0.0f, 0.0f};
  // Padding that no tile overwrites: input columns 12..31, output columns 16..31 of dY.
  for (int k = l; k < kTile * kInPad; k += 64) {
    const int a = k / kInPad, b = k % kInPad;
    if (b >= kIn) L.xs[k] = 0;
    if (a >= kIn) L.xt[k] = 0;
    if (b >= kOutPad) L.dy[k] = 0;
  }
  f32x4_t gw1[4], gw2[16], gw3[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) gw1[k] = gw3[k] = zero;
#pragma unroll
  for (int k = 0; k < 16; ++k) gw2[k] = zero;
  float gb1 = 0.0f, gb2 = 0.0f, gb3 = 0.0f, sq = 0.0f;
  const uint32_t tiles = (batch + kTile - 1) / kTile;
  for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    {  // samples: lane l loads features 6 (l / 32) .. +5 and target l / 32 .. of sample l % 32
      const int s = l & 31, part = l >> 5;
      const uint32_t k = tile * kTile + s;
      const bool valid = k < batch;
      const uint32_t row = valid ? (sampled ? sample_index(seed, step, k, n_data) : k) : 0;
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        const int f = 6 * part + j;
        const uint16_t b = valid ? bf16_bits(x[(size_t)row * kIn + f]) : 0;
        L.xs[s * kInPad + f] = b;
        L.xt[f * kTile + s] = b;
      }
      if (part == 0) {
#pragma unroll
        for (int o = 0; o < kOut; ++o) L.t[s * 4 + o] = valid ? t[(size_t)row * kOut + o] : 0.0f;
        L.t[s * 4 + 3] = valid ? 1.0f : 0.0f;
      }
    }
    wave_sync();
    // Layer 1.
#pragma unroll
    for (int hh = 0; hh < 2; ++hh)
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) {
        const f32x4_t acc = mma(L.xs + 16 * hh * kInPad, kInPad, P->w1 + nb * 16 * kInPad, kInPad, zero);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int s = 16 * hh + 4 * g + j, c = 16 * nb + r;
          const float z = acc[j] + P->b1[c];
          const uint16_t b = bf16_bits(silu(z));
          L.z1[s * kHidden + c] = z;
          L.h1[s * kHidden + c] = b;
          L.h1t[c * kTile + s] = b;
        }
      }
    wave_sync();
    // Layer 2.
#pragma unroll
    for (int hh = 0; hh < 2; ++hh)
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) {
        f32x4_t acc = zero;
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
          acc = mma(L.h1 + 16 * hh * kHidden + 32 * kb, kHidden, P->w2 + nb * 16 * kHidden + 32 * kb, kHidden, acc);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int s = 16 * hh + 4 * g + j, c = 16 * nb + r;
          const float z = acc[j] + P->b2[c];
          const uint16_t b = bf16_bits(silu(z));
          L.z2[s * kHidden + c] = z;
          L.h2[s * kHidden + c] = b;
          L.h2t[c * kTile + s] = b;
        }
      }
    wave_sync();
    // Layer 3 and dY = dy_scale (y - t).
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
      f32x4_t acc = zero;
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) acc = mma(L.h2 + 16 * hh * kHidden + 32 * kb, kHidden, P->w3 + 32 * kb, kHidden, acc);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int s = 16 * hh + 4 * g + j;
        float d = 0.0f;
        if (r < kOut) {
          const float diff = (acc[j] + P->b3[r] - L.t[s * 4 + r]) * L.t[s * 4 + 3];
          sq += diff * diff;
          d = dy_scale * diff;
          gb3 += d;
        }
        const uint16_t b = bf16_bits(d);
        L.dy[s * kInPad + r] = b;
        L.dyt[r * kTile + s] = b;
      }
    }
    wave_sync();
    // dW3 += dY^T H2.
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) gw3[nb] = mma(L.dyt, kTile, L.h2t + nb * 16 * kTile, kTile, gw3[nb]);
    wave_sync();
    // dZ2 = (dY W3) * silu'(Z2), stored over H2 (bf16) and Z2 (f32).
#pragma unroll
    for (int hh = 0; hh < 2; ++hh)
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) {
        const f32x4_t acc = mma(L.dy + 16 * hh * kInPad, kInPad, P->w3t + nb * 16 * kInPad, kInPad, zero);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int s = 16 * hh + 4 * g + j, c = 16 * nb + r;
          const float d = acc[j] * silu_prime(L.z2[s * kHidden + c]);
          const uint16_t b = bf16_bits(d);
          L.z2[s * kHidden + c] = d;
          L.h2[s * kHidden + c] = b;
          L.h2t[c * kTile + s] = b;
        }
      }
    wave_sync();
    // dW2 += dZ2^T H1; db2 += sum over samples of dZ2.
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int nb = 0; nb < 4; ++nb)
        gw2[mb * 4 + nb] = mma(L.h2t + mb * 16 * kTile, kTile, L.h1t + nb * 16 * kTile, kTile, gw2[mb * 4 + nb]);
    for (int s = 0; s < kTile; ++s) gb2 += L.z2[s * kHidden + l];
    // dZ1 = (dZ2 W2) * silu'(Z1), stored over Z1 (f32) and as dZ1^T (bf16).
#pragma unroll
    for (int hh = 0; hh < 2; ++hh)
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) {
        f32x4_t acc = zero;
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
          acc = mma(L.h2 + 16 * hh * kHidden + 32 * kb, kHidden, P->w2t + nb * 16 * kHidden + 32 * kb, kHidden, acc);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int s = 16 * hh + 4 * g + j, c = 16 * nb + r;
          const float d = acc[j] * silu_prime(L.z1[s * kHidden + c]);
          L.z1[s * kHidden + c] = d;
          L.dz1t[c * kTile + s] = bf16_bits(d);
        }
      }
    wave_sync();
    // dW1 += dZ1^T X (inputs 0..15; 12..15 are zero); db1 += sum of dZ1.
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) gw1[mb] = mma(L.dz1t + mb * 16 * kTile, kTile, L.xt, kTile, gw1[mb]);
    for (int s = 0; s < kTile; ++s) gb1 += L.z1[s * kHidden + l];
    wave_sync();
  }
  float* out = partials + (size_t)blockIdx.x * kPartial;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (r < kIn) out[oW1 + (16 * mb + 4 * g + j) * kIn + r] = gw1[mb][j];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int nb = 0; nb < 4; ++nb)
#pragma unroll
      for (int j = 0; j < 4; ++j) out[oW2 + (16 * mb + 4 * g + j) * kHidden + 16 * nb + r] = gw2[mb * 4 + nb][j];
#pragma unroll
  for (int nb = 0; nb < 4; ++nb)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (4 * g + j < kOut) out[oW3 + (4 * g + j) * kHidden + 16 * nb + r] = gw3[nb][j];
  out[oB1 + l] = gb1;
  out[oB2 + l] = gb2;
  gb3 += __shfl_xor(gb3, 16, 64);
  gb3 += __shfl_xor(gb3, 32, 64);
  if (l < kOut) out[oB3 + l] = gb3;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off, 64);
  if (l == 0) out[kParams] = sq;
}

enum { kPackOnly = 0, kGradient = 1, kAdam = 2 };

// One thread per parameter: fixed-order sum of the block partials, optional Adam update of
// the f32 master weight, and the bf16 copies the matrix cores read. Thread kParams sums the
// squared error into *loss (times loss_scale).
__global__ __launch_bounds__(256) void update_kernel(float* __restrict__ master, float* __restrict__ m1,
                                                     float* __restrict__ m2, Packed* __restrict__ P,
                                                     const float* __restrict__ partials, uint32_t blocks, int mode,
                                                     float lr, float bc1, float bc2, float* __restrict__ grads,
                                                     float* __restrict__ loss, float loss_scale) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kParams) return;
  if (p == kParams) {
    if (mode != kPackOnly && loss) {
      float s = 0.0f;
      for (uint32_t b = 0; b < blocks; ++b) s += partials[(size_t)b * kPartial + kParams];
      *loss = s * loss_scale;
    }
    return;
  }
  float w = master[p];
  if (mode != kPackOnly) {
    float gsum = 0.0f;
    for (uint32_t b = 0; b < blocks; ++b) gsum += partials[(size_t)b * kPartial + p];
    if (grads) grads[p] = gsum;
    if (mode == kAdam) {
      // Kingma and Ba 2015, Algorithm 1 (beta1 0.9, beta2 0.999, eps 1e-8).
      const float a = 0.9f * m1[p] + 0.1f * gsum;
      const float b = 0.999f * m2[p] + 0.001f * gsum * gsum;
      m1[p] = a;
      m2[p] = b;
      w -= lr * (a / bc1) / (sqrtf(b / bc2) + 1e-8f);
      master[p] = w;
    }
  }
  const uint16_t h = bf16_bits(w);
  if (p < oB1) {
    const int o = (p - oW1) / kIn, i = (p - oW1) % kIn;
    P->w1[o * kInPad + i] = h;
  } else if (p < oW2) {
    P->b1[p - oB1] = w;
  } else if (p < oB2) {
    const int o = (p - oW2) / kHidden, i = (p - oW2) % kHidden;
    P->w2[o * kHidden + i] = h;
    P->w2t[i * kHidden + o] = h;
  } else if (p < oW3) {
    P->b2[p - oB2] = w;
  } else if (p < oB3) {
    const int o = (p - oW3) / kHidden, i = (p - oW3) % kHidden;
    P->w3[o * kHidden + i] = h;
    P->w3t[i * kInPad + o] = h;
  } else {
    P->b3[p - oB3] = w;
  }
}

}  // namespace mlp
}  // namespace pw

using namespace pw;
using namespace pw::mlp;

struct pw_mlp {
  hipStream_t stream = nullptr;
  float* master = nullptr;
  float* m1 = nullptr;
  float* m2 = nullptr;
  Packed* packed = nullptr;
  float* partials = nullptr;
  float* data_x = nullptr;
  float* data_t = nullptr;
  size_t n_data = 0;
  uint64_t adam_step = 0;
};

namespace pw {
const void* mlp_packed(const pw_mlp* m) { return m ? m->packed : nullptr; }
size_t mlp_packed_bytes() { return sizeof(Packed); }
}  // namespace pw

namespace {

void free_mlp(pw_mlp* m) {
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (void* p : {(void*)m->master, (void*)m->m1, (void*)m->m2, (void*)m->packed, (void*)m->partials,
                  (void*)m->data_x, (void*)m->data_t})
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
}

void launch_update(pw_mlp* m, int mode, uint32_t blocks, float lr, float* grads, float* loss, float loss_scale) {
  float bc1 = 1.0f, bc2 = 1.0f;
  if (mode == kAdam) {
    bc1 = (float)(1.0 - std::pow(0.9, (double)m->adam_step));
    bc2 = (float)(1.0 - std::pow(0.999, (double)m->adam_step));
  }
  hipLaunchKernelGGL(update_kernel, dim3((kParams + 1 + 255) / 256), dim3(256), 0, m->stream, m->master, m->m1, m->m2,
                     m->packed, m->partials, blocks, mode, lr, bc1, bc2, grads, loss, loss_scale);
}

// Gradient of the mean squared error over a minibatch of `batch` samples into the partials;
// returns the number of partial blocks.
uint32_t launch_grad(pw_mlp* m, const float* x, const float* t, size_t n, uint32_t batch, bool sampled,
                     uint32_t seed, uint32_t step) {
  const uint32_t tiles = (batch + kTile - 1) / kTile;
  const uint32_t blocks = tiles < kMaxGradBlocks ? tiles : kMaxGradBlocks;
  hipLaunchKernelGGL(grad_kernel, dim3(blocks), dim3(64), 0, m->stream, m->packed, x, t, (uint32_t)n, batch,
                     sampled ? 1u : 0u, seed, step, 2.0f / (3.0f * (float)batch), m->partials);
  return blocks;
}

}  // namespace

extern "C" {

pw_status pw_mlp_create(pw_mlp** out) {
  if (!out) {
    set_error("Null argument.");
    return PW_EINVAL;
  }
  *out = nullptr;
  pw_mlp* m = new (std::nothrow) pw_mlp();
  if (!m) {
    set_error("Host allocation failed.");
    return PW_ENOMEM;
  }
  hipError_t e;
  auto fail = [&](const char* what) {
    free_mlp(m);
    delete m;
    return device_error(e, what);
  };
  if ((e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking)) != hipSuccess) return fail("network stream");
  if ((e = hipMalloc(&m->master, kParams * sizeof(float))) != hipSuccess ||
      (e = hipMalloc(&m->m1, kParams * sizeof(float))) != hipSuccess ||
      (e = hipMalloc(&m->m2, kParams * sizeof(float))) != hipSuccess ||
      (e = hipMalloc(&m->packed, sizeof(Packed))) != hipSuccess ||
      (e = hipMalloc(&m->partials, (size_t)kMaxGradBlocks * kPartial * sizeof(float))) != hipSuccess)
    return fail("network allocation");
  if ((e = hipMemsetAsync(m->master, 0, kParams * sizeof(float), m->stream)) != hipSuccess ||
      (e = hipMemsetAsync(m->m1, 0, kParams * sizeof(float), m->stream)) != hipSuccess ||
      (e = hipMemsetAsync(m->m2, 0, kParams * sizeof(float), m->stream)) != hipSuccess ||
      (e = hipMemsetAsync(m->packed, 0, sizeof(Packed), m->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(m->stream)) != hipSuccess)
    return fail("network initialisation");
  *out = m;
  return PW_OK;
}

void pw_mlp_destroy(pw_mlp* m) {
  if (!m) return;
  free_mlp(m);
  delete m;
}

pw_status pw_mlp_set_params(pw_mlp* m, const float* params, size_t n) {
  if (!m || !params || n != (size_t)kParams) {
    set_error("Parameters must be one vector of PW_MLP_PARAMS floats.");
    return PW_EINVAL;
  }
  hipError_t e;
  if ((e = hipMemcpyAsync(m->master, params, kParams * sizeof(float), hipMemcpyHostToDevice, m->stream)) != hipSuccess ||
      (e = hipMemsetAsync(m->m1, 0, kParams * sizeof(float), m->stream)) != hipSuccess ||
      (e = hipMemsetAsync(m->m2, 0, kParams * sizeof(float), m->stream)) != hipSuccess)
    return device_error(e, "network parameters");
  m->adam_step = 0;
  launch_update(m, kPackOnly, 0, 0.0f, nullptr, nullptr, 0.0f);
  if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(m->stream)) != hipSuccess)
    return device_error(e, "network parameters");
  return PW_OK;
}

pw_status pw_mlp_get_params(pw_mlp* m, float* params, size_t n) {
  if (!m || !params || n < (size_t)kParams) {
    set_error("Parameter output needs PW_MLP_PARAMS floats.");
    return PW_EINVAL;
  }
  hipError_t e = hipMemcpyAsync(params, m->master, kParams * sizeof(float), hipMemcpyDeviceToHost, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  return e == hipSuccess ? PW_OK : device_error(e, "network parameters");
}

pw_status pw_mlp_forward(pw_mlp* m, const float* x, size_t rows, float* y) {
  if (!m || (!x && rows) || (!y && rows) || rows > 0xffffffffull) {
    set_error("Forward needs rows x 12 inputs and rows x 3 outputs.");
    return PW_EINVAL;
  }
  if (rows == 0) return PW_OK;
  float *dx = nullptr, *dy = nullptr;
  hipError_t e;
  if ((e = hipMallocAsync((void**)&dx, rows * kIn * sizeof(float), m->stream)) != hipSuccess)
    return device_error(e, "forward buffers");
  if ((e = hipMallocAsync((void**)&dy, rows * kOut * sizeof(float), m->stream)) != hipSuccess) {
    (void)hipFreeAsync(dx, m->stream);
    return device_error(e, "forward buffers");
  }
  e = hipMemcpyAsync(dx, x, rows * kIn * sizeof(float), hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(forward_kernel, dim3((unsigned)((rows + 15) / 16)), dim3(64), 0, m->stream, m->packed, dx, dy,
                       (uint32_t)rows);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(y, dy, rows * kOut * sizeof(float), hipMemcpyDeviceToHost, m->stream);
  (void)hipFreeAsync(dx, m->stream);
  (void)hipFreeAsync(dy, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  return e == hipSuccess ? PW_OK : device_error(e, "forward");
}

pw_status pw_mlp_gradients(pw_mlp* m, const float* x, const float* t, size_t rows, float* grads, float* loss) {
  if (!m || !x || !t || !grads || rows == 0 || rows > 0x7fffffffull) {
    set_error("Gradients need rows > 0 samples, targets and a PW_MLP_PARAMS output.");
    return PW_EINVAL;
  }
  float *dx = nullptr, *dt = nullptr, *dg = nullptr;
  hipError_t e;
  const size_t bytes = (rows * (kIn + kOut) + kParams + 1) * sizeof(float);
  if ((e = hipMallocAsync((void**)&dx, bytes, m->stream)) != hipSuccess) return device_error(e, "gradient buffers");
  dt = dx + rows * kIn;
  dg = dt + rows * kOut;
  e = hipMemcpyAsync(dx, x, rows * kIn * sizeof(float), hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dt, t, rows * kOut * sizeof(float), hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) {
    const uint32_t blocks = launch_grad(m, dx, dt, rows, (uint32_t)rows, false, 0, 0);
    launch_update(m, kGradient, blocks, 0.0f, dg, dg + kParams, 1.0f / (3.0f * (float)rows));
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(grads, dg, kParams * sizeof(float), hipMemcpyDeviceToHost, m->stream);
  float l = 0.0f;
  if (e == hipSuccess) e = hipMemcpyAsync(&l, dg + kParams, sizeof(float), hipMemcpyDeviceToHost, m->stream);
  (void)hipFreeAsync(dx, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (e != hipSuccess) return device_error(e, "gradients");
  if (loss) *loss = l;
  return PW_OK;
}

pw_status pw_mlp_set_data(pw_mlp* m, const float* x, const float* t, size_t rows) {
  if (!m || !x || !t || rows == 0 || rows > 0x7fffffffull) {
    set_error("Training data needs rows > 0 samples and targets.");
    return PW_EINVAL;
  }
  hipError_t e;
  if (m->data_x) (void)hipFree(m->data_x);
  if (m->data_t) (void)hipFree(m->data_t);
  m->data_x = m->data_t = nullptr;
  m->n_data = 0;
  if ((e = hipMalloc((void**)&m->data_x, rows * kIn * sizeof(float))) != hipSuccess ||
      (e = hipMalloc((void**)&m->data_t, rows * kOut * sizeof(float))) != hipSuccess)
    return device_error(e, "training data");
  if ((e = hipMemcpyAsync(m->data_x, x, rows * kIn * sizeof(float), hipMemcpyHostToDevice, m->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(m->data_t, t, rows * kOut * sizeof(float), hipMemcpyHostToDevice, m->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(m->stream)) != hipSuccess)
    return device_error(e, "training data");
  m->n_data = rows;
  return PW_OK;
}

pw_status pw_mlp_train(pw_mlp* m, uint32_t steps, uint32_t batch, float lr, uint32_t seed, float* loss_history) {
  if (!m || m->n_data == 0 || !(lr > 0.0f)) {
    set_error("Training needs data (pw_mlp_set_data) and a positive learning rate.");
    return PW_EINVAL;
  }
  if (steps == 0) return PW_OK;
  const bool full = batch == 0;
  const uint32_t b = full ? (uint32_t)m->n_data : batch;
  float* losses = nullptr;
  hipError_t e = hipMallocAsync((void**)&losses, (size_t)steps * sizeof(float), m->stream);
  if (e != hipSuccess) return device_error(e, "training losses");
  for (uint32_t k = 0; k < steps && e == hipSuccess; ++k) {
    m->adam_step += 1;
    const uint32_t blocks = launch_grad(m, m->data_x, m->data_t, m->n_data, b, !full, seed, (uint32_t)m->adam_step);
    launch_update(m, kAdam, blocks, lr, nullptr, losses + k, 1.0f / (3.0f * (float)b));
    e = hipGetLastError();
  }
  if (e == hipSuccess && loss_history)
    e = hipMemcpyAsync(loss_history, losses, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost, m->stream);
  (void)hipFreeAsync(losses, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  return e == hipSuccess ? PW_OK : device_error(e, "training");
}

}  // extern "C"
