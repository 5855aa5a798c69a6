// particles.hip -- streakline particles advected by trilinear sampling of the live velocity.
//
// Follows src/render/shaders/particles_update.wgsl:1-60 (one lane per particle: trilinear
// velocity of the macro field times the solver steps, displacement capped at one cell,
// respawn at x = 2 through the lowbias32 hash when the particle leaves the fluid, reaches the
// outlet or ages past 4000 frames) and src/render/particle-layout.ts:14-112 (seed region,
// inlet candidates, initial records). Particles read the solver's macro buffer and flags on
// the solver's own stream, so an update always sees the macros of the last step batch.
// Arithmetic is f32 with WGSL's mix(a, b, t) = a (1 - t) + b t and no contraction.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "lattice.h"
#include "solver_internal.h"

namespace pw {
namespace part {

constexpr uint32_t kMaxAge = 4000;         // particle-layout.ts:8
constexpr uint32_t kMaxCount = 1u << 24;   // this library's limit; the reference caps at 65536

struct Field {
  const float4* macro;
  const uint32_t* flags;
  uint32_t nx, ny, nz;
};

__host__ __device__ __forceinline__ uint32_t lowbias32(uint32_t h) {
  h = (h ^ (h >> 16)) * 0x7feb352du;
  h = (h ^ (h >> 15)) * 0x846ca68bu;
  return h ^ (h >> 16);
}

__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

__device__ __forceinline__ uint32_t cell_of(const Field& f, uint32_t x, uint32_t y, uint32_t z) {
  return x + f.nx * (y + f.ny * z);
}

// particles_update.wgsl:22-25
__device__ __forceinline__ bool fluid_at(const Field& f, float px, float py, float pz) {
  if (px < 0.0f || py < 0.0f || pz < 0.0f) return false;
  if (px >= (float)(f.nx - 1) || py >= (float)(f.ny - 1) || pz >= (float)(f.nz - 1)) return false;
  const uint32_t x = (uint32_t)floorf(px + 0.5f), y = (uint32_t)floorf(py + 0.5f), z = (uint32_t)floorf(pz + 0.5f);
  return (f.flags[cell_of(f, x, y, z)] & 7u) == 0u;
}

// particles_update.wgsl:26-36
__device__ __forceinline__ float3 velocity_at(const Field& f, float px, float py, float pz) {
  const float lim[3] = {(float)(f.nx - 1), (float)(f.ny - 1), (float)(f.nz - 1)};
  const float p[3] = {px, py, pz};
  float w[3];
  uint32_t lo[3], hi[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float pt = fminf(fmaxf(p[k], 0.0f), lim[k]);
    const float fl = floorf(pt);
    lo[k] = (uint32_t)fl;
    const uint32_t top = (uint32_t)lim[k];
    hi[k] = lo[k] + 1 < top ? lo[k] + 1 : top;
    w[k] = pt - fl;
  }
  const float4 m000 = f.macro[cell_of(f, lo[0], lo[1], lo[2])], m100 = f.macro[cell_of(f, hi[0], lo[1], lo[2])];
  const float4 m010 = f.macro[cell_of(f, lo[0], hi[1], lo[2])], m110 = f.macro[cell_of(f, hi[0], hi[1], lo[2])];
  const float4 m001 = f.macro[cell_of(f, lo[0], lo[1], hi[2])], m101 = f.macro[cell_of(f, hi[0], lo[1], hi[2])];
  const float4 m011 = f.macro[cell_of(f, lo[0], hi[1], hi[2])], m111 = f.macro[cell_of(f, hi[0], hi[1], hi[2])];
  auto tri = [&](float a000, float a100, float a010, float a110, float a001, float a101, float a011, float a111) {
    const float x00 = mixf(a000, a100, w[0]), x10 = mixf(a010, a110, w[0]);
    const float x01 = mixf(a001, a101, w[0]), x11 = mixf(a011, a111, w[0]);
    return mixf(mixf(x00, x10, w[1]), mixf(x01, x11, w[1]), w[2]);
  };
  return make_float3(tri(m000.y, m100.y, m010.y, m110.y, m001.y, m101.y, m011.y, m111.y),
                     tri(m000.z, m100.z, m010.z, m110.z, m001.z, m101.z, m011.z, m111.z),
                     tri(m000.w, m100.w, m010.w, m110.w, m001.w, m101.w, m011.w, m111.w));
}

struct UpdateParams {
  uint32_t count, seed;
  float steps;
  float cs[4];       // seed cross-section (y0, z0, y1, z1)
  float fallback[2]; // (y, z) of the first fluid inlet node
};

// particles_update.wgsl:46-60; a particle record is (pos.xyz, age, prev.xyz, generation).
__global__ __launch_bounds__(256) void update_kernel(float4* __restrict__ particles, Field f, UpdateParams p) {
  const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= p.count) return;
  const float4 pos = particles[2 * (size_t)id], prev = particles[2 * (size_t)id + 1];
  const float3 v = velocity_at(f, pos.x, pos.y, pos.z);
  float dx = v.x * p.steps, dy = v.y * p.steps, dz = v.z * p.steps;
  const float len = sqrtf(dx * dx + dy * dy + dz * dz);
  const float div = fmaxf(1.0f, len);
  dx /= div;
  dy /= div;
  dz /= div;
  const float qx = pos.x + dx, qy = pos.y + dy, qz = pos.z + dz;
  const float age = pos.w + 1.0f;
  if (qx >= (float)(f.nx - 2) || !fluid_at(f, qx, qy, qz) || age > (float)kMaxAge) {
    // particles_update.wgsl:37-45
    const float generation = prev.w + 1.0f;
    const uint32_t h = lowbias32(id + p.seed + (uint32_t)generation * 0x9e3779b9u);
    const float r0 = (float)(h >> 8) / 16777216.0f, r1 = (float)(lowbias32(h) >> 8) / 16777216.0f;
    float y = mixf(p.cs[0], p.cs[2], r0), z = mixf(p.cs[1], p.cs[3], r1);
    if (!fluid_at(f, 2.0f, y, z)) {
      y = p.fallback[0];
      z = p.fallback[1];
    }
    particles[2 * (size_t)id] = make_float4(2.0f, y, z, 0.0f);
    particles[2 * (size_t)id + 1] = make_float4(2.0f, y, z, generation);
  } else {
    particles[2 * (size_t)id] = make_float4(qx, qy, qz, age);
    particles[2 * (size_t)id + 1] = make_float4(pos.x, pos.y, pos.z, prev.w);
  }
}

// particle-layout.ts:39-56
bool needs_reseed(double x, double y, double z, double age, const pw_layout& L, uint32_t flag) {
  return !std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z) || x < 0 || x >= (double)L.nx - 2 ||
         y < 0 || y >= (double)L.ny - 1 || z < 0 || z >= (double)L.nz - 1 || age > kMaxAge ||
         (flag & kTypeMask) != kFluid;
}

double js_round(double v) { return std::floor(v + 0.5); }

}  // namespace part
}  // namespace pw

using namespace pw;
using namespace pw::part;

struct pw_particles {
  pw_solver* solver = nullptr;  // the bound solver, or null when bound to a batched tunnel
  Field field{};                // the macro field and flags the particles sample
  hipStream_t stream = nullptr;  // the stream those are written on
  uint32_t count = 0, seed = 0;
  float region[4] = {0, 0, 0, 0};
  float fallback[3] = {0, 0, 0};
  float4* data = nullptr;
  uint32_t* owned_flags = nullptr;  // the device copy of a caller's flags (field-bound particles)
  bool owns_stream = false;         // the stream was created for these particles
};

extern "C" {

namespace {

// particle-layout.ts:14-36 and :66-112 with the count limit of this library, for the field f
// written on `stream`.
// particle-layout.ts:14-36 (seed region) and the count limit; host checks only.
pw_status seed_arguments(const pw_layout& L, uint32_t count, const double body_min[3], const double body_max[3],
                         double region[4]) {
  const double ey = body_max[1] - body_min[1], ez = body_max[2] - body_min[2];
  region[0] = std::fmax(1, body_min[1] - ey * 0.15);
  region[1] = std::fmax(1, body_min[2] - ez * 0.15);
  region[2] = std::fmin((double)L.ny - 2, body_max[1] + ey * 0.15);
  region[3] = std::fmin((double)L.nz - 2, body_max[2] + ez * 0.15);
  bool finite = true;
  for (int k = 0; k < 3; ++k) finite = finite && std::isfinite(body_min[k]) && std::isfinite(body_max[k]);
  if (L.nx < 5 || L.ny < 5 || L.nz < 5 || !finite || !(ey > 0) || !(ez > 0) || region[0] >= region[2] ||
      region[1] >= region[3]) {
    set_error("Particle seeds need a positive cross-section inside a grid of at least five cells per axis.");
    return PW_EINVAL;
  }
  if (count < 1 || count > kMaxCount) {
    set_error("Particle count, seed or geometry buffer is invalid.");
    return PW_EINVAL;
  }
  return PW_OK;
}

pw_status create_particles(const Field& f, hipStream_t stream, uint32_t count, uint32_t seed,
                           const double body_min[3], const double body_max[3], pw_particles** out) {
  const pw_layout L{f.nx, f.ny, f.nz};
  double region[4];
  if (pw_status st = seed_arguments(L, count, body_min, body_max, region)) return st;
  // Inlet candidates need the cell flags: read back the plane x = 2 rows once.
  std::vector<uint32_t> flags((size_t)L.nx * L.ny * L.nz);
  hipError_t e = hipMemcpyAsync(flags.data(), f.flags, flags.size() * 4, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return device_error(e, "particle flags");
  std::vector<float> inlet;
  for (long z = (long)js_round(region[1]); z <= (long)js_round(region[3]); ++z)
    for (long y = (long)js_round(region[0]); y <= (long)js_round(region[2]); ++y) {
      const float py = (float)std::fmax(region[0], std::fmin((double)y, region[2]));
      const float pz = (float)std::fmax(region[1], std::fmin((double)z, region[3]));
      const long cy = (long)js_round(py), cz = (long)js_round(pz);
      const size_t cell = 2 + (size_t)L.nx * ((size_t)cy + (size_t)L.ny * (size_t)cz);
      const uint32_t flag = cell < flags.size() ? flags[cell] : (uint32_t)kSolid;
      if (needs_reseed(2, py, pz, 0, L, flag)) continue;
      inlet.insert(inlet.end(), {2.0f, py, pz});
    }
  if (inlet.empty()) {
    set_error("The particle seed region has no fluid inlet cells.");
    return PW_EINVAL;
  }
  const size_t n_inlet = inlet.size() / 3;
  std::vector<float4> host(2 * (size_t)count);
  for (uint32_t id = 0; id < count; ++id) {
    const float* p = inlet.data() + 3 * (lowbias32(id + seed) % n_inlet);
    host[2 * (size_t)id] = make_float4(p[0], p[1], p[2], 0.0f);
    host[2 * (size_t)id + 1] = make_float4(p[0], p[1], p[2], 0.0f);
  }
  pw_particles* h = new (std::nothrow) pw_particles();
  if (!h) {
    set_error("Host allocation failed.");
    return PW_ENOMEM;
  }
  h->field = f;
  h->stream = stream;
  h->count = count;
  h->seed = seed;
  for (int k = 0; k < 4; ++k) h->region[k] = (float)region[k];
  std::memcpy(h->fallback, inlet.data(), sizeof h->fallback);
  if ((e = hipMalloc((void**)&h->data, host.size() * sizeof(float4))) != hipSuccess) {
    delete h;
    return device_error(e, "particle allocation");
  }
  e = hipMemcpyAsync(h->data, host.data(), host.size() * sizeof(float4), hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) {
    (void)hipFree(h->data);
    delete h;
    return device_error(e, "particle upload");
  }
  *out = h;
  return PW_OK;
}

}  // namespace

pw_status pw_particles_create(pw_solver* s, uint32_t count, uint32_t seed, const double body_min[3],
                              const double body_max[3], pw_particles** out) {
  if (!solver_live(s) || !out || !body_min || !body_max) {
    set_error("Null argument.");
    return PW_EINVAL;
  }
  *out = nullptr;
  const Field f{s->macro, s->geometry, s->layout.nx, s->layout.ny, s->layout.nz};
  pw_status st = create_particles(f, s->stream, count, seed, body_min, body_max, out);
  if (st == PW_OK) (*out)->solver = s;
  return st;
}

pw_status pw_particles_create_batch(pw_batch* b, uint32_t tunnel, uint32_t count, uint32_t seed,
                                    const double body_min[3], const double body_max[3], pw_particles** out) {
  if (!b || !out || !body_min || !body_max) {
    set_error("Null argument.");
    return PW_EINVAL;
  }
  *out = nullptr;
  const float4* macro = nullptr;
  const uint32_t* geo = nullptr;
  pw_layout L{};
  hipStream_t stream = nullptr;
  if (!batch_field(b, tunnel, &macro, &geo, &L, &stream)) {
    set_error("Invalid tunnel index.");
    return PW_EINVAL;
  }
  return create_particles(Field{macro, geo, L.nx, L.ny, L.nz}, stream, count, seed, body_min, body_max, out);
}

// createParticles(device, format, frame, {grid, bodyBounds, macroBuffer, count, seed})
// (src/render/particles.ts:30-138): the caller's macro buffer on the device and the grid's
// host flags, which are copied to the device. The particles step on a stream of their own.
pw_status pw_particles_create_field(const pw_layout* layout, const uint32_t* geometry, size_t geometry_words,
                                    const float* macro, uint32_t count, uint32_t seed, const double body_min[3],
                                    const double body_max[3], pw_particles** out) {
  if (!layout || !geometry || !macro || !out || !body_min || !body_max) {
    set_error("Null argument.");
    return PW_EINVAL;
  }
  *out = nullptr;
  const size_t N = (size_t)layout->nx * layout->ny * layout->nz;
  if (geometry_words < N) {  // particle-layout.ts:80
    set_error("Particle count, seed or geometry buffer is invalid.");
    return PW_EINVAL;
  }
  double region[4];
  if (pw_status st = seed_arguments(*layout, count, body_min, body_max, region)) return st;
  hipStream_t stream = nullptr;
  uint32_t* flags = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
  if (e != hipSuccess) return device_error(e, "particle stream");
  if ((e = hipMalloc((void**)&flags, N * 4)) != hipSuccess) {
    (void)hipStreamDestroy(stream);
    return device_error(e, "particle flag allocation");
  }
  e = hipMemcpyAsync(flags, geometry, N * 4, hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  pw_status st = e == hipSuccess ? PW_OK : device_error(e, "particle flag upload");
  if (st == PW_OK)
    st = create_particles(Field{(const float4*)macro, flags, layout->nx, layout->ny, layout->nz}, stream, count,
                          seed, body_min, body_max, out);
  if (st != PW_OK) {
    (void)hipFree(flags);
    (void)hipStreamDestroy(stream);
    return st;
  }
  (*out)->owned_flags = flags;
  (*out)->owns_stream = true;
  return PW_OK;
}

// Waits for the particles' queued updates.
pw_status pw_particles_synchronize(pw_particles* h) {
  if (!h) {
    set_error("Null particle handle.");
    return PW_EINVAL;
  }
  hipError_t e = hipStreamSynchronize(h->stream);
  return e == hipSuccess ? PW_OK : device_error(e, "particle synchronize");
}

// particles.update(encoder, steps) (src/render/particles.ts:113-124).
pw_status pw_particles_update(pw_particles* h, float steps) {
  if (!h || (h->solver && !solver_live(h->solver))) {
    set_error("Invalid particle handle.");
    return PW_EINVAL;
  }
  if (!std::isfinite(steps) || steps < 0) {
    set_error("Particle advection needs non-negative finite solver steps.");
    return PW_EINVAL;
  }
  UpdateParams p{h->count, h->seed, steps, {h->region[0], h->region[1], h->region[2], h->region[3]},
                 {h->fallback[1], h->fallback[2]}};
  hipLaunchKernelGGL(update_kernel, dim3((h->count + 255) / 256), dim3(256), 0, h->stream, h->data, h->field, p);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? PW_OK : device_error(e, "particle update");
}

pw_status pw_particles_read(pw_particles* h, float* out, size_t out_floats) {
  if (!h || !out || out_floats < 8 * (size_t)h->count) {
    set_error("Particle output needs 8 floats per particle.");
    return PW_EINVAL;
  }
  hipError_t e = hipMemcpyAsync(out, h->data, 32 * (size_t)h->count, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  return e == hipSuccess ? PW_OK : device_error(e, "particle readback");
}

pw_status pw_particles_write(pw_particles* h, const float* in, size_t in_floats) {
  if (!h || !in || in_floats != 8 * (size_t)h->count) {
    set_error("Particle input must hold 8 floats per particle.");
    return PW_EINVAL;
  }
  hipError_t e = hipMemcpyAsync(h->data, in, 32 * (size_t)h->count, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  return e == hipSuccess ? PW_OK : device_error(e, "particle upload");
}

pw_status pw_particles_info(const pw_particles* h, float region[4], float fallback[3]) {
  if (!h) {
    set_error("Null particle handle.");
    return PW_EINVAL;
  }
  if (region) std::memcpy(region, h->region, sizeof h->region);
  if (fallback) std::memcpy(fallback, h->fallback, sizeof h->fallback);
  return PW_OK;
}

void pw_particles_destroy(pw_particles* h) {
  if (!h) return;
  if (h->owns_stream) (void)hipStreamSynchronize(h->stream);
  (void)hipFree(h->data);  // hipFree waits for work still using the buffer
  if (h->owned_flags) (void)hipFree(h->owned_flags);
  if (h->owns_stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

}  // extern "C"
