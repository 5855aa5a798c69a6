// closure.hip -- the learned wall closure inside a solver: its per-step launch and the
// feature, output and wall-force readbacks that training and evaluation use.
//
// No reference counterpart (the reference's wall treatment is the wall law of
// src/lbm/shaders/stream_collide.wgsl:1-80, which stays the fallback).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "closure.h"
#include "step_row.h"

namespace pw {

using closure::Norm;

namespace {

inline unsigned closure_blocks(size_t n) { return (unsigned)((n + closure::kCellsPerWave - 1) / closure::kCellsPerWave); }

static_assert(sizeof(closure::WallGeo) == kWallGeoBytes, "record size");

// One record per listed wall cell.
__global__ __launch_bounds__(256) void wall_geo_kernel(const uint32_t* __restrict__ geo, StepParams p,
                                                       const uint32_t* __restrict__ cells, uint32_t n_cells,
                                                       closure::WallGeo* __restrict__ out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n_cells) out[k] = closure::wall_geo_of(geo, p, cells[k]);
}

template <class S>
void launch_closure_s(pw_solver* s, hipStream_t stream, const void* state, const void* net, float4* tau,
                      float* records) {
  Norm norm;
  std::memcpy(&norm, s->closure_norm, sizeof norm);
  if (!records && norm.u_ref == 0.0f) norm.u_ref = 1.0f;
  const bool bz = s->params.walls == PW_WALLS_BOUZIDI;
  auto k = bz ? closure::closure_kernel<S, true> : closure::closure_kernel<S, false>;
  hipLaunchKernelGGL(k, dim3(closure_blocks(s->n_wall)), dim3(64), 0, stream, (const typename S::T*)state,
                     s->geometry, s->sp, (const closure::WallGeo*)s->closure_geo, (uint32_t)s->n_wall,
                     (const mlp::Packed*)net, norm, tau, records);
}

void launch_closure_on(pw_solver* s, hipStream_t stream, const void* state, const void* net, float4* tau,
                       float* records) {
  if (!s->closure_geo || s->n_wall == 0) return;  // the records exist from create on; no wall cells, no work
  switch (s->params.storage) {
    case PW_STORAGE_F16: launch_closure_s<StoreF16>(s, stream, state, net, tau, records); break;
    case PW_STORAGE_BF16: launch_closure_s<StoreBF16>(s, stream, state, net, tau, records); break;
    default: launch_closure_s<StoreF32>(s, stream, state, net, tau, records); break;
  }
}

void launch_closure_any(pw_solver* s, const void* state, const void* net, float4* tau, float* records) {
  launch_closure_on(s, s->stream, state, net, tau, records);
}

template <class S>
void launch_wall_forces_s(pw_solver* s, double* out) {
  const bool bz = s->params.walls == PW_WALLS_BOUZIDI;
  auto k = bz ? closure::wall_force_kernel<S, true> : closure::wall_force_kernel<S, false>;
  hipLaunchKernelGGL(k, dim3((unsigned)((s->n_wall + 255) / 256)), dim3(256), 0, s->stream,
                     (const typename S::T*)s->pops[0], (const typename S::T*)s->pops[1], s->geometry, s->sp,
                     s->wall_cells, (uint32_t)s->n_wall, out);
}

// The closure runs inside the step when asked to and the step has instances that carry it: the
// row path, the REG collision with LES and the reference's boundaries.
bool in_step_closure(const pw_solver* s) {
  return s->closure_on == 1 && s->closure_placement == PW_CLOSURE_IN_STEP && s->sp.rows && s->params.op == PW_OP_REG &&
         s->params.les && !s->sp.bc;
}

// The closure runs beside the step when asked to and the step has instances that carry it: the
// REG collision (a learned solver has LES and the reference's boundaries).
bool alongside_closure(const pw_solver* s) {
  return s->closure_on == 1 && s->closure_placement == PW_CLOSURE_ALONGSIDE && s->params.op == PW_OP_REG &&
         s->params.les && !s->sp.bc;
}

// Blocks of a step launch that carry the closure beside the step: one 16-cell tile per
// wavefront, rounded up to a multiple of the XCD count so the step blocks keep their XCD.
uint32_t alongside_blocks(size_t n_wall, uint32_t block) {
  const uint64_t per_block = block / 64;
  const uint64_t blocks = (closure_blocks(n_wall) + per_block - 1) / per_block;
  return (uint32_t)((blocks + kXcds - 1) / kXcds * kXcds);
}

pw_status check(pw_solver* s) {
  if (!solver_live(s)) {
    set_error("Null or destroyed solver handle.");
    return PW_EINVAL;
  }
  return PW_OK;
}

}  // namespace

void launch_wall_geo(hipStream_t stream, const uint32_t* geo, const StepParams& p, const uint32_t* wall_cells,
                     uint32_t n_wall, void* out) {
  if (n_wall == 0) return;
  hipLaunchKernelGGL(wall_geo_kernel, dim3((n_wall + 255) / 256), dim3(256), 0, stream, geo, p, wall_cells, n_wall,
                     static_cast<closure::WallGeo*>(out));
}

void launch_closure(pw_solver* s, StepParams& p) {
  if (!s->closure_on || s->n_wall == 0) return;
  // Evaluated every step on the row path: inside the step's wall wavefronts.
  if (in_step_closure(s) && s->tile_order && s->order_block == s->block && s->order_chunks == s->xcd_chunks) {
    p.closure_net = s->closure_net;
    p.tile_order = s->tile_order;
    p.closure_slot = s->closure_slot;
    p.closure_geo = s->closure_geo;
    return;
  }
  // closure_on 2: the output was prescribed from the host (pw_debug_set_wall_stress).
  if (alongside_closure(s)) {
    // As OVERLAPPED, with the evaluation in the step's own launch (its first blocks) instead of
    // on a second stream: the output of an evaluation beside the previous step replaces the
    // current one before this step.
    settle_closure(s);
    if (!s->closure_fresh) {
      launch_closure_any(s, s->pops[s->current], s->closure_net, s->closure_tau, nullptr);
      s->closure_fresh = 1;
    } else if (s->steps % s->closure_every == 0) {
      p.closure_net = s->closure_net;
      p.closure_geo = s->closure_geo;
      p.closure_out = s->closure_next;
      p.closure_cells = (uint32_t)s->n_wall;
      p.closure_blocks = alongside_blocks(s->n_wall, s->block);
      s->closure_pending = 2;
    }
  } else if (s->closure_on == 1 && s->closure_placement == PW_CLOSURE_OVERLAPPED) {
    // The evaluation started beside the previous step is joined before this one; the first
    // step of a run evaluates in line. An evaluation due at this step reads the state this step
    // reads, on the side stream, while this step runs on the stream with the previous output;
    // the step after it takes the new output (settle_closure).
    settle_closure(s);
    if (!s->closure_fresh) {
      launch_closure_any(s, s->pops[s->current], s->closure_net, s->closure_tau, nullptr);
      s->closure_fresh = 1;
    } else if (s->steps % s->closure_every == 0) {
      (void)hipEventRecord(s->state_ready, s->stream);
      (void)hipStreamWaitEvent(s->side_stream, s->state_ready, 0);
      launch_closure_on(s, s->side_stream, s->pops[s->current], s->closure_net, s->closure_next, nullptr);
      (void)hipEventRecord(s->closure_done, s->side_stream);
      s->closure_pending = 1;
    }
  } else if (s->closure_on == 1 && (!s->closure_fresh || s->steps % s->closure_every == 0)) {
    launch_closure_any(s, s->pops[s->current], s->closure_net, s->closure_tau, nullptr);
    s->closure_fresh = 1;
  }
  p.closure_tau = s->closure_tau;
  p.closure_slot = s->closure_slot;
}

// The momentum sources prescribed through pw_debug_set_wall_stress (w = 2): each listed wall
// cell's stored post-collision populations gain 3 w_i c_i . C, which adds C to its momentum and
// nothing to its mass (sum_i w_i c_i c_i = I / 3). f32 storage only.
__global__ __launch_bounds__(256) void source_kernel(float* __restrict__ set, StepParams p,
                                                     const uint32_t* __restrict__ cells, uint32_t n_cells,
                                                     const float4* __restrict__ source) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_cells) return;
  const float4 c = source[k];
  if (!(c.w > 0.0f)) return;
  float* cell = set + pop_cell(p, cells[k]);
#pragma unroll
  for (int i = 1; i < kQ; ++i) cell[(size_t)i * p.run] += 3.0f * wf(i) * cdot(i, c.x, c.y, c.z);
}

void launch_sources(pw_solver* s, void* set) {
  if (!s->n_wall || !s->closure_source) return;
  hipLaunchKernelGGL(source_kernel, dim3((unsigned)((s->n_wall + 255) / 256)), dim3(256), 0, s->stream,
                     static_cast<float*>(set), s->sp, s->wall_cells, (uint32_t)s->n_wall, s->closure_source);
}

void settle_closure(pw_solver* s) {
  if (!s->closure_pending) return;
  if (s->closure_pending == 1) (void)hipStreamWaitEvent(s->stream, s->closure_done, 0);
  std::swap(s->closure_tau, s->closure_next);
  s->closure_pending = 0;
}

hipError_t closure_tile_order(pw_solver* s) {
  if (s->params.wall_model != PW_WALL_LEARNED || !s->sp.rows) return hipSuccess;
  // The launch's grid (launch_step_t): blocks rounded up to whole chunks on the 8 XCDs.
  const uint64_t blocks = (s->cells + s->block - 1) / s->block;
  const uint32_t parts = s->xcd_chunks;
  const uint64_t per_xcd = ((blocks + kXcds - 1) / kXcds + parts - 1) / parts * parts;
  const uint64_t grid = kXcds * per_xcd;
  std::vector<uint32_t> cells(s->n_wall);
  hipError_t e = s->n_wall ? hipMemcpy(cells.data(), s->wall_cells, s->n_wall * 4, hipMemcpyDeviceToHost) : hipSuccess;
  if (e != hipSuccess) return e;
  std::vector<uint8_t> wall(grid, 0);
  for (uint32_t c : cells) wall[c / s->block] = 1;
  // The logical blocks the launch reaches first are each XCD's first chunk in order; give them
  // the wall blocks, in lattice order, then every other block in lattice order.
  std::vector<uint32_t> first, rest;
  for (uint64_t t = 0; t < grid; ++t) (wall[t] ? first : rest).push_back((uint32_t)t);
  std::vector<uint32_t> order(grid);
  const uint64_t chunk = per_xcd / parts;
  uint64_t k = 0;
  std::vector<uint32_t> by_dispatch;  // logical block of the b-th dispatched block
  by_dispatch.reserve(grid);
  for (uint64_t b = 0; b < grid; ++b) {
    const uint64_t local = b / kXcds;
    by_dispatch.push_back((uint32_t)(((local / chunk) * kXcds + b % kXcds) * chunk + local % chunk));
  }
  for (uint32_t t : first) order[by_dispatch[k++]] = t;
  for (uint32_t t : rest) order[by_dispatch[k++]] = t;
  if (s->tile_order) (void)hipFree(s->tile_order);
  s->tile_order = nullptr;
  if ((e = hipMalloc((void**)&s->tile_order, grid * 4)) != hipSuccess) return e;
  if ((e = hipMemcpy(s->tile_order, order.data(), grid * 4, hipMemcpyHostToDevice)) != hipSuccess) return e;
  s->order_block = s->block;
  s->order_chunks = s->xcd_chunks;
  return hipSuccess;
}

}  // namespace pw

using namespace pw;

extern "C" {

size_t pw_solver_wall_cell_count(const pw_solver* s) { return solver_live(s) ? s->n_wall : 0; }

pw_status pw_solver_wall_cells(pw_solver* s, uint32_t* out, size_t n) {
  if (pw_status st = check(s)) return st;
  if (!out || n < s->n_wall) {
    set_error("Wall cell output needs pw_solver_wall_cell_count entries.");
    return PW_EINVAL;
  }
  hipError_t e = s->n_wall ? hipMemcpyAsync(out, s->wall_cells, s->n_wall * 4, hipMemcpyDeviceToHost, s->stream) : hipSuccess;
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  return e == hipSuccess ? PW_OK : device_error(e, "wall cell readback");
}

pw_status pw_solver_closure_features(pw_solver* s, float u_ref, float* out, size_t n_floats) {
  if (pw_status st = check(s)) return st;
  if (!out || n_floats < s->n_wall * closure::kRecord || !(u_ref > 0.0f)) {
    set_error("Feature output needs 15 floats per wall cell and a positive velocity unit.");
    return PW_EINVAL;
  }
  if (s->n_wall == 0) return PW_OK;
  float* records = nullptr;
  hipError_t e = hipMallocAsync((void**)&records, s->n_wall * closure::kRecord * sizeof(float), s->stream);
  if (e != hipSuccess) return device_error(e, "feature buffer");
  const float keep = s->closure_norm[2 * closure::kFeatures + 2];
  s->closure_norm[2 * closure::kFeatures + 2] = u_ref;
  launch_closure_any(s, s->pops[s->current], nullptr, nullptr, records);
  s->closure_norm[2 * closure::kFeatures + 2] = keep;
  e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, records, s->n_wall * closure::kRecord * sizeof(float), hipMemcpyDeviceToHost, s->stream);
  (void)hipFreeAsync(records, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  return e == hipSuccess ? PW_OK : device_error(e, "features");
}

pw_status pw_solver_closure_output(pw_solver* s, float* out, size_t n_floats) {
  if (pw_status st = check(s)) return st;
  if (!s->closure_tau || !out || n_floats < 4 * s->n_wall) {
    set_error("Closure output needs a learned-closure solver and 4 floats per wall cell.");
    return PW_EINVAL;
  }
  // Evaluated inside the step, the output stays in the step; evaluate it on the current state.
  if (in_step_closure(s) && s->n_wall)
    launch_closure_any(s, s->pops[s->current], s->closure_net, s->closure_tau, nullptr);
  hipError_t e = s->n_wall ? hipMemcpyAsync(out, s->closure_tau, s->n_wall * sizeof(float4), hipMemcpyDeviceToHost, s->stream)
                           : hipSuccess;
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  return e == hipSuccess ? PW_OK : device_error(e, "closure output readback");
}

pw_status pw_solver_wall_forces(pw_solver* s, double* out, size_t n_doubles) {
  if (pw_status st = check(s)) return st;
  if (!out || n_doubles < 3 * s->n_wall) {
    set_error("Wall force output needs 3 doubles per wall cell.");
    return PW_EINVAL;
  }
  if (s->n_wall == 0) return PW_OK;
  double* dev = nullptr;
  hipError_t e = hipMallocAsync((void**)&dev, s->n_wall * 3 * sizeof(double), s->stream);
  if (e != hipSuccess) return device_error(e, "wall force buffer");
  switch (s->params.storage) {
    case PW_STORAGE_F16: launch_wall_forces_s<StoreF16>(s, dev); break;
    case PW_STORAGE_BF16: launch_wall_forces_s<StoreBF16>(s, dev); break;
    default: launch_wall_forces_s<StoreF32>(s, dev); break;
  }
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, dev, s->n_wall * 3 * sizeof(double), hipMemcpyDeviceToHost, s->stream);
  (void)hipFreeAsync(dev, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  return e == hipSuccess ? PW_OK : device_error(e, "wall forces");
}

pw_status pw_solver_set_closure(pw_solver* s, const pw_mlp* net, const float* norm, size_t n_norm) {
  if (pw_status st = check(s)) return st;
  if (s->params.wall_model != PW_WALL_LEARNED || !s->closure_net) {
    set_error("The solver was not created with the learned wall model.");
    return PW_EINVAL;
  }
  if (!net || !norm || n_norm != 2 * closure::kFeatures + 3) {
    set_error("A closure needs a network and 27 normalisation floats (12 means, 12 scales, output "
              "scale, limit, velocity unit).");
    return PW_EINVAL;
  }
  for (size_t k = 0; k < n_norm; ++k)
    if (!std::isfinite(norm[k])) {
      set_error("Closure normalisation must be finite.");
      return PW_EINVAL;
    }
  if (!(norm[2 * closure::kFeatures + 2] > 0.0f) || !(norm[2 * closure::kFeatures + 1] > 0.0f)) {
    set_error("Closure velocity unit and limit must be positive.");
    return PW_EINVAL;
  }
  char block_norm[closure::kNetBlockBytes - closure::kNormOffset] = {};
  std::memcpy(block_norm, norm, n_norm * sizeof(float));
  hipError_t e = hipMemcpy(s->closure_net, mlp_packed(net), mlp_packed_bytes(), hipMemcpyDeviceToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(static_cast<char*>(s->closure_net) + closure::kNormOffset, block_norm, sizeof block_norm,
                  hipMemcpyHostToDevice);
  if (e == hipSuccess) e = closure_tile_order(s);
  if (e != hipSuccess) return device_error(e, "closure network copy");
  std::memcpy(s->closure_norm, norm, n_norm * sizeof(float));
  s->closure_on = 1;
  s->closure_sources = 0;
  s->closure_fresh = 0;
  return PW_OK;
}

pw_status pw_solver_set_closure_placement(pw_solver* s, uint32_t placement) {
  if (pw_status st = check(s)) return st;
  if (placement > PW_CLOSURE_ALONGSIDE) {
    set_error("Unknown closure placement.");
    return PW_EINVAL;
  }
  s->closure_placement = placement;
  s->closure_fresh = 0;
  return PW_OK;
}

pw_status pw_debug_set_wall_stress(pw_solver* s, const float* stress, size_t n_floats) {
  if (pw_status st = check(s)) return st;
  if (s->params.wall_model != PW_WALL_LEARNED || !s->closure_tau) {
    set_error("The solver was not created with the learned wall model.");
    return PW_EINVAL;
  }
  if (!stress || n_floats != 4 * s->n_wall) {
    set_error("Wall stress input needs 4 floats per wall cell.");
    return PW_EINVAL;
  }
  // w = 2 entries are momentum sources: the wall collision keeps the law for them (w = 0 in
  // the stress buffer) and a pass after each step adds them (launch_sources).
  std::vector<float> shear(stress, stress + n_floats), source(n_floats, 0.0f);
  bool sources = false;
  for (size_t k = 0; k < s->n_wall; ++k)
    if (stress[4 * k + 3] > 1.5f) {
      for (int c = 0; c < 3; ++c) source[4 * k + c] = stress[4 * k + c];
      source[4 * k + 3] = 1.0f;
      shear[4 * k + 3] = 0.0f;
      sources = true;
    }
  if (sources && s->params.storage != PW_STORAGE_F32) {
    set_error("Momentum sources need f32 storage.");
    return PW_EINVAL;
  }
  if (s->n_wall) {
    hipError_t e = hipMemcpyAsync(s->closure_tau, shear.data(), n_floats * sizeof(float), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess && sources && !s->closure_source)
      e = hipMalloc((void**)&s->closure_source, s->n_wall * sizeof(float4));
    if (e == hipSuccess && sources)
      e = hipMemcpyAsync(s->closure_source, source.data(), n_floats * sizeof(float), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return device_error(e, "wall stress upload");
  }
  s->closure_on = 2;
  s->closure_sources = sources ? 1 : 0;
  s->closure_fresh = 1;
  return PW_OK;
}

pw_status pw_solver_set_closure_interval(pw_solver* s, uint32_t every) {
  if (pw_status st = check(s)) return st;
  if (every == 0) {
    set_error("The closure interval must be at least one step.");
    return PW_EINVAL;
  }
  s->closure_every = every;
  return PW_OK;
}

}  // extern "C"
