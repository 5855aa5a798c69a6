// solver.hip -- pw_solver: the C-ABI step / force / readback API on one HIP stream.
//
// Host semantics follow src/lbm/gpu-solver.ts:158-465 (create, reset, step with macro on the
// last dispatch of a batch, linear inlet ramp) and src/lbm/gpu-force.ts:45-216 (one force
// slot that is skipped while occupied and returns busy when empty).
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <unordered_set>
#include <vector>

#include "force_kernels.h"
#include "step_launch.h"

namespace pw {

static thread_local std::string g_error;

// Handles returned by pw_solver_create and not yet destroyed. Entry points check membership
// rather than a flag inside the handle, so a destroyed handle is never dereferenced.
static std::mutex g_live_mutex;
static std::unordered_set<const pw_solver*> g_live;

bool solver_live(const pw_solver* s) {
  if (!s) return false;
  std::lock_guard<std::mutex> lock(g_live_mutex);
  return g_live.count(s) != 0;
}

void set_error(const std::string& message) { g_error = message; }

pw_status device_error(hipError_t e, const char* what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? PW_ENOMEM : PW_EDEVICE;
}

// Clears the guard word (a kernel rather than a memset, so a captured step stays valid).
__global__ void guard_clear_kernel(uint32_t* flag) { *flag = 0u; }

__global__ __launch_bounds__(kForceBlock) void force_reduce_kernel(const double* __restrict__ partials,
                                                                   uint32_t n_partials,
                                                                   double* __restrict__ result) {
  force_reduce_block(partials, n_partials, result);
}

namespace {

constexpr int kBlock = 256;

inline bool wall_kind(uint32_t k) { return k >= kSolid && k <= kFloor; }

// geometry-contract.ts:39-66: flags must equal kind | pull wall mask over wrapped
// neighbours, q offsets must be the running count of links, the length must match.
bool validate_geometry(const pw_layout& L, const uint32_t* geo, size_t words) {
  const size_t N = (size_t)L.nx * L.ny * L.nz;
  if (words < 2 * N) {
    set_error("Geometry must match the solver layout and contain flags and q offsets.");
    return false;
  }
  size_t fractions = 0;
  for (size_t cell = 0; cell < N; ++cell) {
    const uint32_t flag = geo[cell];
    const uint32_t kind = flag & kTypeMask;
    if (kind > kEq) {
      set_error("Unknown geometry cell kind at cell " + std::to_string(cell) + ".");
      return false;
    }
    uint32_t mask = 0;
    if (kind == kFluid) {
      const long x = cell % L.nx, y = (cell / L.nx) % L.ny, z = cell / ((size_t)L.nx * L.ny);
      for (int i = 1; i < kQ; ++i) {
        const long sx = (x - cx(i) + L.nx) % L.nx, sy = (y - cy(i) + L.ny) % L.ny,
                   sz = (z - cz(i) + L.nz) % L.nz;
        if (wall_kind(geo[sx + L.nx * (sy + (size_t)L.ny * sz)] & kTypeMask)) mask |= 1u << (3 + i);
      }
    }
    if (flag != (kind | mask)) {
      set_error("Inconsistent wall mask at cell " + std::to_string(cell) + ".");
      return false;
    }
    const uint32_t offset = geo[N + cell];
    if (offset == kNoQ) continue;
    if (mask == 0 || offset != fractions) {
      set_error("Invalid geometry q offset at cell " + std::to_string(cell) + ".");
      return false;
    }
    fractions += __builtin_popcount(mask);
  }
  if (words != 2 * N + fractions) {
    set_error("Geometry length must match its cell planes and packed wall-link fractions.");
    return false;
  }
  return true;
}

// gpu-solver.ts:80-116: the f32 uniform, including the re-derived f32 viscosity.
bool pack_params(const pw_layout& L, const pw_params& prm, double u_in, StepParams* out) {
  const double cells = (double)L.nx * L.ny * L.nz;
  if (cells * 7 > 4294967295.0) {
    set_error("Population buffers exceed u32 shader indexing.");
    return false;
  }
  if (cells * kQ * (prm.storage == PW_STORAGE_F32 ? 4.0 : 2.0) > 4294967295.0) {
    set_error("A population set must stay below 4 GiB (32-bit byte offsets).");
    return false;
  }
  const float tau0 = (float)prm.tau0;
  const float uin = (float)u_in;
  const float nu = (float)(((double)tau0 - 0.5) / 3);
  const float shader_tau = (float)((float)(3 * (double)nu) + 0.5);
  if (!std::isfinite(tau0) || !std::isfinite(shader_tau) || shader_tau <= 0.5f || !std::isfinite(uin)) {
    set_error("Relaxation time must exceed 0.5 and inlet speed must be finite.");
    return false;
  }
  if (!std::isfinite(prm.slip_floor) || prm.slip_floor < 0.0f) {
    set_error("The slip floor must be finite and non-negative.");
    return false;
  }
  out->nx = L.nx;
  out->ny = L.ny;
  out->nz = L.nz;
  out->cells = (uint32_t)cells;
  out->u_in = uin;
  out->nu = nu;
  out->tau0 = 3.0f * nu + 0.5f;
  out->cs_term = (float)(18 * M_SQRT2 * 0.1 * 0.1);
  out->lambda = 3.0f / 16.0f;
  out->inv_tau0 = 1.0f / out->tau0;
  out->wm0 = 1.0f / (out->lambda / (out->tau0 - 0.5f) + 0.5f);
  out->slip_floor = prm.slip_floor > 0.0f ? prm.slip_floor : prm.storage == PW_STORAGE_F32 ? 1e-7f : 1e-4f;
  out->run = pop_run(L.nx);
  out->x_tile = out->run == 64 ? 18u * 64u : 0u;
  out->div_nx = make_fast_div(L.nx);
  out->div_ny = make_fast_div(L.ny);
  out->rows = L.nx % 64 == 0 ? 1u : 0u;
  out->wave_desc = nullptr;
  out->closure_tau = nullptr;
  out->closure_slot = nullptr;
  out->closure_net = nullptr;
  out->tile_order = nullptr;
  out->closure_geo = nullptr;
  out->closure_out = nullptr;
  out->closure_cells = 0;
  out->closure_blocks = 0;
  out->bc = (prm.sides == PW_SIDES_FREE_SLIP ? kBcFreeSlip : 0u) |
            (prm.outlet == PW_OUTLET_ZERO_GRADIENT ? kBcZeroGradient : 0u);
  return true;
}

}  // namespace

std::vector<uint32_t> wave_descriptors(const pw_layout& L, const uint32_t* flags) {
  const size_t runs = (size_t)L.nx * L.ny * L.nz / 64;
  std::vector<uint32_t> desc(runs, 0u);
  uint32_t walls = 0;  // below kMaxRowWallCells (validate_tunnel)
  for (size_t r = 0; r < runs; ++r) {
    const uint32_t* f = flags + 64 * r;
    uint32_t d = 0;
    for (uint32_t l = 0; l < 64 && d != kWaveGeneral; ++l) {
      if (f[l] == 0u) continue;
      if (l == 0 && f[l] == kEq) d |= kWaveEqFirst;
      else if (l == 63 && f[l] == kEq) d |= kWaveEqLast;
      else d = kWaveGeneral;
    }
    desc[r] = d | (walls << kWaveBaseShift);
    for (uint32_t l = 0; l < 64; ++l) walls += (f[l] & kTypeMask) == kFluid && (f[l] >> 3) != 0;
  }
  return desc;
}

namespace {

// Free-slip sides: no wall cell within two cells of a side, so that the wall masks over wrapped
// neighbours equal those over the step's sources and no wall gather crosses a side.
bool walls_clear_of_sides(const pw_layout& L, const uint32_t* geo) {
  for (uint32_t z = 0; z < L.nz; ++z)
    for (uint32_t y = 0; y < L.ny; ++y) {
      if (y >= 2 && y + 2 < L.ny && z >= 2 && z + 2 < L.nz) continue;
      for (uint32_t x = 0; x < L.nx; ++x)
        if (wall_kind(geo[x + (size_t)L.nx * (y + (size_t)L.ny * z)] & kTypeMask)) {
          set_error("Free-slip sides need every wall cell at least two cells from the sides.");
          return false;
        }
    }
  return true;
}

// Zero-gradient outlet: the EQ cells of the plane x = nx - 1 become fluid cells with the pull
// wall mask of their sources (x = nx for c_x = -1 held at nx - 1) and no wall-link fractions
// (halfway bounce-back); every other word is kept.
void open_outlet(const pw_layout& L, uint32_t sides, std::vector<uint32_t>& geo) {
  const size_t N = (size_t)L.nx * L.ny * L.nz;
  const bool slip = sides == PW_SIDES_FREE_SLIP;
  const uint32_t x = L.nx - 1;
  for (uint32_t z = 0; z < L.nz; ++z)
    for (uint32_t y = 0; y < L.ny; ++y) {
      const size_t cell = x + (size_t)L.nx * (y + (size_t)L.ny * z);
      if ((geo[cell] & kTypeMask) != kEq) continue;
      uint32_t mask = 0;
      for (int i = 1; i < kQ; ++i) {
        const long sx = cx(i) < 0 ? (long)x : (long)x - cx(i);
        long sy = (long)y - cy(i), sz = (long)z - cz(i);
        sy = sy < 0 ? (slip ? 0 : L.ny - 1) : (sy >= (long)L.ny ? (slip ? L.ny - 1 : 0) : sy);
        sz = sz < 0 ? (slip ? 0 : L.nz - 1) : (sz >= (long)L.nz ? (slip ? L.nz - 1 : 0) : sz);
        if (wall_kind(geo[sx + L.nx * (sy + (size_t)L.ny * sz)] & kTypeMask)) mask |= 1u << (3 + i);
      }
      geo[cell] = kFluid | mask;
      geo[N + cell] = kNoQ;
    }
}

// gpu-solver.ts:119-123
double ramp_factor(uint64_t step, uint32_t ramp) {
  if (ramp == 0) return 1;
  if (step == 0) return 0;
  return std::fmin(1.0, (double)step / ramp);
}

inline unsigned blocks_for(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

template <class S>
void launch_init(pw_solver* s, void* out, const StepParams& p) {
  hipLaunchKernelGGL(init_kernel<S>, dim3(blocks_for(s->cells)), dim3(kBlock), 0, s->stream,
                     (typename S::T*)out, s->geometry, s->macro, p);
}

}  // namespace

extern template void launch_step_s<StoreF32>(pw_solver*, const void*, void*, bool, const StepParams&, uint32_t);
extern template void launch_step_s<StoreF16>(pw_solver*, const void*, void*, bool, const StepParams&, uint32_t);
extern template void launch_step_s<StoreBF16>(pw_solver*, const void*, void*, bool, const StepParams&, uint32_t);
extern template void launch_step_s<StoreF16Reference>(pw_solver*, const void*, void*, bool, const StepParams&,
                                                      uint32_t);

namespace {

// One step from the current set into the other; the step count seeds the stores' rounding.
void launch_step(pw_solver* s, bool macro, const StepParams& params) {
  const void* in = s->pops[s->current];
  void* out = s->pops[s->current ^ 1];
  const uint32_t seed = (uint32_t)s->steps;
  StepParams p = params;
  launch_closure(s, p);
  switch (s->params.storage) {
    case PW_STORAGE_F16:
      if (s->rounding == PW_ROUNDING_REFERENCE) launch_step_s<StoreF16Reference>(s, in, out, macro, p, seed);
      else launch_step_s<StoreF16>(s, in, out, macro, p, seed);
      break;
    case PW_STORAGE_BF16: launch_step_s<StoreBF16>(s, in, out, macro, p, seed); break;
    default: launch_step_s<StoreF32>(s, in, out, macro, p, seed); break;
  }
  if (s->closure_on == 2 && s->closure_sources) launch_sources(s, out);
}

template <class S>
void launch_edge(pw_solver* s, void* set, const StepParams& p) {
  hipLaunchKernelGGL(edge_moments_kernel<S>, dim3(blocks_for((size_t)p.ny * p.nz)), dim3(kBlock), 0, s->stream,
                     (typename S::T*)set, p);
}

// Recompute the upstream-outlet moments stored with a set from its populations (after the
// populations were written other than by a step).
void launch_edge_any(pw_solver* s, void* set, const StepParams& p) {
  switch (s->params.storage) {
    case PW_STORAGE_F16: launch_edge<StoreF16>(s, set, p); break;
    case PW_STORAGE_BF16: launch_edge<StoreBF16>(s, set, p); break;
    default: launch_edge<StoreF32>(s, set, p); break;
  }
}

void launch_init_both(pw_solver* s, const StepParams& p) {
  for (int k = 0; k < 2; ++k) {
    switch (s->params.storage) {
      case PW_STORAGE_F16:
        if (s->rounding == PW_ROUNDING_REFERENCE) launch_init<StoreF16Reference>(s, s->pops[k], p);
        else launch_init<StoreF16>(s, s->pops[k], p);
        break;
      case PW_STORAGE_BF16: launch_init<StoreBF16>(s, s->pops[k], p); break;
      default: launch_init<StoreF32>(s, s->pops[k], p); break;
    }
    launch_edge_any(s, s->pops[k], p);
  }
}

// Bytes of one population set: 19 N elements, then one float4 of upstream-outlet moments per row.
size_t set_bytes(const pw_solver* s) {
  return edge_offset_bytes(s->sp.cells, s->elem_bytes) + (size_t)s->sp.ny * s->sp.nz * sizeof(float4);
}

template <class S>
void launch_force_s(pw_solver* s, double* dest) {
  const unsigned blocks = (unsigned)s->n_partials;
  if (blocks > 0) {
    auto k = s->params.walls == PW_WALLS_BOUZIDI ? force_partials_kernel<S, true> : force_partials_kernel<S, false>;
    hipLaunchKernelGGL(k, dim3(blocks), dim3(kForceBlock), 0, s->stream, (const typename S::T*)s->pops[0],
                       (const typename S::T*)s->pops[1], s->geometry, s->boundary, (uint32_t)s->n_boundary,
                       s->partials, s->sp);
  }
  hipLaunchKernelGGL(force_reduce_kernel, dim3(1), dim3(kForceBlock), 0, s->stream, s->partials,
                     (uint32_t)s->n_partials, dest);
}

void launch_force(pw_solver* s, double* dest) {
  switch (s->params.storage) {
    case PW_STORAGE_F16: launch_force_s<StoreF16>(s, dest); break;
    case PW_STORAGE_BF16: launch_force_s<StoreBF16>(s, dest); break;
    default: launch_force_s<StoreF32>(s, dest); break;
  }
}

// Time each step variant at each block size on the handle's own buffers and keep the
// fastest. The caller re-initialises (or restores) the populations afterwards, so tuning
// leaves no trace in the state.
// Parse a PW_STEP_KERNEL value into the solver's variant; false when absent or malformed.
bool forced_variant(pw_solver* s, const char* text) {
  if (!text) return false;
  const std::string v(text);
  const size_t a = v.find('/'), b = a == std::string::npos ? a : v.find('/', a + 1);
  const std::string name = v.substr(0, a);
  uint32_t kernel = 0;
  if (name == "cell") kernel = PW_KERNEL_CELL;
  else if (name == "cell-nt") kernel = PW_KERNEL_CELL_NT;
  else if (name == "cell-nt-stores") kernel = PW_KERNEL_CELL_NT_STORES;
  else return false;
  uint32_t block = s->block, chunks = s->xcd_chunks;
  if (a != std::string::npos) block = (uint32_t)std::strtoul(v.c_str() + a + 1, nullptr, 10);
  if (b != std::string::npos) chunks = (uint32_t)std::strtoul(v.c_str() + b + 1, nullptr, 10);
  if ((block != 64 && block != 128 && block != 256) || !valid_xcd_chunks(chunks)) return false;
  s->kernel = kernel;
  s->block = block;
  s->xcd_chunks = chunks;
  return true;
}

hipError_t autotune(pw_solver* s) {
  const uint32_t kernels[3] = {PW_KERNEL_CELL, PW_KERNEL_CELL_NT, PW_KERNEL_CELL_NT_STORES};
  const uint32_t blocks[3] = {64, 128, 256};
  struct Candidate {
    float ms;
    uint32_t kernel, block, chunks;
  };
  // The fastest of `rounds` timed runs of `reps` steps each, after two warm-up steps.
  auto time_variant = [&](const Candidate& c, int rounds, int reps, float* fastest) -> hipError_t {
    s->kernel = c.kernel;
    s->block = c.block;
    s->xcd_chunks = c.chunks;
    for (int w = 0; w < 2; ++w) {
      launch_step(s, false, s->sp);
      s->current ^= 1;
    }
    for (int round = 0; round < rounds; ++round) {
      hipError_t e = hipEventRecord(s->t0, s->stream);
      if (e != hipSuccess) return e;
      for (int r = 0; r < reps; ++r) {
        launch_step(s, false, s->sp);
        s->current ^= 1;
      }
      if ((e = hipEventRecord(s->t1, s->stream)) != hipSuccess) return e;
      if ((e = hipEventSynchronize(s->t1)) != hipSuccess) return e;
      float ms = 0;
      if ((e = hipEventElapsedTime(&ms, s->t0, s->t1)) != hipSuccess) return e;
      if (round == 0 || ms < *fastest) *fastest = ms;
    }
    return hipSuccess;
  };
  // A screen of every candidate (each timed twice over `reps` steps, the faster run kept), then
  // the four fastest timed again over eight times as many steps in four interleaved rounds (each
  // round times every finalist once, so a drift of the clock or of the memory affects them alike),
  // the fastest run of each counting: the variants lie within a few percent of each other, and
  // one short sample can misorder them.
  const int reps = s->cells >= (1u << 20) ? 12 : 2;
  std::vector<Candidate> screened;
  for (uint32_t parts : kXcdChunkings)
    for (uint32_t k : kernels)
      for (uint32_t b : blocks) {
        // A chunking with fewer than two blocks per chunk would only add empty blocks.
        if (parts > 1 && (uint64_t)parts * 2 * kXcds * b > s->cells) continue;
        Candidate c{0.0f, k, b, parts};
        if (hipError_t e = time_variant(c, 2, reps, &c.ms)) return e;
        screened.push_back(c);
      }
  std::sort(screened.begin(), screened.end(), [](const Candidate& x, const Candidate& y) { return x.ms < y.ms; });
  Candidate best = screened.front();
  if (s->cells >= (1u << 20)) {
    std::vector<Candidate> finalists(screened.begin(), screened.begin() + std::min<size_t>(4, screened.size()));
    std::vector<float> fastest(finalists.size(), 0.0f);
    for (int round = 0; round < 4; ++round)
      for (size_t i = 0; i < finalists.size(); ++i) {
        float ms = 0.0f;
        if (hipError_t e = time_variant(finalists[i], 1, 8 * reps, &ms)) return e;
        if (round == 0 || ms < fastest[i]) fastest[i] = ms;
      }
    for (size_t i = 0; i < finalists.size(); ++i)
      if (i == 0 || fastest[i] < best.ms) {
        best = finalists[i];
        best.ms = fastest[i];
      }
  }
  s->kernel = best.kernel;
  s->block = best.block;
  s->xcd_chunks = best.chunks;
  s->current = 0;
  return hipGetLastError();
}

void free_all(pw_solver* s) {
  if (s->stream) hipStreamSynchronize(s->stream);
  for (void* p : s->pops) if (p) hipFree(p);
  if (s->geometry) hipFree(s->geometry);
  if (s->boundary) hipFree(s->boundary);
  if (s->wall_cells) hipFree(s->wall_cells);
  if (s->wave_desc) hipFree(s->wave_desc);
  if (s->tile_order) hipFree(s->tile_order);
  if (s->closure_slot) hipFree(s->closure_slot);
  if (s->side_stream) hipStreamSynchronize(s->side_stream);
  if (s->closure_tau) hipFree(s->closure_tau);
  if (s->closure_next) hipFree(s->closure_next);
  if (s->closure_source) hipFree(s->closure_source);
  if (s->closure_geo) hipFree(s->closure_geo);
  if (s->state_ready) hipEventDestroy(s->state_ready);
  if (s->closure_done) hipEventDestroy(s->closure_done);
  if (s->side_stream) hipStreamDestroy(s->side_stream);
  if (s->closure_net) hipFree(s->closure_net);
  if (s->macro) hipFree(s->macro);
  if (s->partials) hipFree(s->partials);
  if (s->force_dev) hipFree(s->force_dev);
  if (s->force_host) hipHostFree(s->force_host);
  if (s->guard_host) hipHostFree(s->guard_host);
  if (s->force_event) hipEventDestroy(s->force_event);
  if (s->t0) hipEventDestroy(s->t0);
  if (s->t1) hipEventDestroy(s->t1);
  if (s->stream) hipStreamDestroy(s->stream);
}

pw_status check_handle(pw_solver* s) {
  if (!s) {
    set_error("Null solver handle.");
    return PW_EINVAL;
  }
  if (!solver_live(s)) {
    set_error("The solver has been destroyed (or was never created).");
    return PW_ESTATE;
  }
  return PW_OK;
}

#define PW_TRY(expr, what)                                  \
  do {                                                      \
    hipError_t e_ = (expr);                                 \
    if (e_ != hipSuccess) return ::pw::device_error(e_, what); \
  } while (0)

}  // namespace

pw_status validate_tunnel(const pw_layout& L, const uint32_t* geometry, size_t geometry_words,
                          const uint32_t* boundary_cells, size_t n_boundary, const pw_params& params,
                          StepParams* full) {
  if ((!geometry && geometry_words) || (!boundary_cells && n_boundary)) {
    set_error("Null argument.");
    return PW_EINVAL;
  }
  if (L.nx < 1 || L.ny < 1 || L.nz < 1) {
    set_error("Grid dimensions must be positive safe integers.");
    return PW_EINVAL;
  }
  if (params.storage > PW_STORAGE_BF16 || params.op > PW_OP_REG || params.walls > PW_WALLS_BOUZIDI ||
      params.wall_model > PW_WALL_LEARNED) {
    set_error("Unknown operator, wall, storage or wall-model option.");
    return PW_EINVAL;
  }
  if (params.wall_model == PW_WALL_LEARNED && !params.les) {
    set_error("The learned wall closure replaces the LES wall model and needs LES.");
    return PW_EINVAL;
  }
  if (params.sides > PW_SIDES_FREE_SLIP || params.outlet > PW_OUTLET_ZERO_GRADIENT) {
    set_error("Unknown side or outlet boundary option.");
    return PW_EINVAL;
  }
  if ((params.sides != PW_SIDES_PERIODIC || params.outlet != PW_OUTLET_EQUILIBRIUM) &&
      params.wall_model == PW_WALL_LEARNED) {
    set_error("The learned wall closure runs with periodic sides and the equilibrium outlet only.");
    return PW_EINVAL;
  }
  if (params.outlet == PW_OUTLET_ZERO_GRADIENT && L.nx < 2) {
    set_error("A zero-gradient outlet needs nx >= 2.");
    return PW_EINVAL;
  }
  if (!pack_params(L, params, params.u_in, full)) return PW_EINVAL;
  if (!validate_geometry(L, geometry, geometry_words)) return PW_EINVAL;
  if (params.sides == PW_SIDES_FREE_SLIP && !walls_clear_of_sides(L, geometry)) return PW_EINVAL;
  if (full->rows) {
    uint64_t walls = 0;
    for (size_t c = 0; c < full->cells; ++c) walls += (geometry[c] & kTypeMask) == kFluid && (geometry[c] >> 3) != 0;
    if (walls >= kMaxRowWallCells) {
      set_error("A lattice whose nx is a multiple of 64 holds at most 2^29 - 1 wall cells (fluid cells with wall links).");
      return PW_EINVAL;
    }
  }
  const size_t N = (size_t)L.nx * L.ny * L.nz;
  // gpu-force.ts:76-81
  for (size_t k = 0; k < n_boundary; ++k) {
    const uint32_t cell = boundary_cells[k];
    if (cell >= N || (geometry[cell] & kTypeMask) != kFluid || (geometry[cell] >> 4) == 0) {
      set_error("Boundary cells must index fluid cells with wall links.");
      return PW_EINVAL;
    }
  }
  return PW_OK;
}

}  // namespace pw

using namespace pw;

extern "C" {

const char* pw_last_error_message(void) { return g_error.c_str(); }
int pw_abi_version(void) { return PW_ABI_VERSION; }

int pw_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

pw_status pw_pack_params(const pw_layout* layout, const pw_params* params, int write_macro, uint64_t step, void* out,
                         size_t out_bytes) {
  if (!layout || !params || !out || out_bytes < PW_PARAMS_BYTES) {
    set_error("Null argument or an output smaller than the 48-byte uniform.");
    return PW_EINVAL;
  }
  if (layout->nx < 1 || layout->ny < 1 || layout->nz < 1) {
    set_error("Grid dimensions must be positive safe integers.");
    return PW_EINVAL;
  }
  StepParams sp{};
  if (!pack_params(*layout, *params, params->u_in, &sp)) return PW_EINVAL;
  // struct Params (common.wgsl:31-46): four u32, four f32, four u32.
  const uint32_t head[4] = {sp.nx, sp.ny, sp.nz, sp.cells};
  const float mid[4] = {sp.u_in, sp.nu, sp.cs_term, sp.lambda};
  const uint32_t tail[4] = {params->les ? 1u : 0u, params->walls == PW_WALLS_BOUZIDI ? 1u : 0u, write_macro ? 1u : 0u,
                            (uint32_t)step};
  char* o = static_cast<char*>(out);
  std::memcpy(o, head, 16);
  std::memcpy(o + 16, mid, 16);
  std::memcpy(o + 32, tail, 16);
  return PW_OK;
}

pw_status pw_solver_create(const pw_layout* layout, const uint32_t* geometry, size_t geometry_words,
                           const uint32_t* boundary_cells, size_t n_boundary, const pw_params* params,
                           pw_solver** out) {
  if (!layout || !params || !out) {
    set_error("Null argument.");
    return PW_EINVAL;
  }
  *out = nullptr;
  const pw_layout L = *layout;
  StepParams full{};
  if (pw_status st = validate_tunnel(L, geometry, geometry_words, boundary_cells, n_boundary, *params, &full)) return st;
  // gpu-solver.ts:172-177: a ramp starts from rest.
  StepParams sp = full;
  if (params->ramp_steps > 0) sp.u_in = 0.0f;
  const size_t N = (size_t)L.nx * L.ny * L.nz;
  std::vector<uint32_t> opened;
  if (params->outlet == PW_OUTLET_ZERO_GRADIENT) {
    opened.assign(geometry, geometry + geometry_words);
    open_outlet(L, params->sides, opened);
    geometry = opened.data();
  }
  long long first_fluid = -1;
  for (size_t c = 0; c < N && first_fluid < 0; ++c)
    if ((geometry[c] & kTypeMask) == kFluid) first_fluid = (long long)c;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    set_error("No HIP device is available.");
    return PW_EDEVICE;
  }
  pw_solver* s = new (std::nothrow) pw_solver();
  if (!s) {
    set_error("Host allocation failed.");
    return PW_ENOMEM;
  }
  s->layout = L;
  s->cells = N;
  s->params = *params;
  s->sp = full;
  s->elem_bytes = params->storage == PW_STORAGE_F32 ? 4 : 2;
  s->geometry_words = geometry_words;
  s->n_boundary = n_boundary;
  s->first_fluid = first_fluid;
  s->n_partials = (n_boundary + kForceBlock - 1) / kForceBlock;
  std::vector<uint32_t> wall_list;
  for (size_t cell = 0; cell < N; ++cell)
    if ((geometry[cell] & kTypeMask) == kFluid && (geometry[cell] >> 3) != 0) wall_list.push_back((uint32_t)cell);
  s->n_wall = wall_list.size();
  const bool learned = params->wall_model == PW_WALL_LEARNED;
  hipGetDevice(&s->device);
  auto fail = [&](pw_status st) {
    free_all(s);
    delete s;
    return st;
  };
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking)) != hipSuccess) return fail(device_error(e, "hipStreamCreate"));
  const size_t bytes_per_set = edge_offset_bytes(full.cells, s->elem_bytes) + (size_t)full.ny * full.nz * sizeof(float4);
  for (int k = 0; k < 2; ++k)
    if ((e = hipMalloc(&s->pops[k], bytes_per_set)) != hipSuccess) return fail(device_error(e, "population allocation"));
  if ((e = hipMalloc((void**)&s->geometry, geometry_words * 4)) != hipSuccess) return fail(device_error(e, "geometry allocation"));
  if ((e = hipMalloc((void**)&s->macro, (N + 1) * sizeof(float4))) != hipSuccess) return fail(device_error(e, "macro allocation"));
  if ((e = hipMalloc((void**)&s->boundary, (n_boundary ? n_boundary : 1) * 4)) != hipSuccess) return fail(device_error(e, "boundary allocation"));
  if ((e = hipMalloc((void**)&s->wall_cells, (s->n_wall ? s->n_wall : 1) * 4)) != hipSuccess) return fail(device_error(e, "wall cell allocation"));
  if (full.rows) {
    const std::vector<uint32_t> desc = wave_descriptors(L, geometry);
    if ((e = hipMalloc((void**)&s->wave_desc, desc.size() * 4)) != hipSuccess) return fail(device_error(e, "descriptor allocation"));
    if ((e = hipMemcpy(s->wave_desc, desc.data(), desc.size() * 4, hipMemcpyHostToDevice)) != hipSuccess)
      return fail(device_error(e, "descriptor upload"));
    full.wave_desc = s->wave_desc;
    sp.wave_desc = s->wave_desc;
    s->sp.wave_desc = s->wave_desc;
  }
  // Every solver exports closure features (pw_solver_closure_features), so every solver holds
  // its wall cells' static closure geometry.
  if ((e = hipMalloc(&s->closure_geo, (s->n_wall ? s->n_wall : 1) * kWallGeoBytes)) != hipSuccess)
    return fail(device_error(e, "closure geometry allocation"));
  if (learned) {
    std::vector<uint32_t> slot(N, 0xffffffffu);
    for (size_t k = 0; k < wall_list.size(); ++k) slot[wall_list[k]] = (uint32_t)k;
    if ((e = hipMalloc((void**)&s->closure_slot, N * 4)) != hipSuccess ||
        (e = hipMalloc((void**)&s->closure_tau, (s->n_wall ? s->n_wall : 1) * sizeof(float4))) != hipSuccess ||
        (e = hipMalloc((void**)&s->closure_next, (s->n_wall ? s->n_wall : 1) * sizeof(float4))) != hipSuccess ||
        (e = hipMalloc(&s->closure_net, closure::kNetBlockBytes)) != hipSuccess)
      return fail(device_error(e, "closure allocation"));
    if ((e = hipStreamCreateWithFlags(&s->side_stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&s->state_ready, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&s->closure_done, hipEventDisableTiming)) != hipSuccess)
      return fail(device_error(e, "closure stream"));
    if ((e = hipMemcpy(s->closure_slot, slot.data(), N * 4, hipMemcpyHostToDevice)) != hipSuccess)
      return fail(device_error(e, "closure upload"));
  }
  if ((e = hipMalloc((void**)&s->partials, (s->n_partials ? s->n_partials : 1) * 3 * sizeof(double))) != hipSuccess) return fail(device_error(e, "force partial allocation"));
  if ((e = hipMalloc((void**)&s->force_dev, 4 * sizeof(double))) != hipSuccess) return fail(device_error(e, "force allocation"));
  if ((e = hipHostMalloc((void**)&s->force_host, 4 * sizeof(double), hipHostMallocDefault)) != hipSuccess) return fail(device_error(e, "pinned force slot"));
  s->guard_flag = reinterpret_cast<uint32_t*>(s->macro + N);
  if ((e = hipHostMalloc((void**)&s->guard_host, sizeof(uint32_t), hipHostMallocDefault)) != hipSuccess) return fail(device_error(e, "pinned guard slot"));
  if ((e = hipEventCreateWithFlags(&s->force_event, hipEventDisableTiming)) != hipSuccess) return fail(device_error(e, "event"));
  if ((e = hipEventCreate(&s->t0)) != hipSuccess || (e = hipEventCreate(&s->t1)) != hipSuccess) return fail(device_error(e, "event"));
  if ((e = hipMemcpyAsync(s->geometry, geometry, geometry_words * 4, hipMemcpyHostToDevice, s->stream)) != hipSuccess) return fail(device_error(e, "geometry upload"));
  if (n_boundary && (e = hipMemcpyAsync(s->boundary, boundary_cells, n_boundary * 4, hipMemcpyHostToDevice, s->stream)) != hipSuccess) return fail(device_error(e, "boundary upload"));
  if (s->n_wall && (e = hipMemcpyAsync(s->wall_cells, wall_list.data(), s->n_wall * 4, hipMemcpyHostToDevice, s->stream)) != hipSuccess) return fail(device_error(e, "wall cell upload"));
  launch_wall_geo(s->stream, s->geometry, s->sp, s->wall_cells, (uint32_t)s->n_wall, s->closure_geo);
  launch_init_both(s, full);
  // PW_STEP_KERNEL=<cell|cell-nt|cell-nt-stores>[/<threads per block>[/<chunks per XCD>]]
  // fixes the step variant instead of tuning it (profiling runs, reproducible timings).
  if (!forced_variant(s, std::getenv("PW_STEP_KERNEL")) && (e = autotune(s)) != hipSuccess)
    return fail(device_error(e, "step autotune"));
  launch_init_both(s, sp);
  if ((e = hipGetLastError()) != hipSuccess) return fail(device_error(e, "init launch"));
  if ((e = hipStreamSynchronize(s->stream)) != hipSuccess) return fail(device_error(e, "create"));
  {
    std::lock_guard<std::mutex> lock(g_live_mutex);
    g_live.insert(s);
  }
  *out = s;
  return PW_OK;
}

pw_status pw_solver_step(pw_solver* s, uint32_t count, int write_macro_on_last) {
  return step_run(s, count, write_macro_on_last);
}

pw_status pw_solver_advance_guarded(pw_solver* s, uint32_t count, uint64_t* unstable_step) {
  if (pw_status st = check_handle(s)) return st;
  if (!unstable_step) {
    set_error("Null output.");
    return PW_EINVAL;
  }
  *unstable_step = 0;
  // nan-guard.ts:151-196: batches end on multiples of the interval; each batch writes the
  // macros on its last step, and a batch that ends on a multiple is checked.
  uint32_t remaining = count;
  while (remaining > 0) {
    const uint32_t until = PW_GUARD_INTERVAL - (uint32_t)(s->steps % PW_GUARD_INTERVAL);
    const uint32_t steps = remaining < until ? remaining : until;
    const bool check = (s->steps + steps) % PW_GUARD_INTERVAL == 0;
    if (check) hipLaunchKernelGGL(guard_clear_kernel, dim3(1), dim3(1), 0, s->stream, s->guard_flag);
    if (pw_status st = step_run(s, steps, 1)) return st;
    remaining -= steps;
    if (!check) continue;
    PW_TRY(hipMemcpyAsync(s->guard_host, s->guard_flag, sizeof(uint32_t), hipMemcpyDeviceToHost, s->stream), "guard readback");
    PW_TRY(hipStreamSynchronize(s->stream), "guard readback");
    if (*s->guard_host != 0) {
      const uint64_t at = s->steps;
      if (pw_status st = pw_solver_reset(s)) return st;
      PW_TRY(hipStreamSynchronize(s->stream), "reset");
      *unstable_step = at;
      return PW_OK;
    }
  }
  return PW_OK;
}

}  // extern "C"

namespace pw {
pw_status step_run(pw_solver* s, uint32_t count, int write_macro_on_last) {
  if (pw_status st = check_handle(s)) return st;
  for (uint32_t k = 0; k < count; ++k) {
    const bool macro = write_macro_on_last && k + 1 == count;
    StepParams p = s->sp;
    if (s->params.ramp_steps > 0 && s->steps < s->params.ramp_steps) {
      // gpu-solver.ts:401-414: the ramped speed of step `steps + 1`.
      p.u_in = (float)(s->params.u_in * ramp_factor(s->steps + 1, s->params.ramp_steps));
    }
    launch_step(s, macro, p);
    s->current ^= 1;
    s->steps += 1;
  }
  // An overlapped closure evaluation ends with the call that started it, so every other entry
  // point finds the closure joined.
  settle_closure(s);
  if (count) PW_TRY(hipGetLastError(), "step launch");
  return PW_OK;
}
}  // namespace pw

extern "C" {

pw_status pw_solver_encode_force(pw_solver* s) {
  if (pw_status st = check_handle(s)) return st;
  if (s->force_encoded) return PW_OK;  // gpu-force.ts:176: an occupied slot skips the sample
  launch_force(s, s->force_dev);
  PW_TRY(hipGetLastError(), "force launch");
  PW_TRY(hipMemcpyAsync(s->force_host, s->force_dev, 3 * sizeof(double), hipMemcpyDeviceToHost, s->stream), "force readback");
  PW_TRY(hipEventRecord(s->force_event, s->stream), "force event");
  s->force_encoded = 1;
  s->force_step = s->steps;
  return PW_OK;
}

pw_status pw_solver_read_force(pw_solver* s, pw_force_sample* out) {
  if (pw_status st = check_handle(s)) return st;
  if (!out) {
    set_error("Null output.");
    return PW_EINVAL;
  }
  if (!s->force_encoded) return PW_EBUSY;
  s->force_encoded = 0;
  PW_TRY(hipEventSynchronize(s->force_event), "force wait");
  out->step = s->force_step;
  out->fx = s->force_host[0];
  out->fy = s->force_host[1];
  out->fz = s->force_host[2];
  return PW_OK;
}

pw_status pw_solver_reset(pw_solver* s) {
  if (pw_status st = check_handle(s)) return st;
  StepParams p = s->sp;
  if (s->params.ramp_steps > 0) p.u_in = 0.0f;
  launch_init_both(s, p);
  PW_TRY(hipGetLastError(), "reset launch");
  s->current = 0;
  s->steps = 0;
  s->closure_fresh = 0;
  s->ramp_settled = 0;
  s->force_encoded = 0;  // gpu-force.ts:208-210: an unread sample is discarded
  return PW_OK;
}

pw_status pw_solver_set_rounding(pw_solver* s, uint32_t rounding) {
  if (pw_status st = check_handle(s)) return st;
  if (rounding > PW_ROUNDING_REFERENCE) {
    set_error("Unknown f16 rounding.");
    return PW_EINVAL;
  }
  if (s->params.storage != PW_STORAGE_F16) {
    set_error("The rounding option applies to f16 storage.");
    return PW_EINVAL;
  }
  s->rounding = rounding;
  return pw_solver_reset(s);
}

uint64_t pw_solver_step_count(const pw_solver* s) { return solver_live(s) ? s->steps : 0; }

void pw_solver_destroy(pw_solver* s) {
  {
    std::lock_guard<std::mutex> lock(g_live_mutex);
    if (!s || !g_live.erase(s)) return;
  }
  free_all(s);
  delete s;
}

pw_status pw_solver_read_macro(pw_solver* s, float* out, size_t out_floats) {
  if (pw_status st = check_handle(s)) return st;
  if (!out || out_floats < 4 * s->cells) {
    set_error("Macro output needs 4 floats per cell.");
    return PW_EINVAL;
  }
  PW_TRY(hipMemcpyAsync(out, s->macro, 4 * s->cells * sizeof(float), hipMemcpyDeviceToHost, s->stream), "macro readback");
  PW_TRY(hipStreamSynchronize(s->stream), "macro readback");
  return PW_OK;
}

size_t pw_solver_population_bytes(const pw_solver* s) { return solver_live(s) ? (size_t)kQ * s->cells * s->elem_bytes : 0; }

pw_status pw_solver_read_populations(pw_solver* s, void* out, size_t out_bytes) {
  if (pw_status st = check_handle(s)) return st;
  const size_t bytes = pw_solver_population_bytes(s);
  if (!out || out_bytes < bytes) {
    set_error("Population output is smaller than one state.");
    return PW_EINVAL;
  }
  // The device layout keeps each row's directions together; the caller gets the reference's
  // planes back to back.
  PW_TRY(set_to_host_planes(s->pops[s->current], out, s->sp, s->elem_bytes, s->stream), "population readback");
  return PW_OK;
}

pw_status pw_solver_write_populations(pw_solver* s, const void* in, size_t in_bytes) {
  if (pw_status st = check_handle(s)) return st;
  const size_t bytes = pw_solver_population_bytes(s);
  if (!in || in_bytes != bytes) {
    set_error("Population input must hold exactly one state.");
    return PW_EINVAL;
  }
  PW_TRY(set_from_host_planes(s->pops[s->current], in, s->sp, s->elem_bytes, s->stream), "population upload");
  launch_edge_any(s, s->pops[s->current], s->sp);
  PW_TRY(hipGetLastError(), "upstream moments");
  PW_TRY(hipStreamSynchronize(s->stream), "population upload");
  return PW_OK;
}

pw_status pw_debug_inject_nan(pw_solver* s) {
  if (pw_status st = check_handle(s)) return st;
  if (s->first_fluid < 0) return PW_OK;
  const size_t off = (size_t)pop_cell(s->sp, (uint32_t)s->first_fluid) * s->elem_bytes;
  PW_TRY(hipStreamSynchronize(s->stream), "inject");
  if (s->elem_bytes == 4) {
    const uint32_t nan_bits = 0x7fc00000u;
    PW_TRY(hipMemcpy((char*)s->pops[s->current] + off, &nan_bits, 4, hipMemcpyHostToDevice), "inject");
  } else {
    const uint16_t nan_bits = s->params.storage == PW_STORAGE_F16 ? 0x7e00u : 0x7fc0u;
    PW_TRY(hipMemcpy((char*)s->pops[s->current] + off, &nan_bits, 2, hipMemcpyHostToDevice), "inject");
  }
  launch_edge_any(s, s->pops[s->current], s->sp);
  PW_TRY(hipGetLastError(), "upstream moments");
  return PW_OK;
}

pw_status pw_solver_device_state(pw_solver* s, pw_device_state* out) {
  if (pw_status st = check_handle(s)) return st;
  out->populations = s->pops[s->current];
  out->run_elements = s->sp.run;
  out->macro = (const float*)s->macro;
  out->geometry = s->geometry;
  out->stream = (void*)s->stream;
  return PW_OK;
}

pw_status pw_solver_set_kernel(pw_solver* s, uint32_t kernel) {
  if (pw_status st = check_handle(s)) return st;
  if (kernel == PW_KERNEL_AUTO) {
    PW_TRY(hipStreamSynchronize(s->stream), "tune");
    // Tuning steps advance the state: keep a copy of the latest set and restore it.
    const size_t bytes = set_bytes(s);
    void* keep = nullptr;
    PW_TRY(hipMalloc(&keep, bytes), "tune scratch");
    const int cur = s->current;
    hipError_t e = hipMemcpyAsync(keep, s->pops[cur], bytes, hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = autotune(s);
    if (e == hipSuccess) e = hipMemcpyAsync(s->pops[cur], keep, bytes, hipMemcpyDeviceToDevice, s->stream);
    s->current = cur;
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    hipFree(keep);
    if (e == hipSuccess && s->closure_on) e = closure_tile_order(s);
    if (e != hipSuccess) return device_error(e, "tune");
    return PW_OK;
  }
  if (kernel > PW_KERNEL_CELL_NT_STORES) {
    set_error("Unknown step kernel.");
    return PW_EINVAL;
  }
  s->kernel = kernel;
  s->block = kStepBlock;
  if (s->closure_on) PW_TRY(closure_tile_order(s), "closure block order");
  return PW_OK;
}

uint32_t pw_solver_kernel(const pw_solver* s) { return solver_live(s) ? s->kernel : 0; }
uint32_t pw_solver_block_size(const pw_solver* s) { return solver_live(s) ? s->block : 0; }
uint32_t pw_solver_xcd_chunks(const pw_solver* s) { return solver_live(s) ? s->xcd_chunks : 0; }

pw_status pw_debug_cell_path(pw_solver* s, int on) {
  if (pw_status st = check_handle(s)) return st;
  PW_TRY(hipStreamSynchronize(s->stream), "cell path");
  s->sp.rows = (!on && s->layout.nx % 64 == 0) ? 1u : 0u;
  return PW_OK;
}

pw_status pw_solver_set_step_variant(pw_solver* s, uint32_t kernel, uint32_t block, uint32_t xcd_chunks) {
  if (pw_status st = check_handle(s)) return st;
  if (kernel < PW_KERNEL_CELL || kernel > PW_KERNEL_CELL_NT_STORES || (block != 64 && block != 128 && block != 256) ||
      !valid_xcd_chunks(xcd_chunks)) {
    set_error("A step variant needs a kernel 1-3, a block of 64, 128 or 256 and a power of two up to 1024 XCD chunks.");
    return PW_EINVAL;
  }
  s->kernel = kernel;
  s->block = block;
  s->xcd_chunks = xcd_chunks;
  if (s->closure_on) PW_TRY(closure_tile_order(s), "closure block order");
  return PW_OK;
}

pw_status pw_solver_synchronize(pw_solver* s) {
  if (pw_status st = check_handle(s)) return st;
  PW_TRY(hipStreamSynchronize(s->stream), "synchronize");
  return PW_OK;
}

pw_status pw_solver_step_sampled(pw_solver* s, uint32_t count, uint32_t every, double* out, size_t out_doubles) {
  if (pw_status st = check_handle(s)) return st;
  if (every == 0 || !out) {
    set_error("Sampling needs a positive interval and an output buffer.");
    return PW_EINVAL;
  }
  size_t samples = 0;
  for (uint64_t k = s->steps + 1; k <= s->steps + count; ++k) samples += (k % every) == 0;
  if (out_doubles < 4 * samples) {
    set_error("Sample output needs 4 doubles per sample.");
    return PW_EINVAL;
  }
  double* hist = nullptr;
  if (samples) PW_TRY(hipMallocAsync((void**)&hist, samples * 3 * sizeof(double), s->stream), "sample history");
  size_t k = 0;
  std::vector<uint64_t> stamps;
  for (uint32_t n = 0; n < count; ++n) {
    if (pw_status st = pw_solver_step(s, 1, 0)) {
      if (hist) (void)hipFreeAsync(hist, s->stream);
      return st;
    }
    if (s->steps % every == 0) {
      launch_force(s, hist + 3 * k++);
      stamps.push_back(s->steps);
    }
  }
  std::vector<double> host(3 * samples);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && samples) e = hipMemcpyAsync(host.data(), hist, host.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream);
  if (hist) (void)hipFreeAsync(hist, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return device_error(e, "sampled steps");
  for (size_t j = 0; j < samples; ++j) {
    out[4 * j] = (double)stamps[j];
    out[4 * j + 1] = host[3 * j];
    out[4 * j + 2] = host[3 * j + 1];
    out[4 * j + 3] = host[3 * j + 2];
  }
  return PW_OK;
}

}  // extern "C"

namespace pw {
namespace {

// Scratch of pw_solver_step_sampled_noalloc: the FP64 force history (fx, fy, fz per sample)
// of one call, at most kScratchSamples samples (96 KiB). Every GPU holds its own copy for the
// life of the process; all calls on one GPU share it, so calls that overlap on one GPU race.
constexpr size_t kScratchSamples = 4096;
static __device__ __align__(256) double scratch[3 * kScratchSamples];

// Device address of `scratch` on each device, looked up once per device with that device
// current. The host-side name `scratch` is only a stand-in for the symbol; kernels receive
// this address.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

hipError_t scratch_on(int device, double** out) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    int current = 0;
    hipError_t e = hipGetDevice(&current);
    if (e != hipSuccess) return e;
    if (current != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (current != device) (void)hipSetDevice(current);
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

}  // namespace
}  // namespace pw

extern "C" {

pw_status pw_solver_step_sampled_noalloc(pw_solver* s, uint32_t count, uint32_t every, double* out,
                                         size_t out_doubles) {
  if (pw_status st = check_handle(s)) return st;
  if (every == 0 || !out) {
    set_error("Sampling needs a positive interval and an output buffer.");
    return PW_EINVAL;
  }
  // Samples fall on the multiples of `every` in (steps, steps + count].
  const uint64_t first = (s->steps / every + 1) * every;
  const uint64_t last = s->steps + count;
  const size_t samples = last >= first ? (size_t)((last - first) / every + 1) : 0;
  if (samples > kScratchSamples) {
    set_error("The allocation-free sampled stepper holds at most 4096 samples per call.");
    return PW_EINVAL;
  }
  if (out_doubles < 4 * samples) {
    set_error("Sample output needs 4 doubles per sample.");
    return PW_EINVAL;
  }
  double* hist = nullptr;
  PW_TRY(scratch_on(s->device, &hist), "scratch address");
  size_t k = 0;
  for (uint32_t n = 0; n < count; ++n) {
    if (pw_status st = step_run(s, 1, 0)) return st;
    if (s->steps % every == 0) launch_force(s, hist + 3 * k++);
  }
  PW_TRY(hipGetLastError(), "sampled steps");
  // The history lands in the tail of `out` and is spread to (step, fx, fy, fz) records from the
  // front: record j ends at 4j + 3 < samples + 3(j + 1), where unread entries start.
  if (samples) PW_TRY(hipMemcpyAsync(out + samples, hist, 3 * samples * sizeof(double), hipMemcpyDeviceToHost, s->stream), "sample readback");
  PW_TRY(hipStreamSynchronize(s->stream), "sampled steps");
  for (size_t j = 0; j < samples; ++j) {
    const double fx = out[samples + 3 * j], fy = out[samples + 3 * j + 1], fz = out[samples + 3 * j + 2];
    out[4 * j] = (double)(first + j * every);
    out[4 * j + 1] = fx;
    out[4 * j + 2] = fy;
    out[4 * j + 3] = fz;
  }
  return PW_OK;
}

pw_status pw_solver_time_steps(pw_solver* s, uint32_t count, float* elapsed_ms) {
  if (pw_status st = check_handle(s)) return st;
  PW_TRY(hipEventRecord(s->t0, s->stream), "timing event");
  if (pw_status st = pw_solver_step(s, count, 0)) return st;
  PW_TRY(hipEventRecord(s->t1, s->stream), "timing event");
  PW_TRY(hipEventSynchronize(s->t1), "timing wait");
  PW_TRY(hipEventElapsedTime(elapsed_ms, s->t0, s->t1), "timing");
  return PW_OK;
}

}  // extern "C"
