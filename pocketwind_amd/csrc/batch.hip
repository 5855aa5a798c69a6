// batch.hip -- many independent tunnels of uneven size stepped by one launch per step.
//
// The reference runs one solver at a time (src/app/solver-slot.ts:14-44); batching has no
// reference counterpart. Each tunnel keeps the single-solver semantics exactly (the same
// step code, row wavefronts or per cell as the single solver chooses, its own tau, inlet
// speed, geometry, ping-pong sets, macros and learned-closure output), so a tunnel stepped in
// a batch is bit-identical to the same tunnel stepped alone.
//
// Work decomposition: the cells of all tunnels are cut into tiles of one block's worth of
// cells; a tile table maps a tile to its tunnel and the grid holds one block per tile, so one
// launch advances every tunnel by one step. All tunnels step in lockstep, so
// one parity bit selects every tunnel's input set; the two launches of an even/odd step pair
// are captured once into a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "batch_step.h"

namespace pw {
namespace batch {


template <class S>
__global__ __launch_bounds__(kBlock) void batch_init_kernel(const Tunnel* __restrict__ tunnels,
                                                            const uint32_t* __restrict__ tile_tunnel,
                                                            const uint32_t* __restrict__ tile_first,
                                                            uint32_t n_tiles, int set) {
  for (uint32_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
    const uint32_t t = tile_tunnel[tile];
    const Tunnel& tu = tunnels[t];
    const uint32_t n = (tile - tile_first[t]) * blockDim.x + threadIdx.x;
    if (n >= tu.p.cells) continue;
    const uint32_t kind = tu.geo[n] & kTypeMask;
    const float ux = is_wall(kind) ? 0.0f : tu.p.u_in;
    store_equilibrium<S>((typename S::T*)tu.pops[set], tu.p, n, ux, 0.0f, 0.0f, S::cell_word(n, kInitSeed));
    tu.macro[n] = make_float4(1.0f, ux, 0.0f, 0.0f);
  }
}

// Force partials of every tunnel: block b handles 256 entries of one tunnel's boundary list.
template <class S, bool BOUZIDI>
__global__ __launch_bounds__(kForceBlock) void batch_force_kernel(const Tunnel* __restrict__ tunnels,
                                                                  const uint32_t* __restrict__ part_tunnel,
                                                                  double* __restrict__ partials) {
  const uint32_t t = part_tunnel[blockIdx.x];
  const Tunnel& tu = tunnels[t];
  const uint32_t item = (blockIdx.x - tu.partial_base) * kForceBlock + threadIdx.x;
  double fx = 0, fy = 0, fz = 0;
  if (item < tu.n_boundary) {
    const uint32_t n = tu.boundary[item];
    const uint32_t flag = tu.geo[n];
    const StepParams& p = tu.p;
    uint32_t x, y, z;
    cell_coords(n, p, x, y, z);
    const Neighbours nb(x, y, z, p);
    const uint32_t self = pop_x(p, x) + pop_row(p, y + p.ny * z);
    double ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
    cell_exchange<S, BOUZIDI>((const typename S::T*)tu.pops[0], tu.geo, p, nb, n, self, flag, ax, ay, az);
    cell_exchange<S, BOUZIDI>((const typename S::T*)tu.pops[1], tu.geo, p, nb, n, self, flag, bx, by, bz);
    fx = 0.5 * (ax + bx);
    fy = 0.5 * (ay + by);
    fz = 0.5 * (az + bz);
  }
  __shared__ double lds[3][kForceBlock / 64];
  fx = wave_sum(fx);
  fy = wave_sum(fy);
  fz = wave_sum(fz);
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0) {
    lds[0][wave] = fx;
    lds[1][wave] = fy;
    lds[2][wave] = fz;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double s = 0;
    for (int w = 0; w < kForceBlock / 64; ++w) s += lds[threadIdx.x][w];
    partials[3 * (size_t)blockIdx.x + threadIdx.x] = s;
  }
}

// The learned closure of every tunnel in one launch: one wavefront per 16-cell tile of a
// tunnel's wall-cell list (ctile_tunnel / ctile_first map a tile to its tunnel and the tunnel's
// first tile, as the step's tile tables do), on the set of parity `parity`, into the tunnel's
// closure buffer. Tunnels that evaluate the closure inside the step skip their tiles. A tile's
// stresses are the single solver's closure kernel's bit for bit (closure::closure_tile).
template <class S, bool BOUZIDI>
__global__ __launch_bounds__(64) void batch_closure_kernel(const Tunnel* __restrict__ tunnels,
                                                           const uint32_t* __restrict__ ctile_tunnel,
                                                           const uint32_t* __restrict__ ctile_first, int parity,
                                                           const mlp::Packed* __restrict__ P, closure::Norm norm) {
  __shared__ __attribute__((aligned(16))) char lds[closure::kFusedWaveBytes];
  const uint32_t t = ctile_tunnel[blockIdx.x];
  const Tunnel& tu = tunnels[t];
  if (tu.p.closure_net) return;
  closure::closure_tile<S, BOUZIDI>((const typename S::T*)tu.pops[parity], tu.geo, tu.p,
                                    static_cast<const closure::WallGeo*>(tu.wall_geo), tu.n_wall, P, norm,
                                    tu.closure_buf, nullptr, blockIdx.x - ctile_first[t], lds);
}

// One block per tunnel reduces that tunnel's partials in the single-solver order.
__global__ __launch_bounds__(kForceBlock) void batch_reduce_kernel(const Tunnel* __restrict__ tunnels,
                                                                   const double* __restrict__ partials,
                                                                   double* __restrict__ result) {
  const Tunnel& tu = tunnels[blockIdx.x];
  force_reduce_block(partials + 3 * (size_t)tu.partial_base, tu.n_partials, result + 3 * (size_t)blockIdx.x);
}

}  // namespace batch
}  // namespace pw

namespace pw {
namespace batch {
extern template void batch_step_s<StoreF32>(const BatchLaunch&, int, bool);
extern template void batch_step_s<StoreF16>(const BatchLaunch&, int, bool);
extern template void batch_step_s<StoreBF16>(const BatchLaunch&, int, bool);
}  // namespace batch
}  // namespace pw

using namespace pw;
using namespace pw::batch;

struct pw_batch {
  hipStream_t stream = nullptr;
  pw_params params{};
  uint32_t count = 0;
  std::vector<pw_layout> layouts;
  std::vector<Tunnel> host;
  Tunnel* tunnels = nullptr;
  uint32_t* tile_tunnel = nullptr;
  uint32_t* tile_first = nullptr;
  uint32_t n_tiles = 0;
  uint32_t policy = kNtStores;  // step variant (tuned at create)
  uint32_t xcd_chunks = 16;
  uint32_t chunk = 1;           // tiles per XCD chunk of the step grid
  unsigned step_grid = 0;
  uint32_t persistent = 0;      // resident blocks of the persistent step form, 0: one block per tile
  uint32_t block = kBlock;      // threads per step block = cells per tile (the tile table's)
  std::vector<uint32_t> tunnel_cells;  // cells of each tunnel (the tile tables are built from them)
  uint32_t cus = 0;             // compute units of the device
  uint32_t* part_tunnel = nullptr;
  uint32_t n_partials = 0;
  double* partials = nullptr;
  std::vector<void*> allocations;
  uint32_t* seeds = nullptr;  // two step-count words, see batch_step_kernel
  int parity = 0;
  uint64_t steps = 0;
  unsigned grid = 0;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  hipEvent_t t0 = nullptr, t1 = nullptr;
  // Learned wall closure (created with PW_WALL_LEARNED): the network, its normalisation and
  // the evaluation interval, shared by every tunnel.
  void* closure_net = nullptr;
  float closure_norm[32] = {};
  int closure_on = 0;
  uint32_t closure_every = 1;
  uint32_t closure_placement = 0;  // pw_closure_placement
  int closure_fresh = 0;
  std::vector<std::pair<const uint32_t*, float4*>> closure_bufs;  // per tunnel: slot table, output
  // The closure kernel's tiles (16 wall cells each) over all tunnels: tunnel of each tile, first
  // tile of each tunnel.
  uint32_t* ctile_tunnel = nullptr;
  uint32_t* ctile_first = nullptr;
  uint32_t n_ctiles = 0;
  bool closure_kernels = false;  // some tunnel evaluates the closure in the closure kernel
  uint32_t lds = 0;              // dynamic LDS per step block (closure evaluated in the step)
};

namespace {

void launch_step(pw_batch* b, int parity, bool macro) {
  const BatchLaunch L{b->step_grid, b->stream, b->tunnels, b->tile_tunnel, b->tile_first, b->n_tiles, b->seeds,
                      b->params, b->policy, b->chunk, b->lds, b->persistent, b->block};
  switch (b->params.storage) {
    case PW_STORAGE_F16: batch_step_s<StoreF16>(L, parity, macro); break;
    case PW_STORAGE_BF16: batch_step_s<StoreBF16>(L, parity, macro); break;
    default: batch_step_s<StoreF32>(L, parity, macro); break;
  }
}

template <class S>
void launch_edges(pw_batch* b, int set) {
  for (const Tunnel& tu : b->host)
    if (tu.p.nx % 64 == 0)
      hipLaunchKernelGGL(edge_moments_kernel<S>, dim3((tu.p.ny * tu.p.nz + kBlock - 1) / kBlock), dim3(kBlock), 0,
                         b->stream, (typename S::T*)tu.pops[set], tu.p);
}

void launch_init(pw_batch* b) {
  const unsigned grid = b->grid;
  for (int set = 0; set < 2; ++set) {
    switch (b->params.storage) {
      case PW_STORAGE_F16: hipLaunchKernelGGL(batch_init_kernel<StoreF16>, dim3(grid), dim3(b->block), 0, b->stream, b->tunnels, b->tile_tunnel, b->tile_first, b->n_tiles, set); launch_edges<StoreF16>(b, set); break;
      case PW_STORAGE_BF16: hipLaunchKernelGGL(batch_init_kernel<StoreBF16>, dim3(grid), dim3(b->block), 0, b->stream, b->tunnels, b->tile_tunnel, b->tile_first, b->n_tiles, set); launch_edges<StoreBF16>(b, set); break;
      default: hipLaunchKernelGGL(batch_init_kernel<StoreF32>, dim3(grid), dim3(b->block), 0, b->stream, b->tunnels, b->tile_tunnel, b->tile_first, b->n_tiles, set); launch_edges<StoreF32>(b, set); break;
    }
  }
}

// The closure of every tunnel on the state the next step reads.
template <class S>
void launch_closures_s(pw_batch* b) {
  closure::Norm norm;
  std::memcpy(&norm, b->closure_norm, sizeof norm);
  auto k = b->params.walls == PW_WALLS_BOUZIDI ? batch_closure_kernel<S, true> : batch_closure_kernel<S, false>;
  hipLaunchKernelGGL(k, dim3(b->n_ctiles), dim3(64), 0, b->stream, b->tunnels, b->ctile_tunnel, b->ctile_first,
                     b->parity, static_cast<const mlp::Packed*>(b->closure_net), norm);
}

void launch_closures(pw_batch* b) {
  if (!b->n_ctiles) return;
  switch (b->params.storage) {
    case PW_STORAGE_F16: launch_closures_s<StoreF16>(b); break;
    case PW_STORAGE_BF16: launch_closures_s<StoreBF16>(b); break;
    default: launch_closures_s<StoreF32>(b); break;
  }
}

void launch_force(pw_batch* b, double* dest) {
  if (b->n_partials) {
    const bool bz = b->params.walls == PW_WALLS_BOUZIDI;
    switch (b->params.storage) {
      case PW_STORAGE_F16: hipLaunchKernelGGL((bz ? batch_force_kernel<StoreF16, true> : batch_force_kernel<StoreF16, false>), dim3(b->n_partials), dim3(kForceBlock), 0, b->stream, b->tunnels, b->part_tunnel, b->partials); break;
      case PW_STORAGE_BF16: hipLaunchKernelGGL((bz ? batch_force_kernel<StoreBF16, true> : batch_force_kernel<StoreBF16, false>), dim3(b->n_partials), dim3(kForceBlock), 0, b->stream, b->tunnels, b->part_tunnel, b->partials); break;
      default: hipLaunchKernelGGL((bz ? batch_force_kernel<StoreF32, true> : batch_force_kernel<StoreF32, false>), dim3(b->n_partials), dim3(kForceBlock), 0, b->stream, b->tunnels, b->part_tunnel, b->partials); break;
    }
  }
  hipLaunchKernelGGL(batch_reduce_kernel, dim3(b->count), dim3(kForceBlock), 0, b->stream, b->tunnels, b->partials, dest);
}

void free_batch(pw_batch* b) {
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto& g : b->graph)
    if (g) (void)hipGraphExecDestroy(g);
  for (void* p : b->allocations) (void)hipFree(p);
  if (b->t0) (void)hipEventDestroy(b->t0);
  if (b->t1) (void)hipEventDestroy(b->t1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
}

template <class T>
hipError_t dev_alloc(pw_batch* b, T** out, size_t n) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, (n ? n : 1) * sizeof(T));
  if (e == hipSuccess) b->allocations.push_back(p);
  *out = (T*)p;
  return e;
}

// Capture one even/odd step pair (no macro writes) into an executable graph.
hipError_t capture_pair(pw_batch* b, int first_parity, hipGraphExec_t* exec) {
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  launch_step(b, first_parity, false);
  launch_step(b, first_parity ^ 1, false);
  e = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess) return e;
  e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  return e;
}

pw_status step_plain(pw_batch* b, uint32_t count, int write_macro_on_last);

// Where each tunnel's closure runs once a network is attached: inside the step (p.closure_net)
// on row tunnels when the placement asks for it, every step; otherwise in the closure kernel
// into the tunnel's buffer (p.closure_tau). The captured step graphs hold the launch's LDS
// size and are dropped.
hipError_t apply_closure_mode(pw_batch* b) {
  b->closure_kernels = false;
  b->lds = 0;
  for (uint32_t t = 0; t < b->count; ++t) {
    StepParams& p = b->host[t].p;
    const bool fused = b->closure_placement == PW_CLOSURE_IN_STEP && p.rows && b->params.op == PW_OP_REG && b->params.les;
    p.closure_net = fused ? b->closure_net : nullptr;
    p.closure_slot = b->closure_bufs[t].first;  // the kernel's buffer, or the records in the step
    p.closure_tau = fused ? nullptr : b->closure_bufs[t].second;
    p.closure_geo = b->host[t].wall_geo;
    if (fused) b->lds = (b->block / 64) * closure::kFusedWaveBytes;
    else b->closure_kernels = true;
  }
  for (auto& g : b->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  hipError_t e = hipMemcpyAsync(b->tunnels, b->host.data(), b->count * sizeof(Tunnel), hipMemcpyHostToDevice, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  return e;
}

// Tile tables for `block` cells per tile (one block of that many threads per tile): the tiles
// of every tunnel in order, each tunnel's first tile. Uploaded on the batch's stream.
hipError_t set_block(pw_batch* b, uint32_t block) {
  if (block == b->block && b->n_tiles) return hipSuccess;
  std::vector<uint32_t> tt, tf(b->count);
  for (uint32_t t = 0; t < b->count; ++t) {
    tf[t] = (uint32_t)tt.size();
    const size_t n = (b->tunnel_cells[t] + block - 1) / block;
    tt.insert(tt.end(), n, t);
  }
  hipError_t e = hipMemcpyAsync(b->tile_tunnel, tt.data(), tt.size() * 4, hipMemcpyHostToDevice, b->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(b->tile_first, tf.data(), b->count * 4, hipMemcpyHostToDevice, b->stream);
  // the host vectors must outlive the copies
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return e;
  b->block = block;
  b->n_tiles = (uint32_t)tt.size();
  b->grid = (unsigned)std::min<uint64_t>(b->n_tiles, 0x7fffffffu);
  return hipSuccess;
}

// Resident blocks of the persistent step form on this device: the collision's blocks per CU
// times the CUs, a multiple of 8.
uint32_t persistent_grid(const pw_batch* b) {
  const bool reg = b->params.op == PW_OP_REG, les = b->params.les != 0;
  const uint32_t B = b->block;
  const uint32_t per_cu = reg ? (les ? persistent_blocks_per_cu<kOpReg, true>(B) : persistent_blocks_per_cu<kOpReg, false>(B))
                              : (les ? persistent_blocks_per_cu<kOpTrt, true>(B) : persistent_blocks_per_cu<kOpTrt, false>(B));
  return per_cu * b->cus / kXcds * kXcds;
}

bool forced_batch_variant(pw_batch* b, const char* text) {
  if (!text) return false;
  const std::string v(text);
  const size_t a = v.find('/');
  const std::string name = v.substr(0, a);
  uint32_t policy;
  if (name == "cell") policy = kCached;
  else if (name == "cell-nt") policy = kNtAll;
  else if (name == "cell-nt-stores") policy = kNtStores;
  else return false;
  uint32_t chunks = b->xcd_chunks;
  if (a != std::string::npos) chunks = (uint32_t)std::strtoul(v.c_str() + a + 1, nullptr, 10);
  if (!valid_xcd_chunks(chunks)) return false;
  // further tokens: a block size (64, 128, 256) and/or "persistent"
  uint32_t block = kBlock;
  bool persist = false;
  for (size_t c = a == std::string::npos ? a : v.find('/', a + 1); c != std::string::npos;) {
    const size_t d = v.find('/', c + 1);
    const std::string tok = v.substr(c + 1, d == std::string::npos ? std::string::npos : d - c - 1);
    if (tok == "persistent") persist = true;
    else if (tok == "64" || tok == "128" || tok == "256") block = (uint32_t)std::strtoul(tok.c_str(), nullptr, 10);
    else return false;
    c = d;
  }
  if (set_block(b, block) != hipSuccess) return false;
  b->policy = policy;
  b->xcd_chunks = chunks;
  b->persistent = persist ? persistent_grid(b) : 0u;
  return true;
}

// Times the three cache policies with each chunking of kXcdChunkings that leaves every chunk
// two tiles or more, each launched one block per tile and in the persistent form (two runs of a
// few steps each, the faster counting), and keeps the fastest.
hipError_t tune_batch(pw_batch* b) {
  const uint32_t policies[3] = {kCached, kNtAll, kNtStores};
  const uint64_t cells = [&] {
    uint64_t n = 0;
    for (const auto& t : b->host) n += t.p.cells;
    return n;
  }();
  const int reps = cells >= (1u << 22) ? 4 : 16;
  struct Candidate {
    float ms;
    uint32_t block, chunks, policy, persistent;
  };
  // The fastest of `rounds` timed runs of `n` steps each, after two warm-up steps (the tile
  // tables for the candidate's block are in place).
  auto time_variant = [&](Candidate& c, int rounds, int n) -> hipError_t {
    b->policy = c.policy;
    b->xcd_chunks = c.chunks;
    b->persistent = c.persistent;
    b->step_grid = batch_grid(b->n_tiles, c.chunks, &b->chunk);
    for (int w = 0; w < 2; ++w) {
      launch_step(b, b->parity, false);
      b->parity ^= 1;
    }
    for (int round = 0; round < rounds; ++round) {
      hipError_t e = hipEventRecord(b->t0, b->stream);
      if (e != hipSuccess) return e;
      for (int r = 0; r < n; ++r) {
        launch_step(b, b->parity, false);
        b->parity ^= 1;
      }
      if ((e = hipEventRecord(b->t1, b->stream)) != hipSuccess) return e;
      if ((e = hipEventSynchronize(b->t1)) != hipSuccess) return e;
      float ms = 0;
      if ((e = hipEventElapsedTime(&ms, b->t0, b->t1)) != hipSuccess) return e;
      if (round == 0 || ms < c.ms) c.ms = ms;
    }
    return hipSuccess;
  };
  // A screen of every candidate (two runs of `reps` steps), then the four fastest timed again
  // over four times as many steps, three times, as the single solver's tuner does.
  std::vector<Candidate> screened;
  for (uint32_t block : kBatchBlocks) {
    if (hipError_t e = set_block(b, block)) return e;
    const uint32_t resident = persistent_grid(b);
    for (uint32_t chunks : kXcdChunkings)
      for (uint32_t pol : policies)
        for (uint32_t persist : {0u, resident}) {
          if (chunks > 1 && (uint64_t)chunks * 2 * kXcds > b->n_tiles) continue;
          if (persist && (persist >= batch_grid(b->n_tiles, chunks, &b->chunk) || resident == 0)) continue;
          Candidate c{0.0f, block, chunks, pol, persist};
          if (hipError_t e = time_variant(c, 2, reps)) return e;
          screened.push_back(c);
        }
  }
  std::sort(screened.begin(), screened.end(), [](const Candidate& x, const Candidate& y) { return x.ms < y.ms; });
  Candidate best = screened.front();
  best.ms = 0.0f;
  for (size_t i = 0; i < screened.size() && i < 4; ++i) {
    Candidate c = screened[i];
    c.ms = 0.0f;
    if (hipError_t e = set_block(b, c.block)) return e;
    if (hipError_t e = time_variant(c, 3, 4 * reps)) return e;
    if (best.ms == 0.0f || c.ms < best.ms) best = c;
  }
  if (hipError_t e = set_block(b, best.block)) return e;
  b->policy = best.policy;
  b->xcd_chunks = best.chunks;
  b->persistent = best.persistent;
  return hipGetLastError();
}

}  // namespace

extern "C" {

pw_status pw_batch_create(uint32_t count, const pw_layout* layouts, const uint32_t* const* geometries,
                          const size_t* geometry_words, const uint32_t* const* boundary_cells,
                          const size_t* n_boundary, const double* tau0, const double* u_in,
                          const pw_params* shared, pw_batch** out) {
  if (!out || !layouts || !geometries || !geometry_words || !boundary_cells || !n_boundary || !tau0 || !u_in ||
      !shared || count == 0) {
    set_error("Null argument or empty batch.");
    return PW_EINVAL;
  }
  *out = nullptr;
  if (shared->ramp_steps != 0 || shared->wall_model > PW_WALL_LEARNED || shared->storage > PW_STORAGE_BF16 ||
      shared->op > PW_OP_REG || shared->walls > PW_WALLS_BOUZIDI || shared->sides != PW_SIDES_PERIODIC ||
      shared->outlet != PW_OUTLET_EQUILIBRIUM) {
    set_error("Batched tunnels take operator, walls, LES, storage and wall-model options without a ramp, "
              "periodic sides and the equilibrium outlet.");
    return PW_EINVAL;
  }
  // Validate every tunnel exactly as a single solver would, on the host.
  std::vector<StepParams> sps(count);
  size_t tiles = 0, partials = 0;
  for (uint32_t t = 0; t < count; ++t) {
    pw_params prm = *shared;
    prm.tau0 = tau0[t];
    prm.u_in = u_in[t];
    if (pw_status st = validate_tunnel(layouts[t], geometries[t], geometry_words[t], boundary_cells[t],
                                       n_boundary[t], prm, &sps[t])) {
      set_error("Tunnel " + std::to_string(t) + ": " + pw_last_error_message());
      return st;
    }
  }
  pw_batch* b = new (std::nothrow) pw_batch();
  if (!b) {
    set_error("Host allocation failed.");
    return PW_ENOMEM;
  }
  b->params = *shared;
  b->count = count;
  for (uint32_t t = 0; t < count; ++t) {
    tiles += (sps[t].cells + 63) / 64;  // the smallest tile the tuner tries
    partials += (n_boundary[t] + kForceBlock - 1) / kForceBlock;
  }
  if (tiles > 0xffffffffull || partials > 0x7fffffffull) {
    delete b;
    set_error("Batch exceeds 32-bit tile indexing.");
    return PW_EINVAL;
  }
  auto fail = [&](hipError_t e, const char* what) {
    free_batch(b);
    delete b;
    return device_error(e, what);
  };
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) return fail(e, "batch stream");
  {
    int device = 0, cus = 0;
    if ((e = hipGetDevice(&device)) != hipSuccess ||
        (e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess)
      return fail(e, "device attributes");
    b->cus = (uint32_t)cus;
  }
  if ((e = hipEventCreate(&b->t0)) != hipSuccess || (e = hipEventCreate(&b->t1)) != hipSuccess) return fail(e, "batch events");
  const uint32_t elem = shared->storage == PW_STORAGE_F32 ? 4 : 2;
  std::vector<uint32_t> tt, tf(count), pt;
  b->host.resize(count);
  b->layouts.assign(layouts, layouts + count);
  uint32_t part_base = 0;
  for (uint32_t t = 0; t < count; ++t) {
    Tunnel& tu = b->host[t];
    tu.p = sps[t];
    const size_t N = sps[t].cells;
    // 19 N elements, then (row-wavefront tunnels) one float4 of upstream-outlet moments per row.
    const size_t set_bytes = edge_offset_bytes(N, elem) + (size_t)sps[t].ny * sps[t].nz * sizeof(float4);
    for (int k = 0; k < 2; ++k) {
      void* p = nullptr;
      if ((e = hipMalloc(&p, set_bytes)) != hipSuccess) return fail(e, "batch populations");
      b->allocations.push_back(p);
      tu.pops[k] = p;
    }
    if (shared->wall_model == PW_WALL_LEARNED) {
      std::vector<uint32_t> walls, slot(N, 0xffffffffu);
      for (size_t c = 0; c < N; ++c)
        if ((geometries[t][c] & kTypeMask) == kFluid && (geometries[t][c] >> 3) != 0) {
          slot[c] = (uint32_t)walls.size();
          walls.push_back((uint32_t)c);
        }
      uint32_t *wl = nullptr, *sl = nullptr;
      float4 *tau = nullptr, *records = nullptr;
      if ((e = dev_alloc(b, &wl, walls.size())) != hipSuccess || (e = dev_alloc(b, &sl, N)) != hipSuccess ||
          (e = dev_alloc(b, &tau, walls.size())) != hipSuccess ||
          (e = dev_alloc(b, &records, walls.size() * (kWallGeoBytes / sizeof(float4)))) != hipSuccess)
        return fail(e, "batch closure buffers");
      if ((walls.size() && (e = hipMemcpy(wl, walls.data(), walls.size() * 4, hipMemcpyHostToDevice)) != hipSuccess) ||
          (e = hipMemcpy(sl, slot.data(), N * 4, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "batch closure upload");
      tu.wall_cells = wl;
      tu.n_wall = (uint32_t)walls.size();
      tu.wall_geo = records;  // written once the geometry is on the device
      tu.closure_buf = tau;
      b->closure_bufs.push_back({sl, tau});  // given to the step once a network is attached
    } else {
      tu.wall_cells = nullptr;
      tu.n_wall = 0;
      tu.wall_geo = nullptr;
      tu.closure_buf = nullptr;
    }
    uint32_t *geo = nullptr, *bnd = nullptr;
    float4* macro = nullptr;
    if ((e = dev_alloc(b, &geo, geometry_words[t])) != hipSuccess) return fail(e, "batch geometry");
    if ((e = dev_alloc(b, &bnd, n_boundary[t])) != hipSuccess) return fail(e, "batch boundary");
    if ((e = dev_alloc(b, &macro, N + 1)) != hipSuccess) return fail(e, "batch macros");  // + the guard word
    if ((e = hipMemcpyAsync(geo, geometries[t], geometry_words[t] * 4, hipMemcpyHostToDevice, b->stream)) != hipSuccess) return fail(e, "batch upload");
    if (n_boundary[t] && (e = hipMemcpyAsync(bnd, boundary_cells[t], n_boundary[t] * 4, hipMemcpyHostToDevice, b->stream)) != hipSuccess) return fail(e, "batch upload");
    tu.geo = geo;
    tu.macro = macro;
    if (tu.p.rows) {
      const std::vector<uint32_t> desc = wave_descriptors(layouts[t], geometries[t]);
      uint32_t* wd = nullptr;
      if ((e = dev_alloc(b, &wd, desc.size())) != hipSuccess) return fail(e, "batch descriptors");
      if ((e = hipMemcpy(wd, desc.data(), desc.size() * 4, hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "batch upload");
      tu.p.wave_desc = wd;
    }
    if (tu.wall_geo) launch_wall_geo(b->stream, geo, tu.p, tu.wall_cells, tu.n_wall, const_cast<void*>(tu.wall_geo));
    tu.boundary = bnd;
    tu.n_boundary = (uint32_t)n_boundary[t];
    tu.partial_base = part_base;
    tu.n_partials = (uint32_t)((n_boundary[t] + kForceBlock - 1) / kForceBlock);
    for (uint32_t k = 0; k < tu.n_partials; ++k) pt.push_back(t);
    part_base += tu.n_partials;
    tf[t] = (uint32_t)tt.size();
    for (size_t k = 0; k < (N + kBlock - 1) / kBlock; ++k) tt.push_back(t);
    b->tunnel_cells.push_back((uint32_t)N);
  }
  b->n_tiles = (uint32_t)tt.size();
  b->n_partials = (uint32_t)pt.size();
  if ((e = dev_alloc(b, &b->tunnels, count)) != hipSuccess) return fail(e, "batch table");
  if ((e = dev_alloc(b, &b->tile_tunnel, (size_t)tiles)) != hipSuccess) return fail(e, "batch table");  // room for 64-cell tiles
  if ((e = dev_alloc(b, &b->tile_first, count)) != hipSuccess) return fail(e, "batch table");
  if ((e = dev_alloc(b, &b->part_tunnel, pt.size())) != hipSuccess) return fail(e, "batch table");
  if ((e = dev_alloc(b, &b->partials, 3 * (size_t)(pt.size() ? pt.size() : 1))) != hipSuccess) return fail(e, "batch partials");
  if ((e = dev_alloc(b, &b->seeds, 2)) != hipSuccess) return fail(e, "batch seeds");
  const uint32_t seeds0[2] = {0, 0};
  if ((e = hipMemcpyAsync(b->tunnels, b->host.data(), count * sizeof(Tunnel), hipMemcpyHostToDevice, b->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(b->tile_tunnel, tt.data(), tt.size() * 4, hipMemcpyHostToDevice, b->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(b->tile_first, tf.data(), count * 4, hipMemcpyHostToDevice, b->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(b->seeds, seeds0, sizeof seeds0, hipMemcpyHostToDevice, b->stream)) != hipSuccess ||
      (pt.size() && (e = hipMemcpyAsync(b->part_tunnel, pt.data(), pt.size() * 4, hipMemcpyHostToDevice, b->stream)) != hipSuccess))
    return fail(e, "batch tables");
  // One block per tile: a grid the device schedules as blocks retire (a persistent grid larger
  // than the resident capacity would leave its last blocks to run after the others).
  b->grid = (unsigned)std::min<uint64_t>(b->n_tiles, 0x7fffffffu);
  if (shared->wall_model == PW_WALL_LEARNED) {
    if ((e = dev_alloc(b, reinterpret_cast<char**>(&b->closure_net), closure::kNetBlockBytes)) != hipSuccess)
      return fail(e, "batch closure network");
    std::vector<uint32_t> ct, cf(count);
    for (uint32_t t = 0; t < count; ++t) {
      cf[t] = (uint32_t)ct.size();
      for (uint32_t k = 0; k < b->host[t].n_wall; k += closure::kCellsPerWave) ct.push_back(t);
    }
    b->n_ctiles = (uint32_t)ct.size();
    if ((e = dev_alloc(b, &b->ctile_tunnel, ct.size() ? ct.size() : 1)) != hipSuccess ||
        (e = dev_alloc(b, &b->ctile_first, count)) != hipSuccess)
      return fail(e, "batch closure tables");
    if ((ct.size() && (e = hipMemcpyAsync(b->ctile_tunnel, ct.data(), ct.size() * 4, hipMemcpyHostToDevice, b->stream)) != hipSuccess) ||
        (e = hipMemcpyAsync(b->ctile_first, cf.data(), count * 4, hipMemcpyHostToDevice, b->stream)) != hipSuccess)
      return fail(e, "batch closure tables");
  }
  launch_init(b);
  // PW_BATCH_KERNEL=<cell|cell-nt|cell-nt-stores>/<chunks per XCD> fixes the step variant;
  // otherwise each is timed on these tunnels and the fastest kept. Timing leaves the state
  // changed, so the tunnels are initialised again afterwards.
  if (!forced_batch_variant(b, std::getenv("PW_BATCH_KERNEL")) && (e = tune_batch(b)) != hipSuccess)
    return fail(e, "batch tuning");
  b->step_grid = batch_grid(b->n_tiles, b->xcd_chunks, &b->chunk);
  launch_init(b);
  if ((e = hipMemcpyAsync(b->seeds, seeds0, sizeof seeds0, hipMemcpyHostToDevice, b->stream)) != hipSuccess)
    return fail(e, "batch seeds");
  b->parity = 0;
  if ((e = hipGetLastError()) != hipSuccess) return fail(e, "batch init");
  if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return fail(e, "batch init");
  *out = b;
  return PW_OK;
}

uint32_t pw_batch_block_size(const pw_batch* b) { return b ? b->block : 0u; }

uint32_t pw_batch_persistent_grid(const pw_batch* b) {
  return (b && b->persistent && b->persistent < b->step_grid) ? b->persistent : 0u;  // as batch_step_t launches
}

uint32_t pw_batch_step_variant(const pw_batch* b, uint32_t* xcd_chunks) {
  if (!b) return 0;
  if (xcd_chunks) *xcd_chunks = b->xcd_chunks;
  return b->policy == kNtAll ? PW_KERNEL_CELL_NT : b->policy == kNtStores ? PW_KERNEL_CELL_NT_STORES : PW_KERNEL_CELL;
}

pw_status pw_batch_step(pw_batch* b, uint32_t count, int write_macro_on_last) {
  if (!b) {
    set_error("Null batch handle.");
    return PW_EINVAL;
  }
  if (!b->closure_on || !b->closure_kernels) return step_plain(b, count, write_macro_on_last);
  // The closure runs on the state a step reads, before the first step of a run and at every
  // multiple of its interval, as in a single solver.
  uint32_t done = 0;
  while (done < count) {
    if (!b->closure_fresh || b->steps % b->closure_every == 0) {
      launch_closures(b);
      b->closure_fresh = 1;
    }
    const uint32_t chunk =
        (uint32_t)std::min<uint64_t>(b->closure_every - b->steps % b->closure_every, count - done);
    if (pw_status st = step_plain(b, chunk, write_macro_on_last && done + chunk == count)) return st;
    done += chunk;
  }
  return PW_OK;
}

pw_status pw_batch_set_closure(pw_batch* b, const pw_mlp* net, const float* norm, size_t n_norm) {
  if (!b || b->params.wall_model != PW_WALL_LEARNED || !b->closure_net) {
    set_error("The batch was not created with the learned wall model.");
    return PW_EINVAL;
  }
  if (!net || !norm || n_norm != 27) {
    set_error("A closure needs a network and 27 normalisation floats (12 means, 12 scales, output "
              "scale, limit, velocity unit).");
    return PW_EINVAL;
  }
  for (size_t k = 0; k < n_norm; ++k)
    if (!std::isfinite(norm[k])) {
      set_error("Closure normalisation must be finite.");
      return PW_EINVAL;
    }
  if (!(norm[26] > 0.0f) || !(norm[25] > 0.0f)) {
    set_error("Closure velocity unit and limit must be positive.");
    return PW_EINVAL;
  }
  char block_norm[closure::kNetBlockBytes - closure::kNormOffset] = {};
  std::memcpy(block_norm, norm, n_norm * sizeof(float));
  hipError_t e = hipMemcpyAsync(b->closure_net, mlp_packed(net), mlp_packed_bytes(), hipMemcpyDeviceToDevice, b->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(static_cast<char*>(b->closure_net) + closure::kNormOffset, block_norm, sizeof block_norm,
                       hipMemcpyHostToDevice, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return device_error(e, "batch closure attach");
  std::memcpy(b->closure_norm, norm, n_norm * sizeof(float));
  b->closure_on = 1;
  b->closure_fresh = 0;
  e = apply_closure_mode(b);
  return e == hipSuccess ? PW_OK : device_error(e, "batch closure attach");
}

pw_status pw_batch_force(pw_batch* b, double* out, size_t out_doubles) {
  if (!b || !out || out_doubles < 3 * (size_t)b->count) {
    set_error("Force output needs 3 doubles per tunnel.");
    return PW_EINVAL;
  }
  double* dev = nullptr;
  hipError_t e = hipMallocAsync((void**)&dev, 3 * (size_t)b->count * sizeof(double), b->stream);
  if (e != hipSuccess) return device_error(e, "batch force buffer");
  launch_force(b, dev);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, dev, 3 * (size_t)b->count * sizeof(double), hipMemcpyDeviceToHost, b->stream);
  (void)hipFreeAsync(dev, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  return e == hipSuccess ? PW_OK : device_error(e, "batch force");
}

pw_status pw_batch_set_closure_placement(pw_batch* b, uint32_t placement) {
  if (!b || placement > PW_CLOSURE_IN_STEP) {
    set_error("Unknown closure placement.");
    return PW_EINVAL;
  }
  b->closure_placement = placement;
  b->closure_fresh = 0;
  if (!b->closure_on) return PW_OK;
  const hipError_t e = apply_closure_mode(b);
  return e == hipSuccess ? PW_OK : device_error(e, "batch closure placement");
}

pw_status pw_batch_set_closure_interval(pw_batch* b, uint32_t every) {
  if (!b || every == 0) {
    set_error("The closure interval must be at least one step.");
    return PW_EINVAL;
  }
  b->closure_every = every;
  if (!b->closure_on) return PW_OK;
  const hipError_t e = apply_closure_mode(b);
  return e == hipSuccess ? PW_OK : device_error(e, "batch closure interval");
}

}  // extern "C"

namespace {

// `count` steps without closure evaluations: whole even/odd pairs replay a captured graph.
pw_status step_plain(pw_batch* b, uint32_t count, int write_macro_on_last) {
  uint32_t done = 0;
  hipError_t e;
  // Whole even/odd pairs replay a captured graph; the remainder and a macro step launch directly.
  const uint32_t direct_tail = (write_macro_on_last && count) ? 1u : 0u;
  while (count - done - direct_tail >= 2) {
    hipGraphExec_t& g = b->graph[b->parity];
    if (!g && (e = capture_pair(b, b->parity, &g)) != hipSuccess) return device_error(e, "batch graph capture");
    if ((e = hipGraphLaunch(g, b->stream)) != hipSuccess) return device_error(e, "batch graph launch");
    done += 2;
    b->steps += 2;
  }
  while (done < count) {
    const bool macro = write_macro_on_last && done + 1 == count;
    launch_step(b, b->parity, macro);
    b->parity ^= 1;
    done += 1;
    b->steps += 1;
  }
  if ((e = hipGetLastError()) != hipSuccess) return device_error(e, "batch step");
  return PW_OK;
}

}  // namespace

bool pw::batch_field(pw_batch* b, uint32_t tunnel, const float4** macro, const uint32_t** geo, pw_layout* layout,
                     hipStream_t* stream) {
  if (!b || tunnel >= b->count) return false;
  *macro = b->host[tunnel].macro;
  *geo = b->host[tunnel].geo;
  *layout = b->layouts[tunnel];
  *stream = b->stream;
  return true;
}

extern "C" {

pw_status pw_batch_step_sampled(pw_batch* b, uint32_t count, uint32_t every, double* out, size_t out_doubles) {
  if (!b || !out || every == 0) {
    set_error("Sampling needs a batch, a positive interval and an output buffer.");
    return PW_EINVAL;
  }
  size_t samples = 0;
  for (uint64_t k = b->steps + 1; k <= b->steps + count; ++k) samples += (k % every) == 0;
  if (out_doubles < 4 * samples * b->count) {
    set_error("Sample output needs 4 doubles per sample and tunnel.");
    return PW_EINVAL;
  }
  double* hist = nullptr;
  hipError_t e;
  if (samples && (e = hipMallocAsync((void**)&hist, samples * 3 * b->count * sizeof(double), b->stream)) != hipSuccess)
    return device_error(e, "batch sample history");
  std::vector<uint64_t> stamps;
  uint32_t done = 0;
  while (done < count) {
    // Steps up to the next sample, through the graph path.
    const uint64_t next = ((b->steps / every) + 1) * every;
    const uint32_t chunk = (uint32_t)std::min<uint64_t>(next - b->steps, count - done);
    if (pw_status st = pw_batch_step(b, chunk, 0)) {
      if (hist) (void)hipFreeAsync(hist, b->stream);
      return st;
    }
    done += chunk;
    if (b->steps % every == 0) {
      launch_force(b, hist + 3 * (size_t)b->count * stamps.size());
      stamps.push_back(b->steps);
    }
  }
  std::vector<double> host(3 * samples * b->count);
  e = hipGetLastError();
  if (e == hipSuccess && samples) e = hipMemcpyAsync(host.data(), hist, host.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream);
  if (hist) (void)hipFreeAsync(hist, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return device_error(e, "batch samples");
  // Output layout: [tunnel][sample] x (step, fx, fy, fz).
  for (uint32_t t = 0; t < b->count; ++t)
    for (size_t j = 0; j < samples; ++j) {
      double* o = out + 4 * (t * samples + j);
      o[0] = (double)stamps[j];
      o[1] = host[3 * (j * b->count + t)];
      o[2] = host[3 * (j * b->count + t) + 1];
      o[3] = host[3 * (j * b->count + t) + 2];
    }
  return PW_OK;
}

pw_status pw_batch_read_macro(pw_batch* b, uint32_t tunnel, float* out, size_t out_floats) {
  if (!b || tunnel >= b->count || !out || out_floats < 4 * (size_t)b->host[tunnel].p.cells) {
    set_error("Invalid tunnel index or output size.");
    return PW_EINVAL;
  }
  hipError_t e = hipMemcpyAsync(out, b->host[tunnel].macro, 16 * (size_t)b->host[tunnel].p.cells, hipMemcpyDeviceToHost, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  return e == hipSuccess ? PW_OK : device_error(e, "batch macro readback");
}

pw_status pw_batch_read_populations(pw_batch* b, uint32_t tunnel, void* out, size_t out_bytes) {
  const uint32_t elem = b && b->params.storage == PW_STORAGE_F32 ? 4 : 2;
  if (!b || tunnel >= b->count || !out || out_bytes < (size_t)kQ * b->host[tunnel].p.cells * elem) {
    set_error("Invalid tunnel index or output size.");
    return PW_EINVAL;
  }
  const StepParams& p = b->host[tunnel].p;
  const hipError_t e = set_to_host_planes(b->host[tunnel].pops[b->parity], out, p, elem, b->stream);
  return e == hipSuccess ? PW_OK : device_error(e, "batch population readback");
}

uint64_t pw_batch_step_count(const pw_batch* b) { return b ? b->steps : 0; }
uint64_t pw_batch_cells(const pw_batch* b) {
  uint64_t n = 0;
  if (b)
    for (const auto& t : b->host) n += t.p.cells;
  return n;
}

pw_status pw_batch_time_steps(pw_batch* b, uint32_t count, float* elapsed_ms) {
  if (!b || !elapsed_ms) {
    set_error("Null argument.");
    return PW_EINVAL;
  }
  hipError_t e = hipEventRecord(b->t0, b->stream);
  if (e != hipSuccess) return device_error(e, "batch timing");
  if (pw_status st = pw_batch_step(b, count, 0)) return st;
  if ((e = hipEventRecord(b->t1, b->stream)) != hipSuccess || (e = hipEventSynchronize(b->t1)) != hipSuccess ||
      (e = hipEventElapsedTime(elapsed_ms, b->t0, b->t1)) != hipSuccess)
    return device_error(e, "batch timing");
  return PW_OK;
}

void pw_batch_destroy(pw_batch* b) {
  if (!b) return;
  free_batch(b);
  delete b;
}

}  // extern "C"
