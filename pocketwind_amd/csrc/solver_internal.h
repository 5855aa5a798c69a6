// solver_internal.h -- host-side state of a pw_solver handle and shared helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "pw.h"

namespace pw {

// Division of a u32 by a fixed divisor d as a multiply-high and shifts (the round-up
// method with a 33-bit magic number): n / d = (t + ((n - t) >> 1)) >> s with t = mulhi(n, m),
// exact for every u32 n; d == 1 is the identity.
struct FastDiv {
  uint32_t d, m, s;
};

__host__ __device__ __forceinline__ uint32_t fast_div(uint32_t n, const FastDiv& f) {
  if (f.d == 1) return n;
  const uint32_t t = (uint32_t)(((uint64_t)n * f.m) >> 32);
  return (t + ((n - t) >> 1)) >> f.s;
}

// Population layout. Each x-row holds its 19 directions as runs of R cells: R = 64 when nx is
// a multiple of 64 (a run is one wavefront's cells), else R = nx. Element (i, x, y, z) of a
// set lies at
//   i R + x + 18 R floor(x / R) + 19 nx (y + ny z),
// so one direction of a wavefront is one aligned run and a step writes each row's 19 runs
// contiguously, instead of streaming 19 planes that lie far apart (a row-wavefront pull copy
// measured about 16 % faster this way on MI355X). A set is exactly 19 N elements.
inline uint32_t pop_run(uint32_t nx) { return nx % 64 == 0 ? 64u : nx; }

inline FastDiv make_fast_div(uint32_t d) {
  if (d <= 1) return FastDiv{1, 0, 0};
  uint32_t l = 0;
  while (((uint64_t)1 << l) < d) ++l;  // ceil(log2 d)
  const uint64_t m = (((uint64_t)1 << 32) * (((uint64_t)1 << l) - d)) / d + 1;
  return FastDiv{d, (uint32_t)m, l - 1};
}

// Boundary options beyond the reference's (which has periodic y/z sides and a pressure
// outlet). Free-slip sides: a pull that would cross y = 0, ny - 1 or z = 0, nz - 1 takes the
// specular image, i.e. the mirrored direction from the source cell unshifted in that axis
// (halfway specular reflection). Zero-gradient outlet: the plane x = nx - 1 is stepped as fluid
// and its pulls from x = nx take x = nx - 1 instead, which copies the unknown populations from
// the upstream neighbour after streaming.
enum : uint32_t { kBcFreeSlip = 1u, kBcZeroGradient = 2u };

// Per-launch scalars of the step kernel (the reference's 48-byte uniform, struct Params in
// src/lbm/shaders/common.wgsl:31-46, plus constants this kernel precomputes), passed by
// value so a ramped step needs no buffer write.
struct StepParams {
  uint32_t nx, ny, nz;
  uint32_t cells;
  uint32_t run;     // cells per direction run (64, or nx when nx is not a multiple of 64)
  uint32_t x_tile;  // 18 R when R = 64, else 0: element of x = x + x_tile floor(x / 64)
  float u_in;      // inlet speed this step (ramped)
  float nu;        // f32 viscosity, (tau0 - 1/2) / 3 rounded as the reference packs it
  float tau0;      // 3 nu + 1/2 evaluated in f32, as relaxation_time does
  float cs_term;   // 18 sqrt(2) Cs^2
  float lambda;    // TRT magic parameter 3/16
  float inv_tau0;  // 1 / tau0 in f32: the relaxation rate without LES
  float wm0;       // TRT odd rate 1 / (lambda / (tau0 - 1/2) + 1/2) in f32
  float slip_floor;  // smallest slip speed the wall law acts on (pw_params::slip_floor resolved)
  FastDiv div_nx, div_ny;
  uint32_t bc;  // kBcFreeSlip | kBcZeroGradient, or 0 for the reference's periodic y/z sides
  uint32_t rows;  // 1: one wavefront per 64-cell x-row (nx a multiple of 64); 0: one cell per lane
  // Row path only: one descriptor word per 64-cell run (wave_descriptors), or null to read
  // every cell's flag.
  const uint32_t* wave_desc;
  // Learned wall closure (null: the wall law): shear stress per wall cell (xyz, w > 0 when
  // valid) and the wall-cell index of every cell, from the closure kernel; or, on the row path,
  // the network's device block (matrix-core copy, then normalisation), evaluated inside the step.
  const float4* closure_tau;
  const uint32_t* closure_slot;
  const void* closure_net;
  // With closure_net: the block order of the step launch, blocks holding wall cells first.
  const uint32_t* tile_order;
  // The wall cells' static closure geometry (closure::WallGeo records, by wall-cell index);
  // read with closure_slot by the closure evaluated in the step.
  const void* closure_geo;
  // The closure evaluated beside the step (non-null closure_out): the launch's first
  // closure_blocks blocks run the network (closure_net) over the closure_cells wall cells
  // (closure_geo) on the state the step reads, into closure_out, while the other blocks step
  // with closure_tau.
  float4* closure_out;
  uint32_t closure_cells, closure_blocks;
};

// Descriptor of a 64-cell run on the row path. 0: every cell is a fluid cell without wall
// links (flag 0). kWaveEqFirst / kWaveEqLast: the same, except that the first / last cell is an
// EQ cell (flag kEq, the inlet and outlet planes). kWaveGeneral: anything else; the step reads
// each cell's flag. Bits kWaveBaseShift and up: the number of wall cells (fluid cells with wall
// links) in the runs before this one, i.e. the wall-cell list index of the run's first wall
// cell, so a wall lane finds its entry of a per-wall-cell buffer without a per-cell index.
enum : uint32_t { kWaveGeneral = 1u, kWaveEqFirst = 2u, kWaveEqLast = 4u };
constexpr uint32_t kWaveBaseShift = 3;
constexpr uint64_t kMaxRowWallCells = (uint64_t)1 << (32 - kWaveBaseShift);  // wall cells a row lattice may hold
// The descriptors of a lattice whose nx is a multiple of 64, from its cell flags.
std::vector<uint32_t> wave_descriptors(const pw_layout& layout, const uint32_t* flags);

void set_error(const std::string& message);
pw_status device_error(hipError_t e, const char* what);

// Host-side checks of one tunnel exactly as pw_solver_create applies them (geometry contract,
// boundary list, f32 parameters). Fills the full-speed step parameters.
pw_status validate_tunnel(const pw_layout& layout, const uint32_t* geometry, size_t geometry_words,
                          const uint32_t* boundary_cells, size_t n_boundary, const pw_params& params,
                          StepParams* full);

}  // namespace pw

struct pw_solver;
struct pw_mlp;
struct pw_batch;

namespace pw {
// Learned wall closure (closure.hip): the network's device copy for the matrix cores, and
// the per-step closure launch, which reads the state the next step reads and points
// p.closure_tau / p.closure_slot at its output when the solver has a network attached.
const void* mlp_packed(const pw_mlp* m);
size_t mlp_packed_bytes();
void launch_closure(pw_solver* s, StepParams& p);
// Joins an overlapped closure evaluation still running beside the last step: the solver's
// stream waits for it and its output becomes the one the next step reads.
void settle_closure(pw_solver* s);
// Adds the prescribed momentum sources to the wall cells of the set a step just wrote.
void launch_sources(pw_solver* s, void* set);
// Build the step's block order for the closure evaluated in the step (blocks that hold wall
// cells first, so their longer wavefronts start with the launch rather than in its middle)
// for the solver's current block size and XCD chunking. Allocates; called when a network is
// attached and when the step variant changes.
hipError_t closure_tile_order(pw_solver* s);
// Bytes of one wall cell's static closure geometry record, and the kernel that writes the
// records of the listed wall cells from the packed geometry.
constexpr size_t kWallGeoBytes = 48;
void launch_wall_geo(hipStream_t stream, const uint32_t* geo, const StepParams& p, const uint32_t* wall_cells,
                     uint32_t n_wall, void* out);
// The macro field, flags, layout and stream of one batched tunnel (false for a bad index).
bool batch_field(pw_batch* b, uint32_t tunnel, const float4** macro, const uint32_t** geo, pw_layout* layout,
                 hipStream_t* stream);
// True while `s` is a handle returned by pw_solver_create and not yet destroyed.
bool solver_live(const pw_solver* s);
// `count` steps, the last writing the macros (and the guard word after them) when asked.
pw_status step_run(pw_solver* s, uint32_t count, int write_macro_on_last);
}  // namespace pw

struct pw_solver {
  int device = 0;
  hipStream_t stream = nullptr;
  pw_layout layout{};
  size_t cells = 0;
  pw_params params{};
  pw::StepParams sp{};
  uint32_t elem_bytes = 4;
  void* pops[2] = {nullptr, nullptr};
  uint32_t* geometry = nullptr;
  size_t geometry_words = 0;
  uint32_t* boundary = nullptr;
  size_t n_boundary = 0;
  float4* macro = nullptr;
  double* partials = nullptr;
  size_t n_partials = 0;
  double* force_dev = nullptr;    // fx, fy, fz
  double* force_host = nullptr;   // pinned
  hipEvent_t force_event = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  int current = 0;
  uint64_t steps = 0;
  int force_encoded = 0;
  uint64_t force_step = 0;
  int ramp_settled = 0;
  long long first_fluid = -1;
  uint32_t* wall_cells = nullptr;  // fluid cells with wall links, in cell order
  uint32_t* wave_desc = nullptr;   // row path: descriptor of each 64-cell run
  size_t n_wall = 0;
  uint32_t* closure_slot = nullptr;  // wall-cell index of every cell (learned model only)
  float4* closure_tau = nullptr;     // closure output per wall cell
  void* closure_geo = nullptr;       // closure::WallGeo per wall cell
  void* closure_net = nullptr;       // device copy of the network's matrix-core parameters
  float closure_norm[32] = {};       // closure::Norm
  int closure_on = 0;
  uint32_t closure_every = 1;  // steps between closure evaluations
  int closure_fresh = 0;       // closure_tau holds an evaluation of this run
  uint32_t closure_placement = 0;  // pw_closure_placement
  // PW_CLOSURE_OVERLAPPED: the evaluation runs on side_stream beside the step that reads the
  // same state, into closure_next, which replaces closure_tau before the following step
  // (PW_CLOSURE_ALONGSIDE: the same, in the step's own launch).
  hipStream_t side_stream = nullptr;
  hipEvent_t state_ready = nullptr, closure_done = nullptr;
  float4* closure_next = nullptr;
  // closure_next is being written and has not replaced closure_tau: 1 on side_stream (the main
  // stream joins closure_done first), 2 by the last step launch (PW_CLOSURE_ALONGSIDE).
  int closure_pending = 0;
  // pw_debug_set_wall_stress momentum sources (w = 2), added after each step (f32 storage)
  float4* closure_source = nullptr;
  int closure_sources = 0;
  uint32_t* tile_order = nullptr;  // step block order for the closure in the step (closure_tile_order)
  uint32_t order_block = 0, order_chunks = 0;  // the variant tile_order was built for
  uint32_t* guard_flag = nullptr;       // the word after the macros (not a separate allocation)
  uint32_t* guard_host = nullptr;       // pinned readback
  uint32_t kernel = 1;  // pw_kernel actually used by step()
  uint32_t block = 256;  // threads per step block
  uint32_t xcd_chunks = 1;  // chunks of the lattice per XCD in the step launch (valid_xcd_chunks)
  uint32_t rounding = 0;    // pw_rounding of f16 stores
};
