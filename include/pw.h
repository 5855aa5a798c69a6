/*
 * pw.h -- C-ABI of libpocketwind, the MI355X (gfx950) implementation of pocketwind's
 * D3Q19 wind-tunnel hot path. Plain pointers and sizes only; no HIP or torch types.
 *
 * Every entry point names the reference interface it replaces (pocketwind repository,
 * file:line). Calls on one handle are not thread-safe; separate handles are independent.
 * All device work of a handle is ordered on that handle's own HIP stream.
 */
#ifndef PW_H
#define PW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PW_ABI_VERSION 8

/* Errors (src/lbm/gpu-resources.ts:3-27, gpu-force.ts:192-205). EINVAL is the reference's
 * RangeError, ENOMEM its AllocationFailure, EBUSY its 'busy' readback. */
typedef enum {
  PW_OK = 0,
  PW_EINVAL = 1,
  PW_ENOMEM = 2,
  PW_EBUSY = 3,
  PW_EDEVICE = 4,
  PW_ESTATE = 5 /* a solver handle that was destroyed (or never created) */
} pw_status;

/* Cell kinds in bits 0-2 of a geometry flag; bit 3+i marks that incoming direction i pulls
 * from a wall (src/geometry/voxel-grid.ts:3-5, docs/numerics.md "Macro buffer ... walls"). */
enum { PW_FLUID = 0, PW_SOLID = 1, PW_SUPPORT = 2, PW_FLOOR = 3, PW_EQ = 4 };
#define PW_CELL_TYPE_MASK 7u
#define PW_NO_Q 0xffffffffu

typedef enum { PW_OP_TRT = 0, PW_OP_REG = 1 } pw_operator;           /* app/settings Operator */
typedef enum { PW_WALLS_BB = 0, PW_WALLS_BOUZIDI = 1 } pw_walls;     /* app/settings Walls */
/* Population storage: f32, f16 scaled by 2^15, or bf16; both 16-bit forms store f - w with
 * lowbias32 stochastic rounding (src/lbm/shaders/prelude-f16.wgsl:7-22). */
typedef enum { PW_STORAGE_F32 = 0, PW_STORAGE_F16 = 1, PW_STORAGE_BF16 = 2 } pw_storage;
/* Near-wall closure under LES: the reference's Werner-Wengle wall law
 * (src/lbm/shaders/stream_collide.wgsl:8-80), or the learned MLP closure loaded with
 * pw_solver_set_wall_closure, which falls back to the wall law outside its training range. */
typedef enum { PW_WALL_LAW = 0, PW_WALL_LEARNED = 1 } pw_wall_model;
/* Side boundaries y = 0, ny - 1 and z = 0, nz - 1: periodic as in the reference
 * (src/lbm/shaders/common.wgsl:62-68), or free-slip: halfway specular reflection, a pull that
 * would cross a side reads the mirrored direction of the source cell unshifted in that axis.
 * Free-slip sides need every wall cell at least two cells from the four sides (so wall links,
 * interpolated-wall second nodes and wall-function gathers never cross one). */
typedef enum { PW_SIDES_PERIODIC = 0, PW_SIDES_FREE_SLIP = 1 } pw_sides;
/* Outlet plane x = nx - 1: the reference's equilibrium (pressure) outlet on EQ cells
 * (src/lbm/shaders/stream_collide.wgsl:82-107), or zero-gradient: its EQ cells are stepped as
 * fluid cells whose pulls from x = nx take x = nx - 1 (the unknown populations are copied from
 * the upstream neighbour after streaming); pulls from wall cells there bounce back halfway.
 * Needs nx >= 2. The learned wall closure and batched tunnels take the defaults only. */
typedef enum { PW_OUTLET_EQUILIBRIUM = 0, PW_OUTLET_ZERO_GRADIENT = 1 } pw_outlet;

typedef struct {
  uint32_t nx, ny, nz;
} pw_layout;

/* SolverParams (src/lbm/gpu-solver.ts:19-30). */
typedef struct {
  double tau0;         /* baseline relaxation time, > 1/2 after f32 rounding */
  double u_in;         /* inlet speed along +x, cells per step */
  uint32_t les;        /* Smagorinsky sub-grid model (Cs = 0.1) and wall model */
  uint32_t op;         /* pw_operator */
  uint32_t walls;      /* pw_walls */
  uint32_t storage;    /* pw_storage */
  uint32_t ramp_steps; /* linear inlet ramp length; 0 = full speed from the start */
  uint32_t wall_model; /* pw_wall_model */
  uint32_t sides;      /* pw_sides; no reference counterpart (the reference is periodic) */
  uint32_t outlet;     /* pw_outlet; no reference counterpart (the reference's is equilibrium) */
  /* Smallest tangential slip speed at which the wall law prescribes its shear; below it a wall
   * cell takes the bulk collision (stream_collide.wgsl:40 uses 1e-7). 0 selects the default:
   * 1e-7 with f32 storage; 1e-4 with f16 or bf16 storage, whose stored velocities carry
   * rounding noise of about 1e-5 (f16), so that at a stagnation point, where the slip vanishes,
   * the noise does not switch the cell to the law's shear along a random direction. Finite and
   * non-negative. No reference counterpart (the reference fixes 1e-7). */
  float slip_floor;
} pw_params;

/* ForceSample (src/lbm/gpu-force.ts:11); components accumulated and returned in FP64. */
typedef struct {
  uint64_t step;
  double fx, fy, fz;
} pw_force_sample;

typedef struct pw_solver pw_solver;
typedef struct pw_batch pw_batch;
typedef struct pw_mlp pw_mlp;

/* Last error text of the calling thread (the message the reference puts in RangeError). */
const char *pw_last_error_message(void);
int pw_abi_version(void);
/* Number of visible HIP devices (0 without a GPU; no kernel is launched). */
int pw_device_count(void);

/* packSolverParams (src/lbm/gpu-solver.ts:80-116): the 48-byte f32 uniform of struct Params
 * (src/lbm/shaders/common.wgsl:31-46) as the reference packs it: u32 nx, ny, nz, cells; f32
 * u_in, nu = fround((fround(tau0) - 1/2) / 3), 18 sqrt(2) Cs^2, lambda = 3/16; u32 les,
 * walls == Bouzidi, write_macro, step. PW_EINVAL (RangeError) for the same inputs the
 * reference rejects: empty grids, indices outside u32, tau0 not above 1/2 after f32 rounding,
 * non-finite f32 speeds. The step kernels take these values (and constants derived from them)
 * as launch arguments. Host only: no device call. */
#define PW_PARAMS_BYTES 48
pw_status pw_pack_params(const pw_layout *layout, const pw_params *params, int write_macro,
                         uint64_t step, void *out, size_t out_bytes);

/* GpuSolver.create (src/lbm/gpu-solver.ts:158-335): validates the packed geometry
 * [flags N][q offsets N][f32 q per link] (src/lbm/geometry-contract.ts:39-66) and the
 * boundary list (src/lbm/gpu-force.ts:76-81), copies both to the device, allocates two
 * population sets, geometry and macros, and initialises both sets to equilibrium.
 * On allocation failure returns PW_ENOMEM with nothing left allocated. */
pw_status pw_solver_create(const pw_layout *layout, const uint32_t *geometry,
                           size_t geometry_words, const uint32_t *boundary_cells,
                           size_t n_boundary, const pw_params *params, pw_solver **out);

/* GpuSolver.step(encoder, count, writeMacroOnLast) (src/lbm/gpu-solver.ts:378-395).
 * Stream-ordered and asynchronous; count 0 launches nothing. */
pw_status pw_solver_step(pw_solver *s, uint32_t count, int write_macro_on_last);

/* GpuSolver.encodeForce (src/lbm/gpu-solver.ts:448-451, gpu-force.ts:175-189): samples the
 * mean momentum exchange of the two latest states; skipped while a sample is unread. */
pw_status pw_solver_encode_force(pw_solver *s);

/* GpuSolver.readForce (src/lbm/gpu-solver.ts:454-456, gpu-force.ts:192-205): waits for the
 * encoded sample; PW_EBUSY when none is encoded. */
pw_status pw_solver_read_force(pw_solver *s, pw_force_sample *out);

/* GpuSolver.reset (src/lbm/gpu-solver.ts:347-355): both sets back to equilibrium, state 0,
 * step count 0, an unread force sample discarded. */
pw_status pw_solver_reset(pw_solver *s);

/* GpuSolver.stepCount (src/lbm/gpu-solver.ts:342-344). */
uint64_t pw_solver_step_count(const pw_solver *s);

/* GpuSolver.destroy (src/lbm/gpu-solver.ts:459-464). Waits for queued work. The library keeps a
 * set of live solver handles: after destroy, calls with the handle return PW_ESTATE (getters
 * return 0) without dereferencing it, until a later create happens to reuse the address. */
void pw_solver_destroy(pw_solver *s);

/* macroBuffer readback (docs/numerics.md "Macro buffer"): 4*N floats (rho, ux, uy, uz). */
pw_status pw_solver_read_macro(pw_solver *s, float *out, size_t out_floats);

/* populationBuffers readback (src/lbm/gpu-solver.ts:338-340, docs/numerics.md "Population
 * storage"): the latest state as 19 direction-major planes, A (0-6), B (7-12), C (13-18)
 * back to back, in storage format (f32, or raw 16-bit words). */
pw_status pw_solver_read_populations(pw_solver *s, void *out, size_t out_bytes);
/* The test-side queue.writeBuffer into populationBuffers (tests/browser/force.spec.ts:145-153):
 * overwrite the latest state with raw storage-format data of the same size. */
pw_status pw_solver_write_populations(pw_solver *s, const void *in, size_t in_bytes);
size_t pw_solver_population_bytes(const pw_solver *s);

/* injectNaNForRunner (src/lbm/gpu-solver.ts:358-375): write quiet-NaN bits into direction 0
 * of the first fluid cell of the latest state. */
pw_status pw_debug_inject_nan(pw_solver *s);

/* Instability guard (src/lbm/nan-guard.ts:151-196, shaders/nan_guard.wgsl:4-17), fused into
 * the step kernel instead of a separate pass: advances `count` steps in batches that end on
 * multiples of PW_GUARD_INTERVAL; every batch writes the macros on its last step, and on a
 * batch ending at a multiple that step also flags any fluid cell whose density is not finite
 * or lies outside [0.5, 1.5]. A flagged check resets the solver (pw_solver_reset) and
 * *unstable_step receives the step of the check; otherwise it is 0. */
#define PW_GUARD_INTERVAL 200
pw_status pw_solver_advance_guarded(pw_solver *s, uint32_t count, uint64_t *unstable_step);

/* The binary blocks of the legacy VTK field file (src/io/vtk.ts:27-80) for the macro buffer,
 * with velocities scaled as latticeToField does (src/io/field-export.ts:22-34): rho_be gets
 * 4*N bytes, big-endian float32 densities; velocity_be gets 12*N bytes, big-endian float32
 * (ux, uy, uz) * velocity_scale per cell, each product of the float32 value and the double
 * scale rounded once to float32. velocity_scale = speedMps / U_LU must be positive and
 * finite. The host writes the text lines around the blocks. */
pw_status pw_solver_field_vtk_blocks(pw_solver *s, double velocity_scale, void *rho_be, size_t rho_bytes,
                                     void *velocity_be, size_t velocity_bytes);

/* Device views for co-resident kernels (particles, interop). Pointers stay valid until
 * destroy; `populations` is the latest state and changes after every step. On the device each
 * x-row keeps its 19 directions as runs of R = run_elements cells (64 when nx is a multiple of
 * 64, else nx): element (i, x, y, z) lies at i R + x + 18 R floor(x / R) + 19 nx (y + ny z).
 * The readbacks above return the reference's planes instead. */
typedef struct {
  const void *populations;
  const float *macro;   /* float4 per cell */
  const uint32_t *geometry;
  void *stream;         /* hipStream_t of the handle */
  size_t run_elements;  /* R */
} pw_device_state;
pw_status pw_solver_device_state(pw_solver *s, pw_device_state *out);

/* Step-kernel variants: one cell per lane with cached population streams, non-temporal
 * loads and stores, or cached loads with non-temporal stores. AUTO times all three at block
 * sizes 64, 128 and 256 on the handle's buffers (at create, or when set later, restoring the
 * state afterwards) and keeps the fastest; an explicit variant runs at block size 256. No
 * reference counterpart: the reference has one WGSL kernel
 * (src/lbm/shaders/stream_collide.wgsl) with a fixed workgroup size. */
typedef enum {
  PW_KERNEL_AUTO = 0,
  PW_KERNEL_CELL = 1,
  PW_KERNEL_CELL_NT = 2,
  PW_KERNEL_CELL_NT_STORES = 3
} pw_kernel;
pw_status pw_solver_set_kernel(pw_solver *s, uint32_t kernel);
uint32_t pw_solver_kernel(const pw_solver *s);
/* An explicit step variant: kernel PW_KERNEL_CELL .. PW_KERNEL_CELL_NT_STORES, `block` threads
 * per block (64, 128 or 256) and `xcd_chunks` chunks per XCD (a power of two up to 1024),
 * without timing. */
pw_status pw_solver_set_step_variant(pw_solver *s, uint32_t kernel, uint32_t block,
                                     uint32_t xcd_chunks);
/* Test hook: with `on`, step with one cell per lane even when nx is a multiple of 64 (whose
 * steps otherwise give each wavefront one 64-cell x-row), in the same population layout. The
 * two paths run the same collision code; comparing them pins the row path to the per-cell
 * path. No reference counterpart. */
pw_status pw_debug_cell_path(pw_solver *s, int on);
/* Rounding of f16 population stores. HARDWARE (the default): the gfx950 stochastic conversion
 * with this library's per-cell, per-step random words. REFERENCE: encode_population of
 * src/lbm/shaders/prelude-f16.wgsl:7-22, the random word lowbias32 of the value's f32 bits, the
 * cell and the direction, so that stored bits can be compared with the reference's for equal f32
 * values (about 20 integer operations per store instead of one). Setting it resets the solver
 * (both sets to the initial state under the new rounding, step 0). PW_EINVAL for other storage
 * formats. No reference counterpart as an option: the reference has only its own rule. */
typedef enum { PW_ROUNDING_HARDWARE = 0, PW_ROUNDING_REFERENCE = 1 } pw_rounding;
pw_status pw_solver_set_rounding(pw_solver *s, uint32_t rounding);
/* Threads per block of the step launches. */
uint32_t pw_solver_block_size(const pw_solver *s);
/* Chunks of the lattice each of the 8 XCDs steps (1: one contiguous eighth each; more: chunks
 * spread over the lattice, XCD k taking chunks k, k + 8, ...), chosen with the variant by AUTO
 * among 1, 16, 64 and 256. */
uint32_t pw_solver_xcd_chunks(const pw_solver *s);

/* The validation stepper stepWithForces (src/app/case-run.ts:154-169) without a host round
 * trip per sample: queue `count` steps and, after each step whose count is a multiple of
 * `every`, one FP64 force sample (the encodeForce semantics) into a device history, read back
 * once at the end into `out` as 4 doubles per sample (step, fx, fy, fz). Does not touch the
 * encodeForce/readForce slot. Blocks until done. */
pw_status pw_solver_step_sampled(pw_solver *s, uint32_t count, uint32_t every, double *out,
                                 size_t out_doubles);

/* pw_solver_step_sampled without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No reference counterpart. The same steps, force samples and output as
 * pw_solver_step_sampled, with the sample history kept in a fixed device array of the library
 * (4096 samples of 3 doubles, 96 KiB per GPU, held until the process ends) instead of a per-call
 * allocation; a call asking for more than 4096 samples returns PW_EINVAL before stepping.
 * Every call on one GPU shares that array: calls that overlap on one GPU (from several
 * threads or solvers at once) race. Blocks until done. */
pw_status pw_solver_step_sampled_noalloc(pw_solver *s, uint32_t count, uint32_t every,
                                         double *out, size_t out_doubles);

/* Wait for all queued work of the handle. */
pw_status pw_solver_synchronize(pw_solver *s);

/* Measurement (bench/precision-hash.mjs:102-150 timestamps the first and last pass): run
 * `count` steps without macro writes between two HIP events on the handle's stream and
 * return the elapsed milliseconds. */
pw_status pw_solver_time_steps(pw_solver *s, uint32_t count, float *elapsed_ms);

/* measureCopyBandwidth (src/app/bench-page.ts:204-242): device-to-device float4 copy of
 * `bytes`, `passes` times; returns GB/s of read+write traffic. */
pw_status pw_copy_bandwidth(size_t bytes, uint32_t passes, double *gbps);

/* VoxelGrid (src/geometry/voxel-grid.ts:7-15); arrays are malloc'd by pw_voxelize and
 * released by pw_voxel_grid_free. */
typedef struct {
  uint32_t *geometry; /* [flags N][q offsets N][f32 q per link] */
  size_t geometry_words;
  uint32_t *boundary_cells; /* fluid cells with a SOLID pull neighbour, in cell order */
  size_t n_boundary;
  uint64_t solid_cells;
  uint64_t frontal_cells;
  double blockage;
  int32_t watertight;
  uint64_t missed_links;
} pw_voxel_grid;

/* voxelize(body, supports, layout, {floor, linkFractions}) (src/geometry/voxelize.ts:172-199,
 * wall-links.ts:99-131) on the GPU: bit-exact flags, q offsets, link fractions and boundary
 * list. body/supports are 9 floats per triangle in lattice cells; supports may be NULL. */
pw_status pw_voxelize(const float *body, size_t body_tris, const float *supports,
                      size_t support_tris, const pw_layout *layout, int floor, int link_fractions,
                      pw_voxel_grid *out);
void pw_voxel_grid_free(pw_voxel_grid *grid);

/* Streakline particles bound to a solver's macro field and flags (createParticles,
 * src/render/particles.ts:30-138; seeds src/render/particle-layout.ts:66-112). A record is
 * 8 floats: pos.xyz, age, prev.xyz, generation. count may be 1 .. 2^24 (the reference caps
 * it at 65536). Updates run on the solver's stream after its queued steps. Destroy the
 * particles before their solver. */
typedef struct pw_particles pw_particles;
pw_status pw_particles_create(pw_solver *solver, uint32_t count, uint32_t seed,
                              const double body_min[3], const double body_max[3],
                              pw_particles **out);
/* particles.update(encoder, steps) (src/render/particles.ts:113-124,
 * src/render/shaders/particles_update.wgsl:46-60). */
pw_status pw_particles_update(pw_particles *p, float steps);
pw_status pw_particles_read(pw_particles *p, float *out, size_t out_floats);
pw_status pw_particles_write(pw_particles *p, const float *in, size_t in_floats);
/* The seed cross-section (y0, z0, y1, z1) and the fallback inlet node (x, y, z). */
pw_status pw_particles_info(const pw_particles *p, float region[4], float fallback[3]);
void pw_particles_destroy(pw_particles *p);
/* The same particles bound to tunnel `tunnel` of a batch (no reference counterpart: the
 * reference has one solver); updates run on the batch's stream. Destroy them before the batch. */
pw_status pw_particles_create_batch(pw_batch *b, uint32_t tunnel, uint32_t count, uint32_t seed,
                                    const double body_min[3], const double body_max[3],
                                    pw_particles **out);
/* createParticles(device, format, frame, {grid, bodyBounds, macroBuffer, count, seed})
 * (src/render/particles.ts:30-138) over a caller's field: `geometry` is the grid's host flag
 * array (at least nx*ny*nz words, copied to the device), `macro` a device buffer of 4 floats
 * (rho, ux, uy, uz) per cell that the particles read on every update. Updates run on a stream
 * of the particles' own: finish writing `macro` (for example hipDeviceSynchronize) before an
 * update, and call pw_particles_synchronize before changing it again. */
pw_status pw_particles_create_field(const pw_layout *layout, const uint32_t *geometry,
                                    size_t geometry_words, const float *macro, uint32_t count,
                                    uint32_t seed, const double body_min[3],
                                    const double body_max[3], pw_particles **out);
/* Waits for the particles' queued updates (the reference's queue submission completing). */
pw_status pw_particles_synchronize(pw_particles *p);

/* Batched tunnels: `count` independent tunnels of any sizes advanced by one launch per step,
 * even/odd step pairs replayed from a captured hipGraph. No reference counterpart (the
 * reference allows one live solver, src/app/solver-slot.ts:14-44); each tunnel follows
 * pw_solver semantics exactly and is bit-identical to the same tunnel run alone. Tunnel t
 * takes geometries[t] (the packed geometry of pw_solver_create), boundary_cells[t], tau0[t]
 * and u_in[t]; operator, walls, LES, storage and wall model come from `shared` (no ramp,
 * periodic sides, equilibrium outlet). */
pw_status pw_batch_create(uint32_t count, const pw_layout *layouts, const uint32_t *const *geometries,
                          const size_t *geometry_words, const uint32_t *const *boundary_cells,
                          const size_t *n_boundary, const double *tau0, const double *u_in,
                          const pw_params *shared, pw_batch **out);
pw_status pw_batch_step(pw_batch *b, uint32_t count, int write_macro_on_last);
/* The step variant the batch chose at create (timing each cache policy with 1, 16, 64 and 256
 * chunks per XCD, blocks of 64, 128 and 256 threads, one block per tile and persistent, or
 * PW_BATCH_KERNEL=<cell|cell-nt|cell-nt-stores>/<chunks>[/<block>][/persistent]): a pw_kernel value, and
 * the chunks per XCD in *xcd_chunks (may be NULL). */
uint32_t pw_batch_step_variant(const pw_batch *b, uint32_t *xcd_chunks);
/* Resident blocks of the persistent form of the batched step when the tuner chose it (or
 * PW_BATCH_KERNEL=<kernel>/<chunks>/persistent fixed it): that many blocks, the collision's
 * blocks per CU on every CU, walk the tiles of the one-block-per-tile launch in its order, no
 * block waiting for another; 0 when each step launches one block per tile. Steps that write the
 * macros and the closure evaluated in the step always launch one block per tile. No reference
 * counterpart. */
uint32_t pw_batch_persistent_grid(const pw_batch *b);
/* Threads per block of the batched step, which is also the cells per tile (64, 128 or 256),
 * chosen with the variant at create (PW_BATCH_KERNEL=<kernel>/<chunks>/<block> fixes it). No
 * reference counterpart. */
uint32_t pw_batch_block_size(const pw_batch *b);
/* As pw_solver_step_sampled for every tunnel: `out` receives [tunnel][sample] x
 * (step, fx, fy, fz), 4 doubles each. */
pw_status pw_batch_step_sampled(pw_batch *b, uint32_t count, uint32_t every, double *out,
                                size_t out_doubles);
pw_status pw_batch_read_macro(pw_batch *b, uint32_t tunnel, float *out, size_t out_floats);
pw_status pw_batch_read_populations(pw_batch *b, uint32_t tunnel, void *out, size_t out_bytes);
uint64_t pw_batch_step_count(const pw_batch *b);
uint64_t pw_batch_cells(const pw_batch *b);
pw_status pw_batch_time_steps(pw_batch *b, uint32_t count, float *elapsed_ms);
void pw_batch_destroy(pw_batch *b);
/* The learned wall closure of a batch created with wall_model PW_WALL_LEARNED: one network
 * and normalisation for every tunnel, as pw_solver_set_closure / _interval (no reference
 * counterpart). Each tunnel's closure runs on its own wall cells before the steps that
 * evaluate it; a batched tunnel stays bit-identical to the same tunnel run alone with the
 * same closure. */
pw_status pw_batch_set_closure(pw_batch *b, const pw_mlp *net, const float *norm, size_t n_norm);
pw_status pw_batch_set_closure_interval(pw_batch *b, uint32_t every);
/* Closure placement of every tunnel of the batch: PW_CLOSURE_KERNEL (one launch over all the
 * tunnels' wall cells) or PW_CLOSURE_IN_STEP (row tunnels; the others keep the kernel); other
 * values are rejected. */
pw_status pw_batch_set_closure_placement(pw_batch *b, uint32_t placement);
/* The momentum-exchange force of every tunnel now (fx, fy, fz per tunnel, FP64; the mean over
 * both states as a force sample), waiting for the batch's queued work. */
pw_status pw_batch_force(pw_batch *b, double *out, size_t out_doubles);

/* Learned wall closure network: x (12 features) -> 64 -> 64 -> 3 with SiLU activations, bf16
 * matrix-core products accumulated in f32. No reference counterpart: the reference's wall
 * treatment is the wall law of src/lbm/shaders/stream_collide.wgsl:1-80, which stays the
 * closure's fallback. Parameters are one f32 vector: W1 (64 x 12, row-major), b1 (64), W2
 * (64 x 64), b2 (64), W3 (3 x 64), b3 (3). All arrays here are host arrays; a handle is not
 * thread-safe. */
#define PW_MLP_INPUTS 12
#define PW_MLP_HIDDEN 64
#define PW_MLP_OUTPUTS 3
#define PW_MLP_PARAMS 5187
pw_status pw_mlp_create(pw_mlp **out); /* all parameters zero */
void pw_mlp_destroy(pw_mlp *m);
/* Sets the f32 parameters (n == PW_MLP_PARAMS) and clears the Adam state. */
pw_status pw_mlp_set_params(pw_mlp *m, const float *params, size_t n);
pw_status pw_mlp_get_params(pw_mlp *m, float *params, size_t n);
/* y (rows x 3) = network(x (rows x 12)). */
pw_status pw_mlp_forward(pw_mlp *m, const float *x, size_t rows, float *y);
/* Gradient of the mean squared error over all rows with respect to the parameters, and the
 * loss (may be NULL). Does not change the parameters. */
pw_status pw_mlp_gradients(pw_mlp *m, const float *x, const float *t, size_t rows, float *grads,
                           float *loss);
/* Learned wall closure in a solver created with wall_model PW_WALL_LEARNED (which needs LES).
 * A wall cell is a fluid cell with a wall link; the wall cells are listed in cell order. Before
 * every step a closure kernel evaluates the attached network on the 12 features of each wall
 * cell (see the features call) and the LES wall collision uses the resulting wall shear
 * stress along the slip in place of the wall law's rho u_tau^2; until a network is attached,
 * and for cells whose normalised features leave [-limit, limit], the wall law stays. */
size_t pw_solver_wall_cell_count(const pw_solver *s);
pw_status pw_solver_wall_cells(pw_solver *s, uint32_t *out, size_t n);
/* Per wall cell of the latest state, 15 floats: velocity / u_ref, velocity / u_ref of the
 * fluid cell one step along the rounded wall normal, mean wall-link fraction, log10 nu, unit
 * wall normal, 6 x the wall links' weight sum, then the wall law's stress vector. */
pw_status pw_solver_closure_features(pw_solver *s, float u_ref, float *out, size_t n_floats);
/* Momentum-exchange force on the body from each wall cell (links to SOLID only, mean of the
 * two latest states, the per-cell terms of the encodeForce sum), 3 doubles per wall cell. */
pw_status pw_solver_wall_forces(pw_solver *s, double *out, size_t n_doubles);
/* Attaches a copy of `net` with its normalisation: 12 feature means, 12 feature scales
 * (normalised = (feature - mean) x scale), output scale (stress = output x scale), limit,
 * velocity unit u_ref (27 floats). */
pw_status pw_solver_set_closure(pw_solver *s, const pw_mlp *net, const float *norm, size_t n_norm);
/* Evaluate the closure on the steps whose count is a multiple of `every` (default 1) and
 * reuse its stresses in between (the first step after attaching or reset always evaluates). */
pw_status pw_solver_set_closure_interval(pw_solver *s, uint32_t every);
/* Where the closure is evaluated. KERNEL (the default): its own kernel over the wall-cell list,
 * on the state the next step reads, every `every` steps. IN_STEP: inside the step kernel, by
 * each wavefront that holds wall cells, every step (the interval is ignored), for the REG +
 * LES step with the reference's boundaries on grids whose nx is a multiple of 64; other
 * solvers keep the kernel. Both give the same stresses bit for bit. In-step evaluation saves
 * the launch and the stress buffer but lengthens the wall wavefronts by the features' gathers
 * and the network, and measures slower than the kernel every step on the 256x128x128 bluff
 * body. OVERLAPPED: the closure kernel of a step k due an evaluation (k a multiple of `every`)
 * runs on a second stream beside step k, reading the same state; step k still uses the previous
 * output and steps k + 1 .. k + every use the new one, so each output is one step older when
 * used than under KERNEL (the first step after attaching or reset evaluates in line, and its
 * output also serves the step that starts the first overlapped evaluation). The latency-bound
 * closure kernel then runs within the HBM-bound step instead of before it. Each step call joins
 * the evaluation it started before it returns. ALONGSIDE: OVERLAPPED's schedule and outputs bit
 * for bit, with the evaluation in step k's own launch: its first blocks run the network over the
 * wall-cell list while the others step, so no second stream or event is involved (the REG step;
 * TRT solvers keep the kernel). No reference counterpart. */
typedef enum {
  PW_CLOSURE_KERNEL = 0,
  PW_CLOSURE_IN_STEP = 1,
  PW_CLOSURE_OVERLAPPED = 2,
  PW_CLOSURE_ALONGSIDE = 3
} pw_closure_placement;
pw_status pw_solver_set_closure_placement(pw_solver *s, uint32_t placement);
/* Test hook: the wall collision takes the given vectors (x, y, z, w per wall cell, in the
 * order of pw_solver_wall_cells) in place of the network's output, until pw_solver_set_closure
 * attaches a network again: w = 1 a wall shear stress used in the wall law's place; w = 2 a
 * momentum source: the cell keeps the wall law and, after each step, its stored post-collision
 * populations gain 3 w_i c_i . v, so it gains v each step and no mass (f32 storage only; the
 * momentum exchange does not count it, so the body's force is the sampled force minus the sum
 * of the sources); w = 0 the wall law. Used to
 * measure what a perfect closure output would give. Learned wall model only. No reference
 * counterpart. */
pw_status pw_debug_set_wall_stress(pw_solver *s, const float *stress, size_t n_floats);
/* The last closure evaluation: 4 floats per wall cell, stress xyz and 1 (used) or 0 (wall
 * law). */
pw_status pw_solver_closure_output(pw_solver *s, float *out, size_t n_floats);
/* Training set (copied to the device). */
pw_status pw_mlp_set_data(pw_mlp *m, const float *x, const float *t, size_t rows);
/* `steps` Adam steps (beta 0.9 / 0.999, eps 1e-8) at learning rate lr, each on a minibatch of
 * `batch` rows drawn uniformly with replacement (seeded), or on the whole set when batch is
 * 0; loss_history (may be NULL) receives each step's minibatch loss. */
pw_status pw_mlp_train(pw_mlp *m, uint32_t steps, uint32_t batch, float lr, uint32_t seed,
                       float *loss_history);

#ifdef __cplusplus
}
#endif
#endif
