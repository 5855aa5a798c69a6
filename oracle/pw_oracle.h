/*
 * pw_oracle.h -- FP64 CPU restatement of pocketwind's D3Q19 wind-tunnel hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the HIP product is compared
 * against. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product library (pocketwind_amd/libpocketwind.so) never links or calls it.
 *
 * Every function names the reference file:line it restates. Reference files are cited
 * relative to the pocketwind repository root.
 */
#ifndef PW_ORACLE_H
#define PW_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Lattice (src/lbm/lattice.ts:2-52). */
extern const int PWO_C[19][3];
extern const double PWO_W[19];
int pwo_opp(int i);

/* Cell kinds (src/geometry/voxel-grid.ts:3-5). */
enum { PWO_FLUID = 0, PWO_SOLID = 1, PWO_SUPPORT = 2, PWO_FLOOR = 3, PWO_EQ = 4 };
#define PWO_TYPE_MASK 7u
#define PWO_NO_Q 0xffffffffu

/* Operators (src/lbm/reference/collision.ts:5). */
enum { PWO_TRT = 0, PWO_REG = 1, PWO_BGK = 2 };
/* Walls: halfway bounce-back or linear Bouzidi (src/lbm/reference/cpu-solver.ts:23-24). */
enum { PWO_BB = 0, PWO_BOUZIDI = 1 };

/* ---- collision.ts ---- */
void pwo_equilibrium(double *target, double rho, const double u[3]);
void pwo_moments(const double *g, double *rho, double u[3]);
void pwo_collide(double *g, const double *geq, double rho, double tau0, int op, int les,
                 const double *force_terms);
double pwo_friction_velocity(double speed, double viscosity);
int pwo_collide_wall(double *g, const double *geq, double rho, const double u[3],
                     const int *wall_incoming, int n_incoming, double viscosity,
                     const double *force_terms);

/* ---- cpu-solver.ts ---- */
typedef struct pwo_solver pwo_solver;
/*
 * flags: one u32 per cell, or the packed geometry [flags N][q offsets N][f32 q ...]
 * (flags_len = its length). body_force: 3 doubles or NULL. eq_state: (rho,ux,uy,uz) or
 * NULL for rest. initial: 4*N doubles (rho,ux,uy,uz per cell) or NULL.
 * Returns NULL and writes a message into err on invalid input.
 */
pwo_solver *pwo_solver_create(uint32_t nx, uint32_t ny, uint32_t nz, double tau0, int op,
                              int les, int walls, const uint32_t *flags, size_t flags_len,
                              const double *body_force, const double *eq_state,
                              const double *initial, char *err, size_t err_len);
/* Boundary options with no reference counterpart (the reference is periodic in y and z with an
 * equilibrium outlet). Free-slip sides: a pull that would cross y = 0, ny - 1 or z = 0, nz - 1
 * reads the specular image (the direction with that component negated) at the source cell held
 * in that axis. Zero-gradient outlet: EQ cells of x = nx - 1 are fluid cells whose pulls from
 * x = nx take x = nx - 1. pwo_solver_create is this with PERIODIC and EQUILIBRIUM. */
enum { PWO_SIDES_PERIODIC = 0, PWO_SIDES_FREE_SLIP = 1 };
enum { PWO_OUTLET_EQUILIBRIUM = 0, PWO_OUTLET_ZERO_GRADIENT = 1 };
pwo_solver *pwo_solver_create_bc(uint32_t nx, uint32_t ny, uint32_t nz, double tau0, int op,
                                 int les, int walls, int sides, int outlet,
                                 const uint32_t *flags, size_t flags_len,
                                 const double *body_force, const double *eq_state,
                                 const double *initial, char *err, size_t err_len);
/* Flat index into a state that fluid cell `cell` pulls direction i from (-1 for other cells). */
int64_t pwo_solver_source(const pwo_solver *s, size_t cell, int i);
void pwo_solver_destroy(pwo_solver *s);
void pwo_solver_step(pwo_solver *s, int count, int threads);
double *pwo_solver_current(pwo_solver *s);  /* 19*N, direction-major */
double *pwo_solver_previous(pwo_solver *s); /* the other ping-pong state */
void pwo_solver_macroscopic(const pwo_solver *s, size_t cell, double out[4]);
void pwo_solver_macro_all(const pwo_solver *s, double *out4n);
void pwo_solver_momentum_exchange(const pwo_solver *s, double out[3]);
size_t pwo_solver_link_count(const pwo_solver *s);

/* ---- voxelize.ts / wall-links.ts / triangle-bins.ts ---- */
typedef struct {
  uint32_t *geometry;      /* 2N + links words */
  size_t geometry_words;
  uint32_t *boundary_cells;
  size_t n_boundary;
  uint64_t solid_cells;
  uint64_t frontal_cells;
  double blockage;
  int watertight;
  uint64_t missed_links;
} pwo_voxel_grid;
int pwo_voxelize(const float *body, size_t body_tris, const float *supports, size_t support_tris,
                 uint32_t nx, uint32_t ny, uint32_t nz, int floor, int link_fractions,
                 pwo_voxel_grid *out, char *err, size_t err_len);
int pwo_build_wall_links(uint32_t *flags, const float *body, size_t body_tris,
                         const float *supports, size_t support_tris, uint32_t nx, uint32_t ny,
                         uint32_t nz, int floor, int measure, pwo_voxel_grid *out);
void pwo_voxel_grid_free(pwo_voxel_grid *g);
/* nearestLinkHit on a freshly binned mesh; returns 0 for a miss, 1 with *t for a hit. */
int pwo_nearest_link_hit(const float *mesh, size_t tris, uint32_t nx, uint32_t ny, uint32_t nz,
                         const double origin[3], const double direction[3], double *t);

/* ---- presets.ts / transform.ts ---- */
/* kind: 0 sphere, 1 ahmed-25 body, 2 cube, 3 ahmed stilts. Returns triangle count; writes
 * 9*count floats into out when out != NULL. */
size_t pwo_preset_mesh(int kind, float *out);
/* placeMesh with flow +x, up z (src/geometry/transform.ts:118-143). supports may be NULL. */
int pwo_place_mesh(const float *mesh, size_t tris, const float *supports, size_t stris,
                   double length_cells, uint32_t nx, uint32_t ny, uint32_t nz, int floor,
                   float *out_body, float *out_supports);

/* ---- particle-layout.ts / particles_update.wgsl ---- */
uint32_t pwo_particle_hash(uint32_t value);
int pwo_particle_needs_reseed(const double p[3], double age, uint32_t nx, uint32_t ny, uint32_t nz,
                              uint32_t flag);
int pwo_seed_region(const double bmin[3], const double bmax[3], uint32_t nx, uint32_t ny,
                    uint32_t nz, double region[4]);
int pwo_particle_seeds(const uint32_t *flags, uint32_t nx, uint32_t ny, uint32_t nz,
                       const double bmin[3], const double bmax[3], uint32_t count, uint32_t seed,
                       float *data, float fallback[3], double region[4], char *err,
                       size_t err_len);
void pwo_particles_update(float *particles, uint32_t count, const float *macro,
                          const uint32_t *flags, uint32_t nx, uint32_t ny, uint32_t nz,
                          uint32_t seed, float steps, const float cross_section[4],
                          const float fallback[2]);

/* ---- lowbias32 stochastic f16 / bf16 storage (src/lbm/shaders/prelude-f16.wgsl:7-22) ---- */
uint16_t pwo_encode_f16(float value_scaled, uint32_t i, uint32_t n);
void pwo_encode_f16_planes(const float *g, size_t cells, uint16_t *out);
float pwo_decode_f16(uint16_t bits);
uint16_t pwo_encode_bf16(float value, uint32_t i, uint32_t n);
float pwo_decode_bf16(uint16_t bits);

#ifdef __cplusplus
}
#endif
#endif
