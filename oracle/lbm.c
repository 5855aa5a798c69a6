/*
 * lbm.c -- FP64 CPU oracle for the D3Q19 pull-stream/collide solver.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product). Restates, in C:
 *   src/lbm/lattice.ts:2-52             velocity set, weights, opposite pairs
 *   src/lbm/reference/collision.ts      equilibrium, moments, TRT/REG/BGK, Smagorinsky,
 *                                       Werner-Wengle wall law and wall-cell collision
 *   src/lbm/reference/cpu-solver.ts     pull streaming through a source table, EQ inlet and
 *                                       pressure outlet, halfway and Bouzidi walls, body
 *                                       force impulse and two-state momentum exchange
 * Built with -ffp-contract=off so every product is rounded as in IEEE double JS.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pw_oracle.h"

const int PWO_C[19][3] = {
    {0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},  {0, 0, -1},
    {1, 1, 0},  {-1, -1, 0}, {1, 0, 1},  {-1, 0, -1}, {0, 1, 1},  {0, -1, -1}, {1, -1, 0},
    {-1, 1, 0}, {1, 0, -1}, {-1, 0, 1}, {0, 1, -1}, {0, -1, 1}};

const double PWO_W[19] = {1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                          1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                          1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                          1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};

/* src/lbm/lattice.ts:47-52: pairs (1,2), (3,4), ... are opposite. */
int pwo_opp(int i) { return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1); }

/* TRT magic parameter and Smagorinsky constant (collision.ts:7-9). */
static const double LAMBDA = 3.0 / 16.0;
static const double CS = 0.1;

typedef struct {
  double xx, yy, zz, xy, xz, yz;
} stress6;

/* collision.ts:21-31: shifted equilibrium g_eq = w[(rho-1) + rho(3cu + 4.5cu^2 - 1.5u^2)]. */
void pwo_equilibrium(double *target, double rho, const double u[3]) {
  const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  for (int i = 0; i < 19; ++i) {
    const double cu = PWO_C[i][0] * u[0] + PWO_C[i][1] * u[1] + PWO_C[i][2] * u[2];
    target[i] = PWO_W[i] * (rho - 1 + rho * (3 * cu + 4.5 * cu * cu - 1.5 * uu));
  }
}

/* collision.ts:34-47: rho = 1 + sum g, u = sum c g / rho. */
void pwo_moments(const double *g, double *rho, double u[3]) {
  double r = 1, mx = 0, my = 0, mz = 0;
  for (int i = 0; i < 19; ++i) {
    r += g[i];
    mx += PWO_C[i][0] * g[i];
    my += PWO_C[i][1] * g[i];
    mz += PWO_C[i][2] * g[i];
  }
  *rho = r;
  u[0] = mx / r;
  u[1] = my / r;
  u[2] = mz / r;
}

/* collision.ts:49-66: Pi_neq = sum c c (g - geq). */
static stress6 neq_stress(const double *g, const double *geq) {
  stress6 s = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 19; ++i) {
    const double d = g[i] - geq[i];
    const double cx = PWO_C[i][0], cy = PWO_C[i][1], cz = PWO_C[i][2];
    s.xx += cx * cx * d;
    s.yy += cy * cy * d;
    s.zz += cz * cz * d;
    s.xy += cx * cy * d;
    s.xz += cx * cz * d;
    s.yz += cy * cz * d;
  }
  return s;
}

/* collision.ts:68-74: closed-form Smagorinsky relaxation time. */
static double smagorinsky_tau(double tau0, double rho, stress6 s) {
  const double q = s.xx * s.xx + s.yy * s.yy + s.zz * s.zz +
                   2 * (s.xy * s.xy + s.xz * s.xz + s.yz * s.yz);
  return 0.5 * (tau0 + sqrt(tau0 * tau0 + (18 * M_SQRT2 * (CS * CS) * sqrt(q)) / rho));
}

/* collision.ts:77-96: regularized post-collision state from a given stress. */
static void regularize(double *g, const double *geq, double rho, double omega, stress6 s,
                       const double *force_terms) {
  const double third = (s.xx + s.yy + s.zz) / 3;
  for (int i = 0; i < 19; ++i) {
    const double cx = PWO_C[i][0], cy = PWO_C[i][1], cz = PWO_C[i][2];
    const double projected = cx * cx * s.xx + cy * cy * s.yy + cz * cz * s.zz +
                             2 * (cx * cy * s.xy + cx * cz * s.xz + cy * cz * s.yz) - third;
    const double impulse = i == 0 ? 0 : rho * force_terms[i];
    g[i] = geq[i] + (1 - omega) * 4.5 * PWO_W[i] * projected + impulse;
  }
}

/* collision.ts:99-119: two-relaxation-time pairs. */
static void relax_pairs(double *g, const double *geq, double rho, double wp, double wm,
                        const double *force_terms) {
  g[0] = g[0] - wp * (g[0] - geq[0]);
  for (int i = 1; i < 19; i += 2) {
    const int j = i + 1;
    const double gi = g[i], gj = g[j];
    const double even = 0.5 * wp * (gi + gj - geq[i] - geq[j]);
    const double odd = 0.5 * wm * (gi - gj - geq[i] + geq[j]);
    g[i] = gi - even - odd + rho * force_terms[i];
    g[j] = gj - even + odd + rho * force_terms[j];
  }
}

/* collision.ts:122-140. */
void pwo_collide(double *g, const double *geq, double rho, double tau0, int op, int les,
                 const double *force_terms) {
  const int need_stress = les || op == PWO_REG;
  stress6 s = {0, 0, 0, 0, 0, 0};
  if (need_stress) s = neq_stress(g, geq);
  const double tau = les ? smagorinsky_tau(tau0, rho, s) : tau0;
  const double wp = 1 / tau;
  if (op == PWO_REG) {
    regularize(g, geq, rho, wp, s, force_terms);
    return;
  }
  const double wm = op == PWO_BGK ? wp : 1 / (LAMBDA / (tau - 0.5) + 0.5);
  relax_pairs(g, geq, rho, wp, wm, force_terms);
}

/* collision.ts:142-154: Werner-Wengle u+ = 8.3 y+^(1/7), viscous below y+ = 11.81, y = 1/2. */
double pwo_friction_velocity(double speed, double viscosity) {
  const double y = 0.5;
  const double viscous = sqrt((viscosity * speed) / y);
  if ((y * viscous) / viscosity <= 11.81) return viscous;
  const double b = 1.0 / 7.0;
  return pow((speed * pow(viscosity / y, b)) / 8.3, 1 / (1 + b));
}

/* collision.ts:162-224: set the post-collision stress so the wall takes the wall-law shear. */
int pwo_collide_wall(double *g, const double *geq, double rho, const double u[3],
                     const int *wall_incoming, int n_incoming, double viscosity,
                     const double *force_terms) {
  double nx = 0, ny = 0, nz = 0, exposed = 0;
  for (int k = 0; k < n_incoming; ++k) {
    const int i = wall_incoming[k];
    nx += PWO_W[i] * PWO_C[i][0];
    ny += PWO_W[i] * PWO_C[i][1];
    nz += PWO_W[i] * PWO_C[i][2];
    exposed += PWO_W[i];
  }
  double sx = u[0], sy = u[1], sz = u[2];
  const double nlen = hypot(hypot(nx, ny), nz);
  if (nlen > 1e-3) {
    const double normal = (sx * nx + sy * ny + sz * nz) / (nlen * nlen);
    sx -= normal * nx;
    sy -= normal * ny;
    sz -= normal * nz;
  }
  const double speed = hypot(hypot(sx, sy), sz);
  if (speed < 1e-7) return 0;
  const double tx = sx / speed, ty = sy / speed, tz = sz / speed;
  const double ut = pwo_friction_velocity(speed, viscosity);
  const double shear = 6 * exposed * rho * ut * ut;
  double flux = 0;
  stress6 b = {0, 0, 0, 0, 0, 0};
  for (int k = 0; k < n_incoming; ++k) {
    const int j = pwo_opp(wall_incoming[k]);
    const double cx = PWO_C[j][0], cy = PWO_C[j][1], cz = PWO_C[j][2];
    const double along = cx * tx + cy * ty + cz * tz;
    flux += 2 * along * (geq[j] + PWO_W[j]);
    const double s = 9 * PWO_W[j] * along;
    b.xx += s * (cx * cx - 1.0 / 3.0);
    b.yy += s * (cy * cy - 1.0 / 3.0);
    b.zz += s * (cz * cz - 1.0 / 3.0);
    b.xy += s * cx * cy;
    b.xz += s * cx * cz;
    b.yz += s * cy * cz;
  }
  const double norm = b.xx * b.xx + b.yy * b.yy + b.zz * b.zz +
                      2 * (b.xy * b.xy + b.xz * b.xz + b.yz * b.yz);
  if (norm < 0.05) return 0;
  double post[19];
  regularize(post, geq, rho, 1 - (shear - flux) / norm, b, force_terms);
  for (int i = 0; i < 19; ++i)
    if (post[i] + PWO_W[i] < 0) return 0;
  memcpy(g, post, sizeof post);
  return 1;
}

/* ------------------------------------------------------------------------------------ */
/* CpuSolver (cpu-solver.ts:69-323)                                                      */

typedef struct {
  size_t cell;
  int incoming;
  int type;
  double q;
  long long second; /* -1: no second node (halfway fallback) */
} wall_link;

struct pwo_solver {
  uint32_t nx, ny, nz;
  size_t cells;
  double tau0;
  int op, les, walls;
  int sides, outlet;
  uint8_t *types;
  int64_t *sources; /* 19*N: flat index into a population state */
  wall_link *links;
  size_t n_links;
  size_t *first_link; /* N+1 */
  double eq_cell[19];
  double force_terms[19];
  double *cur, *nxt;
};

static void fail(char *err, size_t n, const char *msg) {
  if (err && n) snprintf(err, n, "%s", msg);
}

static void coords(const pwo_solver *s, size_t cell, long *x, long *y, long *z) {
  *x = (long)(cell % s->nx);
  *y = (long)((cell / s->nx) % s->ny);
  *z = (long)(cell / ((size_t)s->nx * s->ny));
}

/* cpu-solver.ts:179-183: every axis wraps. With a zero-gradient outlet (not in the reference)
 * x = nx is held at nx - 1; with free-slip sides (not in the reference) y and z outside the
 * grid are held at the cell's own coordinate, and *mirror_y / *mirror_z (when given) report
 * that the pull's direction takes its specular image in that axis. */
static long along(long c, int d, long n, int hold, int *held) {
  const long v = c + d;
  if (v >= 0 && v < n) return v;
  if (held) *held = hold;
  return hold ? c : (v + n) % n;
}
static size_t neighbour(const pwo_solver *s, size_t cell, int cx, int cy, int cz, int *mirror_y,
                        int *mirror_z) {
  long x, y, z;
  coords(s, cell, &x, &y, &z);
  const long nx = s->nx, ny = s->ny, nz = s->nz;
  const long sx = (s->outlet == PWO_OUTLET_ZERO_GRADIENT && x + cx == nx) ? x : (x + cx + nx) % nx;
  const int slip = s->sides == PWO_SIDES_FREE_SLIP;
  const long sy = along(y, cy, ny, slip, mirror_y), sz = along(z, cz, nz, slip, mirror_z);
  return (size_t)(sx + nx * (sy + ny * sz));
}
static size_t wrapped(const pwo_solver *s, size_t cell, int cx, int cy, int cz) {
  return neighbour(s, cell, cx, cy, cz, NULL, NULL);
}

/* Index of the velocity (x, y, z). */
static int velocity_index(int x, int y, int z) {
  for (int i = 0; i < 19; ++i)
    if (PWO_C[i][0] == x && PWO_C[i][1] == y && PWO_C[i][2] == z) return i;
  return -1;
}

static int is_wall_type(int t) { return t >= PWO_SOLID && t <= PWO_FLOOR; }

static int valid_state(const double *st) {
  return isfinite(st[0]) && st[0] > 0 && isfinite(st[1]) && isfinite(st[2]) && isfinite(st[3]);
}

pwo_solver *pwo_solver_create(uint32_t nx, uint32_t ny, uint32_t nz, double tau0, int op,
                              int les, int walls, const uint32_t *flags, size_t flags_len,
                              const double *body_force, const double *eq_state,
                              const double *initial, char *err, size_t err_len) {
  return pwo_solver_create_bc(nx, ny, nz, tau0, op, les, walls, PWO_SIDES_PERIODIC,
                              PWO_OUTLET_EQUILIBRIUM, flags, flags_len, body_force, eq_state,
                              initial, err, err_len);
}

pwo_solver *pwo_solver_create_bc(uint32_t nx, uint32_t ny, uint32_t nz, double tau0, int op,
                                 int les, int walls, int sides, int outlet,
                                 const uint32_t *flags, size_t flags_len,
                                 const double *body_force, const double *eq_state,
                                 const double *initial, char *err, size_t err_len) {
  if (sides < PWO_SIDES_PERIODIC || sides > PWO_SIDES_FREE_SLIP || outlet < PWO_OUTLET_EQUILIBRIUM ||
      outlet > PWO_OUTLET_ZERO_GRADIENT || (outlet == PWO_OUTLET_ZERO_GRADIENT && nx < 2)) {
    fail(err, err_len, "Unknown side or outlet boundary option.");
    return NULL;
  }
  if (nx < 1 || ny < 1 || nz < 1) {
    fail(err, err_len, "Grid dimensions must be positive safe integers.");
    return NULL;
  }
  if (!isfinite(tau0) || tau0 <= 0.5) {
    fail(err, err_len, "Relaxation time must be finite and greater than 0.5.");
    return NULL;
  }
  const size_t cells = (size_t)nx * ny * nz;
  if (cells * 19 > 0x7fffffffULL) {
    fail(err, err_len, "Reference grid exceeds signed population indexing.");
    return NULL;
  }
  if (flags && flags_len != cells && flags_len < 2 * cells) {
    fail(err, err_len, "Flags must contain one entry per cell or the packed geometry.");
    return NULL;
  }
  if (walls == PWO_BOUZIDI && (!flags || flags_len < 2 * cells)) {
    fail(err, err_len, "Interpolated walls need the packed geometry with q offsets.");
    return NULL;
  }
  static const double rest[4] = {1, 0, 0, 0};
  const double *eqs = eq_state ? eq_state : rest;
  if (!valid_state(eqs)) {
    fail(err, err_len, "Initial density and velocity must be finite, density positive.");
    return NULL;
  }
  const double zero3[3] = {0, 0, 0};
  const double *a = body_force ? body_force : zero3;
  if (!isfinite(a[0]) || !isfinite(a[1]) || !isfinite(a[2])) {
    fail(err, err_len, "Body acceleration must be finite.");
    return NULL;
  }
  pwo_solver *s = calloc(1, sizeof *s);
  s->nx = nx;
  s->ny = ny;
  s->nz = nz;
  s->cells = cells;
  s->tau0 = tau0;
  s->op = op;
  s->les = les;
  s->walls = walls;
  s->sides = sides;
  s->outlet = outlet;
  s->types = calloc(cells, 1);
  s->sources = malloc(sizeof(int64_t) * cells * 19);
  s->first_link = calloc(cells + 1, sizeof(size_t));
  s->cur = calloc(cells * 19, sizeof(double));
  s->nxt = calloc(cells * 19, sizeof(double));
  pwo_equilibrium(s->eq_cell, eqs[0], eqs + 1);
  for (int i = 0; i < 19; ++i)
    s->force_terms[i] =
        3 * PWO_W[i] * (PWO_C[i][0] * a[0] + PWO_C[i][1] * a[1] + PWO_C[i][2] * a[2]);

  /* cpu-solver.ts:114-132: walls stay zero (rest); EQ cells hold the inflow equilibrium. */
  double geq[19];
  for (size_t cell = 0; cell < cells; ++cell) {
    const uint32_t t = (flags ? flags[cell] : 0u) & PWO_TYPE_MASK;
    if (t > PWO_EQ) {
      fail(err, err_len, "Unknown cell type.");
      pwo_solver_destroy(s);
      return NULL;
    }
    /* The zero-gradient outlet steps the EQ cells of x = nx - 1 as fluid cells. */
    const int opened = outlet == PWO_OUTLET_ZERO_GRADIENT && t == PWO_EQ && cell % nx == nx - 1;
    s->types[cell] = (uint8_t)(opened ? PWO_FLUID : t);
    const double *st = initial ? initial + 4 * cell : eqs;
    if (!valid_state(st)) {
      fail(err, err_len, "Initial density and velocity must be finite, density positive.");
      pwo_solver_destroy(s);
      return NULL;
    }
    if (is_wall_type((int)t)) continue;
    pwo_equilibrium(geq, st[0], st + 1);
    for (int i = 0; i < 19; ++i)
      s->cur[(size_t)i * cells + cell] = s->types[cell] == PWO_EQ ? s->eq_cell[i] : geq[i];
  }
  memcpy(s->nxt, s->cur, sizeof(double) * cells * 19);

  /* cpu-solver.ts:144-177: source table and per-cell wall links in cell order. */
  size_t cap = 1024;
  s->links = malloc(cap * sizeof(wall_link));
  for (size_t cell = 0; cell < cells; ++cell) {
    s->first_link[cell] = s->n_links;
    if (s->types[cell] != PWO_FLUID) continue;
    int rank = 0;
    for (int i = 0; i < 19; ++i) {
      int my = 0, mz = 0;
      const size_t nb = neighbour(s, cell, -PWO_C[i][0], -PWO_C[i][1], -PWO_C[i][2], &my, &mz);
      const int plane = velocity_index(PWO_C[i][0], my ? -PWO_C[i][1] : PWO_C[i][1],
                                       mz ? -PWO_C[i][2] : PWO_C[i][2]);
      const int t = s->types[nb];
      const int wall = is_wall_type(t);
      s->sources[(size_t)i * cells + cell] =
          wall ? (int64_t)pwo_opp(i) * (int64_t)cells + (int64_t)cell
               : (int64_t)plane * (int64_t)cells + (int64_t)nb;
      if (!wall) continue;
      double q = 0.5;
      if (walls == PWO_BOUZIDI) {
        const uint32_t base = flags[cells + cell];
        if (base != PWO_NO_Q) {
          const size_t index = 2 * cells + base + (size_t)rank;
          if (index >= flags_len) {
            char msg[128];
            snprintf(msg, sizeof msg, "Wall-link fraction is missing at cell %zu.", cell);
            fail(err, err_len, msg);
            pwo_solver_destroy(s);
            return NULL;
          }
          float qf;
          memcpy(&qf, flags + index, 4);
          q = qf;
          if (!(q > 0 && q <= 1)) {
            char msg[128];
            snprintf(msg, sizeof msg, "Wall-link fraction at cell %zu must be in (0, 1].", cell);
            fail(err, err_len, msg);
            pwo_solver_destroy(s);
            return NULL;
          }
        }
      }
      const size_t second = wrapped(s, cell, PWO_C[i][0], PWO_C[i][1], PWO_C[i][2]);
      const int short_link = q < 0.5 && s->types[second] == PWO_FLUID;
      if (s->n_links == cap) {
        cap *= 2;
        s->links = realloc(s->links, cap * sizeof(wall_link));
      }
      s->links[s->n_links++] = (wall_link){cell, i, t, q, short_link ? (long long)second : -1};
      rank += 1;
    }
  }
  s->first_link[cells] = s->n_links;
  return s;
}

void pwo_solver_destroy(pwo_solver *s) {
  if (!s) return;
  free(s->types);
  free(s->sources);
  free(s->links);
  free(s->first_link);
  free(s->cur);
  free(s->nxt);
  free(s);
}

double *pwo_solver_current(pwo_solver *s) { return s->cur; }
double *pwo_solver_previous(pwo_solver *s) { return s->nxt; }
size_t pwo_solver_link_count(const pwo_solver *s) { return s->n_links; }
int64_t pwo_solver_source(const pwo_solver *s, size_t cell, int i) {
  return s->types[cell] == PWO_FLUID ? s->sources[(size_t)i * s->cells + cell] : -1;
}

/* cpu-solver.ts:200-208: linear Bouzidi from a post-collision state. */
static double wall_value(const pwo_solver *s, const wall_link *l, const double *state) {
  const size_t N = s->cells;
  const double reflected = state[(size_t)pwo_opp(l->incoming) * N + l->cell];
  if (l->q == 0.5) return reflected;
  if (l->q < 0.5) {
    if (l->second < 0) return reflected;
    return 2 * l->q * reflected +
           (1 - 2 * l->q) * state[(size_t)pwo_opp(l->incoming) * N + (size_t)l->second];
  }
  return (reflected + (2 * l->q - 1) * state[(size_t)l->incoming * N + l->cell]) / (2 * l->q);
}

static void macro_of(const pwo_solver *s, const double *state, size_t cell, double out[4]) {
  double g[19];
  for (int i = 0; i < 19; ++i) g[i] = state[(size_t)i * s->cells + cell];
  pwo_moments(g, out, out + 1);
}

void pwo_solver_macroscopic(const pwo_solver *s, size_t cell, double out[4]) {
  macro_of(s, s->cur, cell, out);
}

void pwo_solver_macro_all(const pwo_solver *s, double *out4n) {
  for (size_t c = 0; c < s->cells; ++c) macro_of(s, s->cur, c, out4n + 4 * c);
}

/* cpu-solver.ts:246-282 for one fluid cell. */
static void advance_fluid(const pwo_solver *s, size_t cell) {
  const size_t N = s->cells;
  double g[19], geq[19], rho, u[3];
  for (int i = 0; i < 19; ++i) g[i] = s->cur[s->sources[(size_t)i * N + cell]];
  const size_t first = s->first_link[cell], end = s->first_link[cell + 1];
  if (s->walls == PWO_BOUZIDI)
    for (size_t k = first; k < end; ++k) g[s->links[k].incoming] = wall_value(s, &s->links[k], s->cur);
  pwo_moments(g, &rho, u);
  pwo_equilibrium(geq, rho, u);
  int walled = 0;
  if (s->les && end > first) {
    int incoming[18];
    int n = 0;
    for (size_t k = first; k < end; ++k) incoming[n++] = s->links[k].incoming;
    walled = pwo_collide_wall(g, geq, rho, u, incoming, n, (s->tau0 - 0.5) / 3, s->force_terms);
  }
  if (!walled) pwo_collide(g, geq, rho, s->tau0, s->op, s->les, s->force_terms);
  for (int i = 0; i < 19; ++i) s->nxt[(size_t)i * N + cell] = g[i];
}

/* cpu-solver.ts:222-244. */
static void advance_cell(const pwo_solver *s, size_t cell) {
  const size_t N = s->cells;
  const int t = s->types[cell];
  if (t == PWO_EQ) {
    double geq[19];
    const double *src = s->eq_cell;
    if (s->nx >= 2 && cell % s->nx == s->nx - 1) {
      double m[4];
      macro_of(s, s->cur, cell - 1, m);
      pwo_equilibrium(geq, 1.0, m + 1);
      src = geq;
    }
    for (int i = 0; i < 19; ++i) s->nxt[(size_t)i * N + cell] = src[i];
  } else if (t == PWO_FLUID) {
    advance_fluid(s, cell);
  }
}

void pwo_solver_step(pwo_solver *s, int count, int threads) {
  for (int k = 0; k < count; ++k) {
    const long long N = (long long)s->cells;
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1) if (threads > 1)
    for (long long cell = 0; cell < N; ++cell) advance_cell(s, (size_t)cell);
    double *t = s->cur;
    s->cur = s->nxt;
    s->nxt = t;
  }
}

/* cpu-solver.ts:296-322: mean over the two latest states of sum c_j (g*_j + wall value). */
static void exchange(const pwo_solver *s, const double *state, double f[3]) {
  const size_t N = s->cells;
  f[0] = f[1] = f[2] = 0;
  for (size_t k = 0; k < s->n_links; ++k) {
    const wall_link *l = &s->links[k];
    if (l->type != PWO_SOLID) continue;
    const int j = pwo_opp(l->incoming);
    const double out = state[(size_t)j * N + l->cell];
    const double ex = out + (s->walls == PWO_BOUZIDI ? wall_value(s, l, state) : out);
    f[0] += PWO_C[j][0] * ex;
    f[1] += PWO_C[j][1] * ex;
    f[2] += PWO_C[j][2] * ex;
  }
}

void pwo_solver_momentum_exchange(const pwo_solver *s, double out[3]) {
  double a[3], b[3];
  exchange(s, s->cur, a);
  exchange(s, s->nxt, b);
  for (int k = 0; k < 3; ++k) out[k] = 0.5 * (a[k] + b[k]);
}

/* ------------------------------------------------------------------------------------ */
/* Population storage encodings.                                                          */

static uint32_t f2u(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
static float u2f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}

/* lowbias32 mix of (value bits, cell, direction) (prelude-f16.wgsl:8-13). */
static uint32_t storage_hash(uint32_t bits, uint32_t i, uint32_t n) {
  uint32_t r = bits ^ (n * 0x9e3779b9u) ^ (i * 0x85ebca6bu);
  r = (r ^ (r >> 16)) * 0x7feb352du;
  r = (r ^ (r >> 15)) * 0x846ca68bu;
  r ^= r >> 16;
  return r;
}

/* IEEE binary16 from a float, round to nearest even (the native cast). */
static uint16_t f32_to_f16_rne(float f) {
  const uint32_t x = f2u(f);
  const uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t mag = x & 0x7fffffffu;
  if (mag >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (mag > 0x7f800000u ? 0x200u : 0u));
  if (mag >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* rounds past 65504 */
  if (mag < 0x38800000u) {                                   /* f16 subnormal or zero */
    if (mag < 0x33000000u) return (uint16_t)sign;            /* below half the smallest subnormal */
    const uint32_t e = mag >> 23;
    const uint32_t m = (mag & 0x7fffffu) | 0x800000u;
    const uint32_t shift = 126 - e; /* value / 2^-24 = m * 2^(e - 126) */
    uint32_t r = m >> (shift);
    const uint32_t rem = m & ((1u << shift) - 1);
    const uint32_t half = 1u << (shift - 1);
    if (rem > half || (rem == half && (r & 1))) r += 1;
    return (uint16_t)(sign | r);
  }
  uint32_t r = ((mag - 0x38000000u) >> 13);
  const uint32_t rem = mag & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (r & 1))) r += 1;
  return (uint16_t)(sign | r);
}

float pwo_decode_f16(uint16_t h) {
  const int s = (h >> 15) & 1, e = (h >> 10) & 0x1f, m = h & 0x3ff;
  double v;
  if (e == 0)
    v = ldexp((double)m, -24);
  else if (e == 31)
    v = m ? NAN : INFINITY;
  else
    v = ldexp((double)(m | 0x400), e - 25);
  return (float)(s ? -v : v);
}

/* prelude-f16.wgsl:7-22: proportional stochastic rounding of the 13 dropped bits. */
uint16_t pwo_encode_f16(float value, uint32_t i, uint32_t n) {
  const uint32_t bits = f2u(value);
  const uint32_t r = storage_hash(bits, i, n);
  const uint32_t mag = bits & 0x7fffffffu;
  if (mag < 0x38800000u) return f32_to_f16_rne(value);
  const uint32_t inc = (r & 0x1fffu) < (mag & 0x1fffu) ? 0x2000u : 0u;
  const uint32_t rounded = (mag & 0xffffe000u) + inc;
  return f32_to_f16_rne(u2f((bits & 0x80000000u) | rounded));
}

/* common.wgsl:134-135 over a whole state: the planes of 19 x cells deviations g (direction
 * major, the reference's cell order), each stored as encode_population(g * POP_SCALE, i, n). */
void pwo_encode_f16_planes(const float *g, size_t cells, uint16_t *out) {
  for (uint32_t i = 0; i < 19; ++i)
    for (size_t n = 0; n < cells; ++n) out[i * cells + n] = pwo_encode_f16(g[i * cells + n] * 32768.0f, i, (uint32_t)n);
}

/* bfloat16 deviation storage: the same hash decides the 16 dropped bits (build-only format). */
uint16_t pwo_encode_bf16(float value, uint32_t i, uint32_t n) {
  const uint32_t bits = f2u(value);
  const uint32_t mag = bits & 0x7fffffffu;
  if (mag >= 0x7f800000u) return (uint16_t)((bits >> 16) | (mag > 0x7f800000u ? 0x40u : 0u));
  const uint32_t r = storage_hash(bits, i, n);
  const uint32_t inc = (r & 0xffffu) < (mag & 0xffffu) ? 0x10000u : 0u;
  const uint32_t rounded = (mag & 0xffff0000u) + inc; /* may carry into the exponent */
  return (uint16_t)(((bits & 0x80000000u) | rounded) >> 16);
}

float pwo_decode_bf16(uint16_t b) { return u2f((uint32_t)b << 16); }
