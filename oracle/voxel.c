/*
 * voxel.c -- CPU oracle for the triangle-mesh voxelizer, wall links and preset bodies.
 *
 * TEST INFRASTRUCTURE ONLY (the checker). Restates, in IEEE double with no contraction:
 *   src/geometry/voxelize.ts:23-199      three-projection parity vote, EQ/floor planes, counts
 *   src/geometry/wall-links.ts:36-131    pull wall masks, q offsets, boundary cells, q fractions
 *   src/geometry/triangle-bins.ts:14-162 4-cell triangle bins and Moller-Trumbore link hits
 *   src/geometry/presets.ts:4-189        sphere, Ahmed body, stilts and cube meshes
 *   src/geometry/transform.ts:118-143    placement in lattice cells (flow +x, up z)
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pw_oracle.h"

static void fail(char *err, size_t n, const char *msg) {
  if (err && n) snprintf(err, n, "%s", msg);
}

/* voxelize.ts:38-45 */
static double edge_fn(const double *a, const double *b, double u, double v) {
  return (b[0] - a[0]) * (v - a[1]) - (b[1] - a[1]) * (u - a[0]);
}
static int owns_edge(double value, const double *a, const double *b) {
  return value > 0 || (value == 0 && (b[1] < a[1] || (b[1] == a[1] && b[0] > a[0])));
}

typedef struct {
  size_t *count;  /* hits per column */
  size_t *offset; /* prefix sums */
  double *hits;
} columns_t;

/* voxelize.ts:47-75; mode 0 counts hits per column, mode 1 stores them. */
static void rasterize(const double *a, const double *second, const double *third, long nu, long nv,
                      columns_t *cols, size_t *cursor, int mode) {
  const double signed_area = edge_fn(a, second, third[0], third[1]);
  if (signed_area == 0) return;
  const double *b = signed_area > 0 ? second : third;
  const double *c = signed_area > 0 ? third : second;
  const double area = fabs(signed_area);
  const double umin = fmax(0, ceil(fmin(fmin(a[0], b[0]), c[0])));
  const double umax = fmin((double)(nu - 1), floor(fmax(fmax(a[0], b[0]), c[0])));
  const double vmin = fmax(0, ceil(fmin(fmin(a[1], b[1]), c[1])));
  const double vmax = fmin((double)(nv - 1), floor(fmax(fmax(a[1], b[1]), c[1])));
  if (umax < umin || vmax < vmin) return;
  for (long v = (long)vmin; v <= (long)vmax; ++v) {
    for (long u = (long)umin; u <= (long)umax; ++u) {
      const double wa = edge_fn(b, c, (double)u, (double)v);
      const double wb = edge_fn(c, a, (double)u, (double)v);
      const double wc = edge_fn(a, b, (double)u, (double)v);
      if (owns_edge(wa, b, c) && owns_edge(wb, c, a) && owns_edge(wc, a, b)) {
        const size_t col = (size_t)(u + nu * v);
        if (mode == 0)
          cols->count[col] += 1;
        else
          cols->hits[cols->offset[col] + cursor[col]++] = (wa * a[2] + wb * b[2] + wc * c[2]) / area;
      }
    }
  }
}

static int cmp_double(const void *x, const void *y) {
  const double a = *(const double *)x, b = *(const double *)y;
  return (a > b) - (a < b);
}

/* voxelize.ts:103-136: one mesh votes; cells with >= 2 votes that are still FLUID take `type`. */
static int vote_mesh(const float *pos, size_t tris, const uint32_t size[3], uint32_t *flags,
                     uint32_t type) {
  const size_t cells = (size_t)size[0] * size[1] * size[2];
  const size_t strides[3] = {1, size[0], (size_t)size[0] * size[1]};
  static const int proj[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};
  uint8_t *votes = calloc(cells, 1);
  int watertight = 1;
  for (int p = 0; p < 3; ++p) {
    const int U = proj[p][0], V = proj[p][1], A = proj[p][2];
    const long nu = size[U], nv = size[V];
    const size_t ncol = (size_t)nu * (size_t)nv;
    columns_t cols;
    cols.count = calloc(ncol, sizeof(size_t));
    cols.offset = calloc(ncol + 1, sizeof(size_t));
    size_t *cursor = calloc(ncol, sizeof(size_t));
    for (int mode = 0; mode < 2; ++mode) {
      if (mode == 1) {
        for (size_t k = 0; k < ncol; ++k) cols.offset[k + 1] = cols.offset[k] + cols.count[k];
        cols.hits = malloc(sizeof(double) * (cols.offset[ncol] + 1));
      }
      for (size_t t = 0; t < tris; ++t) {
        const float *q = pos + 9 * t;
        const double a[3] = {q[U], q[V], q[A]};
        const double b[3] = {q[3 + U], q[3 + V], q[3 + A]};
        const double c[3] = {q[6 + U], q[6 + V], q[6 + A]};
        rasterize(a, b, c, nu, nv, &cols, cursor, mode);
      }
    }
    /* voxelize.ts:77-101 */
    const long depth = size[A];
    for (size_t col = 0; col < ncol; ++col) {
      const size_t n = cols.count[col];
      if (n == 0) continue;
      double *h = cols.hits + cols.offset[col];
      qsort(h, n, sizeof(double), cmp_double);
      if (n % 2) watertight = 0;
      const size_t base = (col % (size_t)nu) * strides[U] + (col / (size_t)nu) * strides[V];
      for (size_t k = 0; k + 1 < n; k += 2) {
        const double start = fmax(0, ceil(h[k]));
        const double end = fmin((double)depth, ceil(h[k + 1]));
        if (!(end > start)) continue;
        for (long z = (long)start; z < (long)end; ++z) votes[base + (size_t)z * strides[A]] += 1;
      }
    }
    free(cols.count);
    free(cols.offset);
    free(cols.hits);
    free(cursor);
  }
  for (size_t c = 0; c < cells; ++c)
    if (votes[c] >= 2 && flags[c] == PWO_FLUID) flags[c] = type;
  free(votes);
  return watertight;
}

/* ---- triangle bins (triangle-bins.ts) ---- */
typedef struct {
  const float *pos;
  float *bounds; /* 6 per triangle */
  size_t *start; /* per bin, CSR */
  size_t *items; /* triangle indices */
  long nbx, nby, nbz;
} bins_t;

static void bin_triangles(bins_t *b, const float *pos, size_t tris, uint32_t nx, uint32_t ny,
                          uint32_t nz) {
  b->pos = pos;
  b->nbx = (nx + 3) / 4;
  b->nby = (ny + 3) / 4;
  b->nbz = (nz + 3) / 4;
  const size_t nbins = (size_t)b->nbx * b->nby * b->nbz;
  b->bounds = malloc(sizeof(float) * 6 * (tris + 1));
  size_t *count = calloc(nbins + 1, sizeof(size_t));
  long *range = malloc(sizeof(long) * 6 * (tris + 1));
  for (size_t t = 0; t < tris; ++t) {
    float lo[3], hi[3];
    for (int ax = 0; ax < 3; ++ax) {
      const float p = pos[9 * t + ax], q = pos[9 * t + 3 + ax], r = pos[9 * t + 6 + ax];
      lo[ax] = fminf(fminf(p, q), r);
      hi[ax] = fmaxf(fmaxf(p, q), r);
      b->bounds[6 * t + ax] = lo[ax];
      b->bounds[6 * t + 3 + ax] = hi[ax];
    }
    const long nb[3] = {b->nbx, b->nby, b->nbz};
    for (int ax = 0; ax < 3; ++ax) {
      long mn = (long)floor((double)lo[ax] / 4), mx = (long)floor((double)hi[ax] / 4);
      if (mn < 0) mn = 0;
      if (mx > nb[ax] - 1) mx = nb[ax] - 1;
      range[6 * t + ax] = mn;
      range[6 * t + 3 + ax] = mx;
    }
    for (long z = range[6 * t + 2]; z <= range[6 * t + 5]; ++z)
      for (long y = range[6 * t + 1]; y <= range[6 * t + 4]; ++y)
        for (long x = range[6 * t + 0]; x <= range[6 * t + 3]; ++x) count[x + b->nbx * (y + b->nby * z)]++;
  }
  b->start = calloc(nbins + 1, sizeof(size_t));
  for (size_t k = 0; k < nbins; ++k) b->start[k + 1] = b->start[k] + count[k];
  b->items = malloc(sizeof(size_t) * (b->start[nbins] + 1));
  memset(count, 0, sizeof(size_t) * nbins);
  for (size_t t = 0; t < tris; ++t)
    for (long z = range[6 * t + 2]; z <= range[6 * t + 5]; ++z)
      for (long y = range[6 * t + 1]; y <= range[6 * t + 4]; ++y)
        for (long x = range[6 * t + 0]; x <= range[6 * t + 3]; ++x) {
          const size_t k = (size_t)(x + b->nbx * (y + b->nby * z));
          b->items[b->start[k] + count[k]++] = t;
        }
  free(count);
  free(range);
}

static void free_bins(bins_t *b) {
  free(b->bounds);
  free(b->start);
  free(b->items);
}

/* triangle-bins.ts:60-94: Moller-Trumbore; t is the fraction of the unnormalised link. */
static int intersect(const float *p, const double o[3], const double d[3], double *t_out) {
  const double ax = p[0], ay = p[1], az = p[2];
  const double e1x = (double)p[3] - ax, e1y = (double)p[4] - ay, e1z = (double)p[5] - az;
  const double e2x = (double)p[6] - ax, e2y = (double)p[7] - ay, e2z = (double)p[8] - az;
  const double px = d[1] * e2z - d[2] * e2y;
  const double py = d[2] * e2x - d[0] * e2z;
  const double pz = d[0] * e2y - d[1] * e2x;
  const double det = e1x * px + e1y * py + e1z * pz;
  if (fabs(det) < DBL_EPSILON) return 0;
  const double tx = o[0] - ax, ty = o[1] - ay, tz = o[2] - az;
  const double u = (tx * px + ty * py + tz * pz) / det;
  if (u < -1e-9 || u > 1 + 1e-9) return 0;
  const double qx = ty * e1z - tz * e1y;
  const double qy = tz * e1x - tx * e1z;
  const double qz = tx * e1y - ty * e1x;
  const double v = (d[0] * qx + d[1] * qy + d[2] * qz) / det;
  if (v < -1e-9 || u + v > 1 + 1e-9) return 0;
  const double t = (e2x * qx + e2y * qy + e2z * qz) / det;
  if (!(t >= -1e-9 && t <= 1 + 1e-9)) return 0;
  *t_out = t;
  return 1;
}

/* triangle-bins.ts:98-162 */
static int nearest_hit(const bins_t *b, const double o[3], const double d[3], double *t_out) {
  double lo[3], hi[3];
  long bmin[3], bmax[3];
  const long nb[3] = {b->nbx, b->nby, b->nbz};
  for (int ax = 0; ax < 3; ++ax) {
    lo[ax] = fmin(o[ax], o[ax] + d[ax]);
    hi[ax] = fmax(o[ax], o[ax] + d[ax]);
    bmin[ax] = (long)floor(lo[ax] / 4);
    bmax[ax] = (long)floor(hi[ax] / 4);
    if (bmin[ax] < 0) bmin[ax] = 0;
    if (bmax[ax] > nb[ax] - 1) bmax[ax] = nb[ax] - 1;
  }
  double nearest = INFINITY;
  for (long z = bmin[2]; z <= bmax[2]; ++z)
    for (long y = bmin[1]; y <= bmax[1]; ++y)
      for (long x = bmin[0]; x <= bmax[0]; ++x) {
        const size_t k = (size_t)(x + b->nbx * (y + b->nby * z));
        for (size_t e = b->start[k]; e < b->start[k + 1]; ++e) {
          const size_t t = b->items[e];
          const float *bd = b->bounds + 6 * t;
          if (!(bd[0] <= hi[0] && bd[3] >= lo[0] && bd[1] <= hi[1] && bd[4] >= lo[1] &&
                bd[2] <= hi[2] && bd[5] >= lo[2]))
            continue;
          double tt;
          if (intersect(b->pos + 9 * t, o, d, &tt) && tt < nearest) nearest = tt;
        }
      }
  if (nearest == INFINITY) return 0;
  *t_out = nearest;
  return 1;
}

int pwo_nearest_link_hit(const float *mesh, size_t tris, uint32_t nx, uint32_t ny, uint32_t nz,
                         const double origin[3], const double direction[3], double *t) {
  bins_t b;
  bin_triangles(&b, mesh, tris, nx, ny, nz);
  const int hit = nearest_hit(&b, origin, direction, t);
  free_bins(&b);
  return hit;
}

static int is_wall_type(uint32_t t) { return t >= PWO_SOLID && t <= PWO_FLOOR; }

/* wall-links.ts:36-40: x never wraps, z wraps only without a floor, y always wraps. */
static long neighbour_cell(long x, long y, long z, uint32_t nx, uint32_t ny, uint32_t nz, int floor) {
  if (x < 0 || x >= (long)nx || (floor && (z < 0 || z >= (long)nz))) return -1;
  return x + (long)nx * (((y + (long)ny) % (long)ny) + (long)ny * ((z + (long)nz) % (long)nz));
}

/* wall-links.ts:99-131 */
int pwo_build_wall_links(uint32_t *flags, const float *body, size_t body_tris,
                         const float *supports, size_t support_tris, uint32_t nx, uint32_t ny,
                         uint32_t nz, int floor, int measure, pwo_voxel_grid *out) {
  const size_t cells = (size_t)nx * ny * nz;
  uint8_t *body_boundary = calloc(cells, 1);
  /* markWalls (wall-links.ts:46-63) */
  for (size_t w = 0; w < cells; ++w) {
    const uint32_t t = flags[w] & PWO_TYPE_MASK;
    if (!is_wall_type(t)) continue;
    const long x = (long)(w % nx), y = (long)((w / nx) % ny), z = (long)(w / ((size_t)nx * ny));
    for (int i = 0; i < 19; ++i) {
      const long f = neighbour_cell(x + PWO_C[i][0], y + PWO_C[i][1], z + PWO_C[i][2], nx, ny, nz, floor);
      if (f < 0 || (flags[f] & PWO_TYPE_MASK) != PWO_FLUID) continue;
      flags[f] |= 1u << (3 + i);
      if (t == PWO_SOLID) body_boundary[f] = 1;
    }
  }
  bins_t bb, sb;
  const int have_body_bins = measure;
  const int have_support_bins = supports != NULL;
  if (have_body_bins) bin_triangles(&bb, body, body_tris, nx, ny, nz);
  if (have_support_bins) bin_triangles(&sb, supports, support_tris, nx, ny, nz);
  uint32_t *bases = malloc(sizeof(uint32_t) * cells);
  size_t cap = 1024, nfrac = 0, nboundary = 0, bcap = 1024;
  float *fractions = malloc(sizeof(float) * cap);
  uint32_t *boundary = malloc(sizeof(uint32_t) * bcap);
  uint64_t missed = 0;
  for (size_t cell = 0; cell < cells; ++cell) {
    const uint32_t flag = flags[cell];
    bases[cell] = PWO_NO_Q;
    if ((flag >> 3) == 0) continue;
    bases[cell] = (uint32_t)nfrac;
    if (body_boundary[cell]) {
      if (nboundary == bcap) boundary = realloc(boundary, sizeof(uint32_t) * (bcap *= 2));
      boundary[nboundary++] = (uint32_t)cell;
    }
    const long x = (long)(cell % nx), y = (long)((cell / nx) % ny), z = (long)(cell / ((size_t)nx * ny));
    for (int i = 0; i < 19; ++i) {
      if (!(flag & (1u << (3 + i)))) continue;
      /* wall-links.ts:65-92 */
      double q = 0.5;
      int hit = 1;
      if (have_body_bins) {
        const long w = neighbour_cell(x - PWO_C[i][0], y - PWO_C[i][1], z - PWO_C[i][2], nx, ny, nz, floor);
        const uint32_t wt = flags[w] & PWO_TYPE_MASK;
        if (wt != PWO_FLOOR) {
          const bins_t *bins = wt == PWO_SUPPORT ? &sb : &bb;
          const double o[3] = {(double)(w % nx) + PWO_C[i][0], (double)((w / nx) % ny) + PWO_C[i][1],
                               (double)(w / ((long)nx * ny)) + PWO_C[i][2]};
          const double d[3] = {-PWO_C[i][0], -PWO_C[i][1], -PWO_C[i][2]};
          double t;
          hit = nearest_hit(bins, o, d, &t);
          if (hit) q = t;
        }
      }
      missed += hit ? 0 : 1;
      if (!hit) q = 0.5;
      if (nfrac == cap) fractions = realloc(fractions, sizeof(float) * (cap *= 2));
      fractions[nfrac++] = (float)fmax(1e-3, fmin(1, q));
    }
  }
  out->geometry_words = 2 * cells + nfrac;
  out->geometry = malloc(sizeof(uint32_t) * out->geometry_words);
  memcpy(out->geometry, flags, sizeof(uint32_t) * cells);
  memcpy(out->geometry + cells, bases, sizeof(uint32_t) * cells);
  memcpy(out->geometry + 2 * cells, fractions, sizeof(float) * nfrac);
  out->boundary_cells = boundary;
  out->n_boundary = nboundary;
  out->missed_links = missed;
  if (have_body_bins) free_bins(&bb);
  if (have_support_bins) free_bins(&sb);
  free(bases);
  free(fractions);
  free(body_boundary);
  return 0;
}

static int valid_mesh(const float *p, size_t tris) {
  for (size_t k = 0; k < 9 * tris; ++k)
    if (!isfinite(p[k])) return 0;
  return 1;
}

/* voxelize.ts:172-199 */
int pwo_voxelize(const float *body, size_t body_tris, const float *supports, size_t support_tris,
                 uint32_t nx, uint32_t ny, uint32_t nz, int floor, int link_fractions,
                 pwo_voxel_grid *out, char *err, size_t err_len) {
  memset(out, 0, sizeof *out);
  if (nx < 2 || ny < 2 || nz < 2) {
    fail(err, err_len, "Grid dimensions must be integers of at least two cells.");
    return -1;
  }
  if (!valid_mesh(body, body_tris) || (supports && !valid_mesh(supports, support_tris))) {
    fail(err, err_len, "Mesh coordinates must be finite.");
    return -1;
  }
  const size_t cells = (size_t)nx * ny * nz;
  const uint32_t size[3] = {nx, ny, nz};
  uint32_t *flags = calloc(cells, sizeof(uint32_t));
  const int body_closed = vote_mesh(body, body_tris, size, flags, PWO_SOLID);
  const int supports_closed = supports ? vote_mesh(supports, support_tris, size, flags, PWO_SUPPORT) : 1;
  /* voxelize.ts:138-150 */
  for (size_t z = 0; z < nz; ++z)
    for (size_t y = 0; y < ny; ++y) {
      const size_t row = nx * (y + ny * z);
      flags[row] = PWO_EQ;
      flags[row + nx - 1] = PWO_EQ;
    }
  if (floor) {
    for (size_t c = 0; c < (size_t)nx * ny; ++c) flags[c] = PWO_FLOOR;
    for (size_t c = (size_t)nx * ny * (nz - 1); c < cells; ++c) flags[c] = PWO_EQ;
  }
  /* voxelize.ts:152-169 */
  uint8_t *bodyp = calloc((size_t)ny * nz, 1), *wallp = calloc((size_t)ny * nz, 1);
  uint64_t solid = 0;
  for (size_t c = 0; c < cells; ++c) {
    const uint32_t t = flags[c] & PWO_TYPE_MASK;
    if (t == PWO_SOLID) {
      solid += 1;
      bodyp[c / nx] = 1;
    }
    if (t == PWO_SOLID || t == PWO_SUPPORT) wallp[c / nx] = 1;
  }
  uint64_t frontal = 0, blocked = 0;
  for (size_t k = 0; k < (size_t)ny * nz; ++k) {
    frontal += bodyp[k];
    blocked += wallp[k];
  }
  free(bodyp);
  free(wallp);
  out->solid_cells = solid;
  out->frontal_cells = frontal;
  out->blockage = (double)blocked / ((double)ny * nz);
  pwo_build_wall_links(flags, body, body_tris, supports, support_tris, nx, ny, nz, floor,
                       link_fractions, out);
  out->watertight = body_closed && supports_closed;
  free(flags);
  return 0;
}

void pwo_voxel_grid_free(pwo_voxel_grid *g) {
  free(g->geometry);
  free(g->boundary_cells);
  memset(g, 0, sizeof *g);
}

/* ---- preset bodies (presets.ts) ---- */
typedef struct {
  float *out;
  size_t n;
} tri_sink;

static void emit(tri_sink *s, const double a[3], const double b[3], const double c[3]) {
  if (s->out) {
    float *p = s->out + 9 * s->n;
    for (int k = 0; k < 3; ++k) {
      p[k] = (float)a[k];
      p[3 + k] = (float)b[k];
      p[6 + k] = (float)c[k];
    }
  }
  s->n += 1;
}

static void quad(tri_sink *s, const double *a, const double *b, const double *c, const double *d) {
  emit(s, a, b, c);
  emit(s, a, c, d);
}

typedef struct {
  double x, y_half, z_min, z_max;
} station;

static void corners(const station *st, double c[4][3]) {
  const double v[4][3] = {{st->x, -st->y_half, st->z_min},
                          {st->x, st->y_half, st->z_min},
                          {st->x, st->y_half, st->z_max},
                          {st->x, -st->y_half, st->z_max}};
  memcpy(c, v, sizeof v);
}

/* presets.ts:63-90 */
static void loft(tri_sink *s, const station *sts, int n) {
  double near[4][3], far[4][3];
  for (int k = 0; k + 1 < n; ++k) {
    corners(&sts[k], near);
    corners(&sts[k + 1], far);
    for (int c = 0; c < 4; ++c) {
      const int nxt = (c + 1) % 4;
      quad(s, near[c], far[c], far[nxt], near[nxt]);
    }
  }
  double first[4][3], last[4][3];
  corners(&sts[0], first);
  corners(&sts[n - 1], last);
  quad(s, first[0], first[1], first[2], first[3]);
  quad(s, last[0], last[3], last[2], last[1]);
}

/* presets.ts:92-115 (dimensions in millimetres, presets.ts:4-17) */
static void ahmed(tri_sink *s) {
  const double radius = 100, half_width = 389.0 / 2, bottom = 50, top = 50 + 288;
  station sts[16];
  int n = 0;
  for (int seg = 0; seg <= 12; ++seg) {
    const double angle = ((double)seg / 12) * (M_PI / 2);
    const double inset = radius * (1 - sin(angle));
    sts[n++] = (station){radius * (1 - cos(angle)), half_width - inset, bottom + inset, top - inset};
  }
  const double slant = (25 * M_PI) / 180;
  const double slant_start = 1044 - 222 * cos(slant);
  const double slant_drop = 222 * sin(slant);
  sts[n++] = (station){slant_start, half_width, bottom, top};
  sts[n++] = (station){1044, half_width, bottom, top - slant_drop};
  loft(s, sts, n);
}

/* presets.ts:118-133 */
static void cylinder(tri_sink *s, double x, double y, double diameter, double height) {
  const double r = diameter / 2;
  for (int seg = 0; seg < 16; ++seg) {
    const int nxt = seg + 1;
    const double t0 = ((double)seg / 16) * 2 * M_PI, t1 = ((double)nxt / 16) * 2 * M_PI;
    const double b0[3] = {x + r * cos(t0), y + r * sin(t0), 0};
    const double b1[3] = {x + r * cos(t1), y + r * sin(t1), 0};
    const double h0[3] = {x + r * cos(t0), y + r * sin(t0), height};
    const double h1[3] = {x + r * cos(t1), y + r * sin(t1), height};
    quad(s, b0, b1, h1, h0);
    const double cb[3] = {x, y, 0}, ch[3] = {x, y, height};
    emit(s, cb, b1, b0);
    emit(s, ch, h0, h1);
  }
}

static void stilts(tri_sink *s) {
  const double xs[2] = {202, 672};
  for (int a = 0; a < 2; ++a) {
    cylinder(s, xs[a], -163.5, 30, 50);
    cylinder(s, xs[a], 163.5, 30, 50);
  }
}

/* presets.ts:152-177 */
static void sphere(tri_sink *s) {
  const double radius = 50;
  for (int ring = 0; ring < 16; ++ring)
    for (int seg = 0; seg < 24; ++seg) {
      const int nxt = (seg + 1) % 24;
      double v[4][3];
      const int rs[4] = {ring, ring + 1, ring + 1, ring};
      const int ss[4] = {seg, seg, nxt, nxt};
      for (int k = 0; k < 4; ++k) {
        const double phi = ((double)rs[k] / 16) * M_PI;
        const double theta = ((double)ss[k] / 24) * 2 * M_PI;
        v[k][0] = radius * sin(phi) * cos(theta);
        v[k][1] = radius * sin(phi) * sin(theta);
        v[k][2] = radius * cos(phi);
      }
      /* top = v0, bottom = v1, bottomNext = v2, topNext = v3 */
      if (ring != 0) emit(s, v[0], v[1], v[3]);
      if (ring != 15) emit(s, v[3], v[1], v[2]);
    }
}

static void cube(tri_sink *s) {
  const station sts[2] = {{0, 50, 0, 100}, {100, 50, 0, 100}};
  loft(s, sts, 2);
}

size_t pwo_preset_mesh(int kind, float *out) {
  tri_sink s = {out, 0};
  switch (kind) {
    case 0: sphere(&s); break;
    case 1: ahmed(&s); break;
    case 2: cube(&s); break;
    case 3: stilts(&s); break;
    default: return 0;
  }
  return s.n;
}

/* transform.ts:118-143 with flowAxis x, flowSign +1, upAxis z (lateral y, sign +1). */
int pwo_place_mesh(const float *mesh, size_t tris, const float *supports, size_t stris,
                   double length_cells, uint32_t nx, uint32_t ny, uint32_t nz, int floor,
                   float *out_body, float *out_supports) {
  if (tris < 1 || !(length_cells > 0)) return -1;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t k = 0; k < 3 * tris; ++k)
    for (int ax = 0; ax < 3; ++ax) {
      lo[ax] = fmin(lo[ax], mesh[3 * k + ax]);
      hi[ax] = fmax(hi[ax], mesh[3 * k + ax]);
    }
  double lowest = lo[2];
  for (size_t k = 0; supports && k < 3 * stris; ++k) lowest = fmin(lowest, supports[3 * k + 2]);
  const double longest = fmax(fmax(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
  if (longest == 0) return -1;
  const double scale = length_cells / longest;
  const double nose = lo[0];
  const double cy = (lo[1] + hi[1]) / 2, cz = (lo[2] + hi[2]) / 2;
  for (int m = 0; m < 2; ++m) {
    const float *src = m == 0 ? mesh : supports;
    float *dst = m == 0 ? out_body : out_supports;
    const size_t n = m == 0 ? tris : stris;
    if (!src || !dst) continue;
    for (size_t k = 0; k < 3 * n; ++k) {
      const double vx = src[3 * k], vy = src[3 * k + 1], vz = src[3 * k + 2];
      dst[3 * k] = (float)(0.25 * nx + 1 * (vx - nose) * scale);
      dst[3 * k + 1] = (float)((double)ny / 2 + 1 * (vy - cy) * scale);
      dst[3 * k + 2] = (float)(floor ? 0.5 + (vz - lowest) * scale : (double)nz / 2 + (vz - cz) * scale);
    }
  }
  return 0;
}
