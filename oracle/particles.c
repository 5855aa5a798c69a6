/*
 * particles.c -- CPU oracle for streakline particles.
 *
 * TEST INFRASTRUCTURE ONLY (the checker). Restates:
 *   src/render/particle-layout.ts:14-112      seed region, reseed rule, lowbias32 hash, seeds
 *   src/render/shaders/particles_update.wgsl  trilinear advection step, in f32 arithmetic with
 *                                             WGSL's mix(a, b, t) = a (1 - t) + b t
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pw_oracle.h"

/* particle-layout.ts:58-63 */
uint32_t pwo_particle_hash(uint32_t value) {
  uint32_t r = value;
  r = (r ^ (r >> 16)) * 0x7feb352du;
  r = (r ^ (r >> 15)) * 0x846ca68bu;
  return r ^ (r >> 16);
}

/* particle-layout.ts:14-36 */
int pwo_seed_region(const double bmin[3], const double bmax[3], uint32_t nx, uint32_t ny,
                    uint32_t nz, double region[4]) {
  const double ey = bmax[1] - bmin[1], ez = bmax[2] - bmin[2];
  region[0] = fmax(1, bmin[1] - ey * 0.15);
  region[1] = fmax(1, bmin[2] - ez * 0.15);
  region[2] = fmin((double)ny - 2, bmax[1] + ey * 0.15);
  region[3] = fmin((double)nz - 2, bmax[2] + ez * 0.15);
  for (int k = 0; k < 3; ++k)
    if (!isfinite(bmin[k]) || !isfinite(bmax[k])) return -1;
  if (nx < 5 || ny < 5 || nz < 5 || !(ey > 0) || !(ez > 0) || region[0] >= region[2] ||
      region[1] >= region[3])
    return -1;
  return 0;
}

/* particle-layout.ts:39-56 */
static int needs_reseed(const double p[3], double age, uint32_t nx, uint32_t ny, uint32_t nz,
                        uint32_t flag) {
  return !isfinite(p[0]) || !isfinite(p[1]) || !isfinite(p[2]) || p[0] < 0 ||
         p[0] >= (double)nx - 2 || p[1] < 0 || p[1] >= (double)ny - 1 || p[2] < 0 ||
         p[2] >= (double)nz - 1 || age > 4000 || (flag & PWO_TYPE_MASK) != PWO_FLUID;
}

int pwo_particle_needs_reseed(const double p[3], double age, uint32_t nx, uint32_t ny, uint32_t nz,
                              uint32_t flag) {
  return needs_reseed(p, age, nx, ny, nz, flag);
}

/* JS Math.round: round half up. */
static double js_round(double v) { return floor(v + 0.5); }

/* particle-layout.ts:66-112 (count limit left to the caller). */
int pwo_particle_seeds(const uint32_t *flags, uint32_t nx, uint32_t ny, uint32_t nz,
                       const double bmin[3], const double bmax[3], uint32_t count, uint32_t seed,
                       float *data, float fallback[3], double region[4], char *err,
                       size_t err_len) {
  if (pwo_seed_region(bmin, bmax, nx, ny, nz, region) != 0) {
    if (err) snprintf(err, err_len, "Particle seeds need a positive cross-section.");
    return -1;
  }
  /* Inlet candidates in (z, y) order. */
  size_t n_inlet = 0;
  const long z0 = (long)js_round(region[1]), z1 = (long)js_round(region[3]);
  const long y0 = (long)js_round(region[0]), y1 = (long)js_round(region[2]);
  float *inlet = NULL;
  size_t cap = 0;
  for (long z = z0; z <= z1; ++z)
    for (long y = y0; y <= y1; ++y) {
      const float py = (float)fmax(region[0], fmin((double)y, region[2]));
      const float pz = (float)fmax(region[1], fmin((double)z, region[3]));
      const double p[3] = {2, py, pz};
      const long cy = (long)js_round(py), cz = (long)js_round(pz);
      const size_t cell = 2 + (size_t)nx * ((size_t)cy + (size_t)ny * (size_t)cz);
      const uint32_t flag = cell < (size_t)nx * ny * nz ? flags[cell] : PWO_SOLID;
      if (needs_reseed(p, 0, nx, ny, nz, flag)) continue;
      if (n_inlet == cap) {
        cap = cap ? 2 * cap : 256;
        inlet = (float *)realloc(inlet, sizeof(float) * 3 * cap);
      }
      inlet[3 * n_inlet] = 2;
      inlet[3 * n_inlet + 1] = py;
      inlet[3 * n_inlet + 2] = pz;
      n_inlet += 1;
    }
  if (n_inlet == 0) {
    if (err) snprintf(err, err_len, "The particle seed region has no fluid inlet cells.");
    return -1;
  }
  memcpy(fallback, inlet, sizeof(float) * 3);
  for (uint32_t id = 0; id < count; ++id) {
    const float *p = inlet + 3 * (pwo_particle_hash(id + seed) % n_inlet);
    float *d = data + 8 * (size_t)id;
    d[0] = p[0], d[1] = p[1], d[2] = p[2], d[3] = 0;
    d[4] = p[0], d[5] = p[1], d[6] = p[2], d[7] = 0;
  }
  free(inlet);
  return 0;
}

/* ---- particles_update.wgsl, f32 ---- */
typedef struct {
  uint32_t nx, ny, nz;
  const float *macro;
  const uint32_t *flags;
} field;

static float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

static size_t cell_of(const field *f, uint32_t x, uint32_t y, uint32_t z) {
  return x + (size_t)f->nx * (y + (size_t)f->ny * z);
}

/* particles_update.wgsl:22-25 */
static int is_fluid(const field *f, const float p[3]) {
  if (p[0] < 0 || p[1] < 0 || p[2] < 0) return 0;
  if (p[0] >= (float)(f->nx - 1) || p[1] >= (float)(f->ny - 1) || p[2] >= (float)(f->nz - 1)) return 0;
  const uint32_t x = (uint32_t)floorf(p[0] + 0.5f), y = (uint32_t)floorf(p[1] + 0.5f),
                 z = (uint32_t)floorf(p[2] + 0.5f);
  return (f->flags[cell_of(f, x, y, z)] & 7u) == 0u;
}

/* particles_update.wgsl:26-36 */
static void velocity_at(const field *f, const float p[3], float out[3]) {
  const float lim[3] = {(float)(f->nx - 1), (float)(f->ny - 1), (float)(f->nz - 1)};
  float pt[3], w[3];
  uint32_t lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    pt[k] = fminf(fmaxf(p[k], 0.0f), lim[k]);
    const float fl = floorf(pt[k]);
    lo[k] = (uint32_t)fl;
    hi[k] = lo[k] + 1 < (uint32_t)lim[k] ? lo[k] + 1 : (uint32_t)lim[k];
    w[k] = pt[k] - fl;
  }
  for (int c = 0; c < 3; ++c) {
    const float *m = f->macro + 1 + c;
#define M(X, Y, Z) m[4 * cell_of(f, X, Y, Z)]
    const float x00 = mixf(M(lo[0], lo[1], lo[2]), M(hi[0], lo[1], lo[2]), w[0]);
    const float x10 = mixf(M(lo[0], hi[1], lo[2]), M(hi[0], hi[1], lo[2]), w[0]);
    const float x01 = mixf(M(lo[0], lo[1], hi[2]), M(hi[0], lo[1], hi[2]), w[0]);
    const float x11 = mixf(M(lo[0], hi[1], hi[2]), M(hi[0], hi[1], hi[2]), w[0]);
#undef M
    out[c] = mixf(mixf(x00, x10, w[1]), mixf(x01, x11, w[1]), w[2]);
  }
}

void pwo_particles_update(float *particles, uint32_t count, const float *macro,
                          const uint32_t *flags, uint32_t nx, uint32_t ny, uint32_t nz,
                          uint32_t seed, float steps, const float cs[4], const float fallback[2]) {
  const field f = {nx, ny, nz, macro, flags};
  for (uint32_t id = 0; id < count; ++id) {
    float *q = particles + 8 * (size_t)id;
    float v[3];
    velocity_at(&f, q, v);
    float d[3] = {v[0] * steps, v[1] * steps, v[2] * steps};
    const float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const float div = fmaxf(1.0f, len);
    for (int k = 0; k < 3; ++k) d[k] /= div;
    const float p[3] = {q[0] + d[0], q[1] + d[1], q[2] + d[2]};
    const float age = q[3] + 1.0f;
    if (p[0] >= (float)(nx - 2) || !is_fluid(&f, p) || age > 4000.0f) {
      /* particles_update.wgsl:37-45 */
      const float generation = q[7] + 1.0f;
      const uint32_t h = pwo_particle_hash(id + seed + (uint32_t)generation * 0x9e3779b9u);
      const float r0 = (float)(h >> 8) / 16777216.0f, r1 = (float)(pwo_particle_hash(h) >> 8) / 16777216.0f;
      float pt[3] = {2.0f, mixf(cs[0], cs[2], r0), mixf(cs[1], cs[3], r1)};
      if (!is_fluid(&f, pt)) {
        pt[1] = fallback[0];
        pt[2] = fallback[1];
      }
      q[0] = pt[0], q[1] = pt[1], q[2] = pt[2], q[3] = 0.0f;
      q[4] = pt[0], q[5] = pt[1], q[6] = pt[2], q[7] = generation;
    } else {
      q[4] = q[0], q[5] = q[1], q[6] = q[2];
      q[0] = p[0], q[1] = p[1], q[2] = p[2], q[3] = age;
    }
  }
}
