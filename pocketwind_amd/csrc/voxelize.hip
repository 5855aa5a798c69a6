// voxelize.hip -- triangle-mesh voxelizer, pull wall masks and interpolated-wall link fractions
// on the GPU, bit-exact against the reference's float64 JavaScript.
//
// Follows src/geometry/voxelize.ts:23-199 (three-projection parity vote with half-open edge
// ownership, EQ/floor planes, counts), src/geometry/wall-links.ts:36-131 (pull masks, q
// offsets, body boundary list, clamped q) and src/geometry/triangle-bins.ts:14-162 (triangle
// bins, Moller-Trumbore link hits; the bins here are kBinCells wide, not 4: a link's nearest hit
// is the minimum over every triangle whose box meets the link's box, which any binning finds). Every expression keeps the reference's operand
// order and is evaluated in IEEE double with contraction disabled, so no product is fused
// into an FMA; division is the correctly rounded f64 sequence.
//
// GPU mapping: one lane per triangle rasterises its columns (a count pass, a scan, a fill pass
// into per-column hit lists), one lane per column sorts its hits and votes its cells, one lane
// per cell classifies and marks pull walls, one lane per wall link casts its ray through the bins.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "lattice.h"
#include "solver_internal.h"

namespace pw {
namespace vox {

constexpr int kBlock = 256;
// Width of a triangle bin in cells. A dense mesh puts hundreds of triangles in a 4-cell bin;
// one-cell bins leave a link a few dozen candidates (measured on the 217 088-triangle body at
// 256x128x128: the link-fraction pass 28.8 ms with a lane per cell and 4-cell bins, 14.4 ms
// with a lane per link, and the whole voxelisation 26 ms with 2-cell bins, 23 ms with 1-cell).
constexpr int kBinCells = 1;

struct Proj {
  int u, v, a;  // axis indices of column u, column v and depth
  uint32_t nu, nv, depth;
  uint32_t su, sv, sa;  // flat-index strides of the three axes
};

__device__ __forceinline__ double edge_fn(const double* a, const double* b, double u, double v) {
  return (b[0] - a[0]) * (v - a[1]) - (b[1] - a[1]) * (u - a[0]);
}

__device__ __forceinline__ bool owns_edge(double value, const double* a, const double* b) {
  return value > 0 || (value == 0 && (b[1] < a[1] || (b[1] == a[1] && b[0] > a[0])));
}

// voxelize.ts:47-75. MODE 0 counts hits per column, MODE 1 writes them.
template <int MODE>
__global__ __launch_bounds__(kBlock) void raster_kernel(const float* __restrict__ tris, uint32_t n_tris, Proj pr,
                                                        uint32_t* __restrict__ count,
                                                        const uint32_t* __restrict__ offset,
                                                        uint32_t* __restrict__ cursor, double* __restrict__ hits) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tris) return;
  const float* q = tris + 9 * (size_t)t;
  const double p0[3] = {q[pr.u], q[pr.v], q[pr.a]};
  const double p1[3] = {q[3 + pr.u], q[3 + pr.v], q[3 + pr.a]};
  const double p2[3] = {q[6 + pr.u], q[6 + pr.v], q[6 + pr.a]};
  const double signed_area = edge_fn(p0, p1, p2[0], p2[1]);
  if (signed_area == 0) return;
  const double* a = p0;
  const double* b = signed_area > 0 ? p1 : p2;
  const double* c = signed_area > 0 ? p2 : p1;
  const double area = fabs(signed_area);
  const double umin = fmax(0.0, ceil(fmin(fmin(a[0], b[0]), c[0])));
  const double umax = fmin((double)pr.nu - 1, floor(fmax(fmax(a[0], b[0]), c[0])));
  const double vmin = fmax(0.0, ceil(fmin(fmin(a[1], b[1]), c[1])));
  const double vmax = fmin((double)pr.nv - 1, floor(fmax(fmax(a[1], b[1]), c[1])));
  if (umax < umin || vmax < vmin) return;
  for (long v = (long)vmin; v <= (long)vmax; ++v) {
    for (long u = (long)umin; u <= (long)umax; ++u) {
      const double wa = edge_fn(b, c, (double)u, (double)v);
      const double wb = edge_fn(c, a, (double)u, (double)v);
      const double wc = edge_fn(a, b, (double)u, (double)v);
      if (owns_edge(wa, b, c) && owns_edge(wb, c, a) && owns_edge(wc, a, b)) {
        const uint32_t col = (uint32_t)(u + (long)pr.nu * v);
        if (MODE == 0) {
          atomicAdd(count + col, 1u);
        } else {
          const uint32_t slot = atomicAdd(cursor + col, 1u);
          hits[offset[col] + slot] = (wa * a[2] + wb * b[2] + wc * c[2]) / area;
        }
      }
    }
  }
}

// voxelize.ts:77-101: sort each column's depths, vote cells in [ceil(h0), ceil(h1)) per pair.
__global__ __launch_bounds__(kBlock) void column_vote_kernel(Proj pr, const uint32_t* __restrict__ count,
                                                             const uint32_t* __restrict__ offset,
                                                             double* __restrict__ hits,
                                                             uint8_t* __restrict__ votes, int* __restrict__ open) {
  const uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= pr.nu * pr.nv) return;
  const uint32_t n = count[col];
  if (n == 0) return;
  double* h = hits + offset[col];
  for (uint32_t k = 1; k < n; ++k) {
    const double key = h[k];
    uint32_t j = k;
    while (j > 0 && h[j - 1] > key) {
      h[j] = h[j - 1];
      --j;
    }
    h[j] = key;
  }
  if (n % 2) atomicOr(open, 1);
  const size_t base = (size_t)(col % pr.nu) * pr.su + (size_t)(col / pr.nu) * pr.sv;
  for (uint32_t k = 0; k + 1 < n; k += 2) {
    const double start = fmax(0.0, ceil(h[k]));
    const double end = fmin((double)pr.depth, ceil(h[k + 1]));
    if (!(end > start)) continue;
    for (long z = (long)start; z < (long)end; ++z) votes[base + (size_t)z * pr.sa] += 1;
  }
}

// voxelize.ts:132-134
__global__ void classify_kernel(uint32_t* __restrict__ flags, const uint8_t* __restrict__ votes, size_t n,
                                uint32_t type) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n && votes[c] >= 2 && flags[c] == kFluid) flags[c] = type;
}

// voxelize.ts:138-150: EQ inlet and outlet planes; with a floor, z = 0 is FLOOR (also at the
// x ends) and z = nz - 1 is EQ.
__global__ void boundary_planes_kernel(uint32_t* __restrict__ flags, uint32_t nx, uint32_t ny, uint32_t nz,
                                       int floor) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (size_t)nx * ny * nz) return;
  const uint32_t x = c % nx, z = (uint32_t)(c / ((size_t)nx * ny));
  uint32_t t = flags[c];
  if (x == 0 || x == nx - 1) t = kEq;
  if (floor) {
    if (z == 0) t = kFloor;
    else if (z == nz - 1) t = kEq;
  }
  flags[c] = t;
}

// voxelize.ts:152-169: one lane per (y, z) row.
__global__ void count_rows_kernel(const uint32_t* __restrict__ flags, uint32_t nx, uint32_t rows,
                                  unsigned long long* __restrict__ totals) {
  const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  unsigned long long solid = 0;
  bool body = false, wall = false;
  for (uint32_t x = 0; x < nx; ++x) {
    const uint32_t t = flags[(size_t)row * nx + x] & kTypeMask;
    if (t == kSolid) {
      ++solid;
      body = true;
    }
    if (t == kSolid || t == kSupport) wall = true;
  }
  if (solid) atomicAdd(totals + 0, solid);
  if (body) atomicAdd(totals + 1, 1ull);
  if (wall) atomicAdd(totals + 2, 1ull);
}

// wall-links.ts:36-63 in pull form: incoming i of fluid cell f pulls from f - c_i, where x
// never wraps, y wraps, and z wraps only without a floor.
__global__ void mark_walls_kernel(const uint32_t* __restrict__ kinds, uint32_t* __restrict__ flags,
                                  uint32_t* __restrict__ link_count, uint32_t* __restrict__ body_boundary,
                                  uint32_t nx, uint32_t ny, uint32_t nz, int floor) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t N = (size_t)nx * ny * nz;
  if (c >= N) return;
  const uint32_t kind = kinds[c] & kTypeMask;
  uint32_t mask = 0, body = 0;
  if (kind == kFluid) {
    const long x = c % nx, y = (c / nx) % ny, z = c / ((size_t)nx * ny);
#pragma unroll
    for (int i = 1; i < kQ; ++i) {
      const long wx = x - cx(i), wy = y - cy(i), wz = z - cz(i);
      if (wx < 0 || wx >= (long)nx || (floor && (wz < 0 || wz >= (long)nz))) continue;
      const size_t w = wx + (size_t)nx * (((wy + ny) % ny) + (size_t)ny * ((wz + nz) % nz));
      const uint32_t wt = kinds[w] & kTypeMask;
      if (wt >= kSolid && wt <= kFloor) {
        mask |= 1u << (3 + i);
        if (wt == kSolid) body = 1;
      }
    }
  }
  flags[c] = kind | mask;
  link_count[c] = __builtin_popcount(mask);
  body_boundary[c] = body;
}

struct Bins {
  const float* tris;
  const float* bounds;
  const uint32_t* start;
  const uint32_t* items;
  long nbx, nby, nbz;
};

// triangle-bins.ts:17-54: f32 bounds and the bin range of one triangle.
__global__ void bin_bounds_kernel(const float* __restrict__ tris, uint32_t n_tris, float* __restrict__ bounds,
                                  int* __restrict__ range, long nbx, long nby, long nbz,
                                  uint32_t* __restrict__ count, int mode, const uint32_t* __restrict__ start,
                                  uint32_t* __restrict__ cursor, uint32_t* __restrict__ items) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tris) return;
  const long nb[3] = {nbx, nby, nbz};
  long lo_b[3], hi_b[3];
  for (int ax = 0; ax < 3; ++ax) {
    const float p = tris[9 * (size_t)t + ax], q = tris[9 * (size_t)t + 3 + ax], r = tris[9 * (size_t)t + 6 + ax];
    const float lo = fminf(fminf(p, q), r), hi = fmaxf(fmaxf(p, q), r);
    bounds[6 * (size_t)t + ax] = lo;
    bounds[6 * (size_t)t + 3 + ax] = hi;
    long mn = (long)floor((double)lo / kBinCells), mx = (long)floor((double)hi / kBinCells);
    lo_b[ax] = mn < 0 ? 0 : mn;
    hi_b[ax] = mx > nb[ax] - 1 ? nb[ax] - 1 : mx;
  }
  for (long z = lo_b[2]; z <= hi_b[2]; ++z)
    for (long y = lo_b[1]; y <= hi_b[1]; ++y)
      for (long x = lo_b[0]; x <= hi_b[0]; ++x) {
        const size_t k = (size_t)(x + nbx * (y + nby * z));
        if (mode == 0) atomicAdd(count + k, 1u);
        else items[start[k] + atomicAdd(cursor + k, 1u)] = t;
      }
}

// triangle-bins.ts:60-94
__device__ bool intersect(const float* p, const double o[3], const double d[3], double* t_out) {
  const double ax = p[0], ay = p[1], az = p[2];
  const double e1x = (double)p[3] - ax, e1y = (double)p[4] - ay, e1z = (double)p[5] - az;
  const double e2x = (double)p[6] - ax, e2y = (double)p[7] - ay, e2z = (double)p[8] - az;
  const double px = d[1] * e2z - d[2] * e2y;
  const double py = d[2] * e2x - d[0] * e2z;
  const double pz = d[0] * e2y - d[1] * e2x;
  const double det = e1x * px + e1y * py + e1z * pz;
  if (fabs(det) < DBL_EPSILON) return false;
  const double tx = o[0] - ax, ty = o[1] - ay, tz = o[2] - az;
  const double u = (tx * px + ty * py + tz * pz) / det;
  if (u < -1e-9 || u > 1 + 1e-9) return false;
  const double qx = ty * e1z - tz * e1y;
  const double qy = tz * e1x - tx * e1z;
  const double qz = tx * e1y - ty * e1x;
  const double v = (d[0] * qx + d[1] * qy + d[2] * qz) / det;
  if (v < -1e-9 || u + v > 1 + 1e-9) return false;
  const double t = (e2x * qx + e2y * qy + e2z * qz) / det;
  if (!(t >= -1e-9 && t <= 1 + 1e-9)) return false;
  *t_out = t;
  return true;
}

// triangle-bins.ts:98-162
__device__ bool nearest_hit(const Bins& b, const double o[3], const double d[3], double* t_out) {
  double lo[3], hi[3];
  long bmin[3], bmax[3];
  const long nb[3] = {b.nbx, b.nby, b.nbz};
  for (int ax = 0; ax < 3; ++ax) {
    lo[ax] = fmin(o[ax], o[ax] + d[ax]);
    hi[ax] = fmax(o[ax], o[ax] + d[ax]);
    bmin[ax] = (long)floor(lo[ax] / kBinCells);
    bmax[ax] = (long)floor(hi[ax] / kBinCells);
    if (bmin[ax] < 0) bmin[ax] = 0;
    if (bmax[ax] > nb[ax] - 1) bmax[ax] = nb[ax] - 1;
  }
  double nearest = INFINITY;
  for (long z = bmin[2]; z <= bmax[2]; ++z)
    for (long y = bmin[1]; y <= bmax[1]; ++y)
      for (long x = bmin[0]; x <= bmax[0]; ++x) {
        const size_t k = (size_t)(x + b.nbx * (y + b.nby * z));
        for (uint32_t e = b.start[k]; e < b.start[k + 1]; ++e) {
          const uint32_t t = b.items[e];
          const float* bd = b.bounds + 6 * (size_t)t;
          if (!(bd[0] <= hi[0] && bd[3] >= lo[0] && bd[1] <= hi[1] && bd[4] >= lo[1] && bd[2] <= hi[2] &&
                bd[5] >= lo[2]))
            continue;
          double tt;
          if (intersect(b.tris + 9 * (size_t)t, o, d, &tt) && tt < nearest) nearest = tt;
        }
      }
  if (nearest == INFINITY) return false;
  *t_out = nearest;
  return true;
}

// The cell of every wall link: link base[c] + k belongs to cell c (links are packed per cell in
// cell order, ascending direction order within a cell).
__global__ void link_cells_kernel(const uint32_t* __restrict__ flags, const uint32_t* __restrict__ bases,
                                  uint32_t* __restrict__ link_cell, size_t n) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const uint32_t mask = flags[c] >> 3;
  if (mask == 0) return;
  const uint32_t base = bases[c], count = __builtin_popcount(mask);
  for (uint32_t k = 0; k < count; ++k) link_cell[base + k] = (uint32_t)c;
}

// wall-links.ts:65-92 and :114-125, one lane per wall link (the links are sparse in the lattice:
// a lane per cell would leave most lanes of a wavefront idle while a few cast their rays).
__global__ void link_fraction_kernel(const uint32_t* __restrict__ flags, const uint32_t* __restrict__ bases,
                                     const uint32_t* __restrict__ link_cell, uint32_t n_links,
                                     float* __restrict__ q, uint32_t nx, uint32_t ny, uint32_t nz, int floor,
                                     int measure, Bins body, Bins supports, int have_supports,
                                     unsigned long long* __restrict__ missed) {
  const uint32_t link = blockIdx.x * blockDim.x + threadIdx.x;
  if (link >= n_links) return;
  const size_t c = link_cell[link];
  const uint32_t flag = flags[c];
  // Direction of the link: the (link - base)-th set wall bit of the cell, in ascending order.
  uint32_t rank = link - bases[c];
  int i = 1;
  for (; i < kQ; ++i)
    if (has_wall(flag, i) && rank-- == 0) break;
  const long x = c % nx, y = (c / nx) % ny, z = c / ((size_t)nx * ny);
  double frac = 0.5;
  if (measure) {
    const long wx = x - cx(i), wy = ((y - cy(i)) % (long)ny + ny) % ny, wz = ((z - cz(i)) % (long)nz + nz) % nz;
    const size_t w = wx + (size_t)nx * (wy + (size_t)ny * wz);
    const uint32_t wt = flags[w] & kTypeMask;
    if (wt != kFloor) {
      const Bins& bins = (wt == kSupport && have_supports) ? supports : body;
      const double o[3] = {(double)wx + cx(i), (double)wy + cy(i), (double)wz + cz(i)};
      const double d[3] = {(double)-cx(i), (double)-cy(i), (double)-cz(i)};
      double t;
      if (nearest_hit(bins, o, d, &t)) frac = t;
      else atomicAdd(missed, 1ull);
    }
  }
  q[link] = (float)fmax(1e-3, fmin(1.0, frac));
}

__global__ void bases_kernel(const uint32_t* __restrict__ count, const uint32_t* __restrict__ prefix,
                             uint32_t* __restrict__ bases, size_t n) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) bases[c] = count[c] ? prefix[c] : kNoQ;
}

__global__ void compact_kernel(const uint32_t* __restrict__ keep, const uint32_t* __restrict__ prefix,
                               uint32_t* __restrict__ out, size_t n) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n && keep[c]) out[prefix[c]] = (uint32_t)c;
}

inline unsigned grid_for(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// Device allocations of one voxelize call, released together.
struct Arena {
  std::vector<void*> ptrs;
  hipError_t err = hipSuccess;
  template <class T>
  T* alloc(size_t n) {
    void* p = nullptr;
    if (err != hipSuccess) return nullptr;
    err = hipMalloc(&p, (n ? n : 1) * sizeof(T));
    if (err != hipSuccess) return nullptr;
    ptrs.push_back(p);
    return (T*)p;
  }
  ~Arena() {
    for (void* p : ptrs) (void)hipFree(p);
  }
};

// Exclusive prefix sum of n u32 values; returns the total through *total.
hipError_t exclusive_scan(Arena& ar, const uint32_t* in, uint32_t* out, size_t n, hipStream_t st, uint64_t* total) {
  size_t temp = 0;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, temp, in, out, (int)n, st);
  if (e != hipSuccess) return e;
  void* tmp = ar.alloc<uint8_t>(temp);
  if (ar.err != hipSuccess) return ar.err;
  if ((e = hipcub::DeviceScan::ExclusiveSum(tmp, temp, in, out, (int)n, st)) != hipSuccess) return e;
  uint32_t last_in = 0, last_out = 0;
  if (n) {
    if ((e = hipMemcpyAsync(&last_in, in + n - 1, 4, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(&last_out, out + n - 1, 4, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  }
  *total = (uint64_t)last_in + last_out;
  return hipSuccess;
}

// voxelize.ts:103-136 for one mesh.
hipError_t vote_mesh(Arena& ar, const float* tris, uint32_t n_tris, const uint32_t size[3], uint32_t* flags,
                     uint32_t type, hipStream_t st, int* open_dev) {
  const size_t N = (size_t)size[0] * size[1] * size[2];
  const uint32_t strides[3] = {1, size[0], size[0] * size[1]};
  static const int proj[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};
  uint8_t* votes = ar.alloc<uint8_t>(N);
  if (ar.err != hipSuccess) return ar.err;
  hipError_t e = hipMemsetAsync(votes, 0, N, st);
  if (e != hipSuccess) return e;
  for (int p = 0; p < 3; ++p) {
    Proj pr{proj[p][0], proj[p][1], proj[p][2], size[proj[p][0]], size[proj[p][1]], size[proj[p][2]],
            strides[proj[p][0]], strides[proj[p][1]], strides[proj[p][2]]};
    const size_t ncol = (size_t)pr.nu * pr.nv;
    uint32_t* count = ar.alloc<uint32_t>(ncol);
    uint32_t* offset = ar.alloc<uint32_t>(ncol);
    uint32_t* cursor = ar.alloc<uint32_t>(ncol);
    if (ar.err != hipSuccess) return ar.err;
    if ((e = hipMemsetAsync(count, 0, ncol * 4, st)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(cursor, 0, ncol * 4, st)) != hipSuccess) return e;
    if (n_tris) hipLaunchKernelGGL(raster_kernel<0>, dim3(grid_for(n_tris)), dim3(kBlock), 0, st, tris, n_tris, pr, count, offset, cursor, nullptr);
    uint64_t total = 0;
    if ((e = exclusive_scan(ar, count, offset, ncol, st, &total)) != hipSuccess) return e;
    double* hits = ar.alloc<double>(total);
    if (ar.err != hipSuccess) return ar.err;
    if (n_tris && total) hipLaunchKernelGGL(raster_kernel<1>, dim3(grid_for(n_tris)), dim3(kBlock), 0, st, tris, n_tris, pr, count, offset, cursor, hits);
    hipLaunchKernelGGL(column_vote_kernel, dim3(grid_for(ncol)), dim3(kBlock), 0, st, pr, count, offset, hits, votes, open_dev);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(classify_kernel, dim3(grid_for(N)), dim3(kBlock), 0, st, flags, votes, N, type);
  return hipGetLastError();
}

hipError_t build_bins(Arena& ar, const float* tris, uint32_t n_tris, const uint32_t size[3], hipStream_t st, Bins* out) {
  Bins b{};
  b.tris = tris;
  b.nbx = (size[0] + kBinCells - 1) / kBinCells;
  b.nby = (size[1] + kBinCells - 1) / kBinCells;
  b.nbz = (size[2] + kBinCells - 1) / kBinCells;
  const size_t nbins = (size_t)b.nbx * b.nby * b.nbz;
  float* bounds = ar.alloc<float>(6 * (size_t)n_tris);
  uint32_t* count = ar.alloc<uint32_t>(nbins + 1);
  uint32_t* start = ar.alloc<uint32_t>(nbins + 1);
  uint32_t* cursor = ar.alloc<uint32_t>(nbins);
  if (ar.err != hipSuccess) return ar.err;
  hipError_t e;
  if ((e = hipMemsetAsync(count, 0, (nbins + 1) * 4, st)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(cursor, 0, nbins * 4, st)) != hipSuccess) return e;
  if (n_tris) hipLaunchKernelGGL(bin_bounds_kernel, dim3(grid_for(n_tris)), dim3(kBlock), 0, st, tris, n_tris, bounds, nullptr, b.nbx, b.nby, b.nbz, count, 0, nullptr, nullptr, nullptr);
  uint64_t total = 0;
  if ((e = exclusive_scan(ar, count, start, nbins + 1, st, &total)) != hipSuccess) return e;
  uint32_t* items = ar.alloc<uint32_t>(total);
  if (ar.err != hipSuccess) return ar.err;
  if (n_tris && total) hipLaunchKernelGGL(bin_bounds_kernel, dim3(grid_for(n_tris)), dim3(kBlock), 0, st, tris, n_tris, bounds, nullptr, b.nbx, b.nby, b.nbz, count, 1, start, cursor, items);
  b.bounds = bounds;
  b.start = start;
  b.items = items;
  *out = b;
  return hipGetLastError();
}

bool finite_mesh(const float* p, size_t tris) {
  for (size_t k = 0; k < 9 * tris; ++k)
    if (!std::isfinite(p[k])) return false;
  return true;
}

}  // namespace vox
}  // namespace pw

using namespace pw;
using namespace pw::vox;

namespace pw {
namespace vox {

// Pinned staging for the results' device-to-host copies: copies into pageable memory ran at
// 2-4 GB/s (7.3 + 3.9 ms for the 2N geometry words of 256x128x128), into pinned memory at
// PCIe rate. One buffer per process, grown on demand and kept; callers take it in turn.
std::mutex g_staging_mutex;
void* g_staging = nullptr;
size_t g_staging_bytes = 0;

hipError_t staging(size_t bytes, void** out) {
  if (bytes > g_staging_bytes) {
    if (g_staging) (void)hipHostFree(g_staging);
    g_staging = nullptr;
    g_staging_bytes = 0;
    hipError_t e = hipHostMalloc(&g_staging, bytes, hipHostMallocDefault);
    if (e != hipSuccess) return e;
    g_staging_bytes = bytes;
  }
  *out = g_staging;
  return hipSuccess;
}

}  // namespace vox
}  // namespace pw

extern "C" pw_status pw_voxelize(const float* body, size_t body_tris, const float* supports, size_t support_tris,
                                 const pw_layout* layout, int floor, int link_fractions, pw_voxel_grid* out) {
  if (!layout || !out || (!body && body_tris) || (!supports && support_tris)) {
    set_error("Null argument.");
    return PW_EINVAL;
  }
  std::memset(out, 0, sizeof *out);
  const uint32_t size[3] = {layout->nx, layout->ny, layout->nz};
  if (size[0] < 2 || size[1] < 2 || size[2] < 2) {
    set_error("Grid dimensions must be integers of at least two cells.");
    return PW_EINVAL;
  }
  if (!finite_mesh(body, body_tris) || (supports && !finite_mesh(supports, support_tris))) {
    set_error("Mesh coordinates must be finite.");
    return PW_EINVAL;
  }
  const size_t N = (size_t)size[0] * size[1] * size[2];
  if (N > 0xffffffffull / 2 || body_tris > 0x7fffffff || support_tris > 0x7fffffff) {
    set_error("Grid or mesh exceeds 32-bit indexing.");
    return PW_EINVAL;
  }
  int devices = 0;
  if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) {
    set_error("No HIP device is available.");
    return PW_EDEVICE;
  }
  hipStream_t st = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e != hipSuccess) return device_error(e, "voxelize stream");
  pw_status status = PW_OK;
  {
    Arena ar;
    float* d_body = ar.alloc<float>(9 * body_tris);
    float* d_sup = supports ? ar.alloc<float>(9 * support_tris) : nullptr;
    uint32_t* kinds = ar.alloc<uint32_t>(N);
    uint32_t* flags = ar.alloc<uint32_t>(N);
    uint32_t* links = ar.alloc<uint32_t>(N);
    uint32_t* prefix = ar.alloc<uint32_t>(N);
    uint32_t* bases = ar.alloc<uint32_t>(N);
    uint32_t* boundary = ar.alloc<uint32_t>(N);
    uint32_t* bprefix = ar.alloc<uint32_t>(N);
    unsigned long long* totals = ar.alloc<unsigned long long>(4);
    int* open = ar.alloc<int>(1);
    std::vector<uint32_t> host;
    uint64_t n_links = 0, n_bound = 0;
    unsigned long long tot[4] = {0, 0, 0, 0};
    int open_h = 0;
    Bins bb{}, sb{};
    float* q = nullptr;
    uint32_t* bcells = nullptr;
    if ((e = ar.err) != hipSuccess) goto fail;
    if (body_tris && (e = hipMemcpyAsync(d_body, body, 36 * body_tris, hipMemcpyHostToDevice, st)) != hipSuccess) goto fail;
    if (supports && support_tris && (e = hipMemcpyAsync(d_sup, supports, 36 * support_tris, hipMemcpyHostToDevice, st)) != hipSuccess) goto fail;
    if ((e = hipMemsetAsync(kinds, 0, N * 4, st)) != hipSuccess) goto fail;
    if ((e = hipMemsetAsync(totals, 0, 32, st)) != hipSuccess) goto fail;
    if ((e = hipMemsetAsync(open, 0, 4, st)) != hipSuccess) goto fail;
    if ((e = vote_mesh(ar, d_body, (uint32_t)body_tris, size, kinds, kSolid, st, open)) != hipSuccess) goto fail;
    if (supports && (e = vote_mesh(ar, d_sup, (uint32_t)support_tris, size, kinds, kSupport, st, open)) != hipSuccess) goto fail;
    hipLaunchKernelGGL(boundary_planes_kernel, dim3(grid_for(N)), dim3(kBlock), 0, st, kinds, size[0], size[1], size[2], floor);
    hipLaunchKernelGGL(count_rows_kernel, dim3(grid_for((size_t)size[1] * size[2])), dim3(kBlock), 0, st, kinds, size[0], size[1] * size[2], totals);
    hipLaunchKernelGGL(mark_walls_kernel, dim3(grid_for(N)), dim3(kBlock), 0, st, kinds, flags, links, boundary, size[0], size[1], size[2], floor);
    if ((e = hipGetLastError()) != hipSuccess) goto fail;
    if ((e = exclusive_scan(ar, links, prefix, N, st, &n_links)) != hipSuccess) goto fail;
    if ((e = exclusive_scan(ar, boundary, bprefix, N, st, &n_bound)) != hipSuccess) goto fail;
    hipLaunchKernelGGL(bases_kernel, dim3(grid_for(N)), dim3(kBlock), 0, st, links, prefix, bases, N);
    q = ar.alloc<float>(n_links);
    bcells = ar.alloc<uint32_t>(n_bound);
    if ((e = ar.err) != hipSuccess) goto fail;
    hipLaunchKernelGGL(compact_kernel, dim3(grid_for(N)), dim3(kBlock), 0, st, boundary, bprefix, bcells, N);
    if (link_fractions) {
      if ((e = build_bins(ar, d_body, (uint32_t)body_tris, size, st, &bb)) != hipSuccess) goto fail;
      if (supports && (e = build_bins(ar, d_sup, (uint32_t)support_tris, size, st, &sb)) != hipSuccess) goto fail;
    }
    if (n_links) {
      uint32_t* link_cell = ar.alloc<uint32_t>(n_links);
      if ((e = ar.err) != hipSuccess) goto fail;
      hipLaunchKernelGGL(link_cells_kernel, dim3(grid_for(N)), dim3(kBlock), 0, st, flags, bases, link_cell, N);
      hipLaunchKernelGGL(link_fraction_kernel, dim3(grid_for(n_links)), dim3(kBlock), 0, st, flags, bases, link_cell,
                         (uint32_t)n_links, q, size[0], size[1], size[2], floor, link_fractions, bb, sb,
                         supports ? 1 : 0, totals + 3);
    }
    if ((e = hipGetLastError()) != hipSuccess) goto fail;
    out->geometry_words = 2 * N + n_links;
    out->geometry = (uint32_t*)std::malloc(out->geometry_words * 4);
    out->n_boundary = n_bound;
    out->boundary_cells = (uint32_t*)std::malloc((n_bound ? n_bound : 1) * 4);
    if (!out->geometry || !out->boundary_cells) {
      std::free(out->geometry);
      std::free(out->boundary_cells);
      std::memset(out, 0, sizeof *out);
      set_error("Host allocation failed.");
      status = PW_ENOMEM;
      goto done;
    }
    {
      // geometry (2N + links words) and boundary list through the pinned staging buffer
      std::lock_guard<std::mutex> lock(g_staging_mutex);
      const size_t words = out->geometry_words + n_bound;
      uint32_t* stage = nullptr;
      if ((e = staging(words * 4, reinterpret_cast<void**>(&stage))) != hipSuccess) goto fail;
      if ((e = hipMemcpyAsync(stage, flags, N * 4, hipMemcpyDeviceToHost, st)) != hipSuccess) goto fail;
      if ((e = hipMemcpyAsync(stage + N, bases, N * 4, hipMemcpyDeviceToHost, st)) != hipSuccess) goto fail;
      if (n_links && (e = hipMemcpyAsync(stage + 2 * N, q, n_links * 4, hipMemcpyDeviceToHost, st)) != hipSuccess) goto fail;
      if (n_bound && (e = hipMemcpyAsync(stage + out->geometry_words, bcells, n_bound * 4, hipMemcpyDeviceToHost, st)) != hipSuccess) goto fail;
      if ((e = hipMemcpyAsync(tot, totals, 32, hipMemcpyDeviceToHost, st)) != hipSuccess) goto fail;
      if ((e = hipMemcpyAsync(&open_h, open, 4, hipMemcpyDeviceToHost, st)) != hipSuccess) goto fail;
      if ((e = hipStreamSynchronize(st)) != hipSuccess) goto fail;
      std::memcpy(out->geometry, stage, out->geometry_words * 4);
      if (n_bound) std::memcpy(out->boundary_cells, stage + out->geometry_words, n_bound * 4);
    }
    out->solid_cells = tot[0];
    out->frontal_cells = tot[1];
    out->blockage = (double)tot[2] / ((double)size[1] * size[2]);
    out->missed_links = tot[3];
    out->watertight = open_h ? 0 : 1;
    goto done;
  fail:
    pw_voxel_grid_free(out);
    status = device_error(e, "voxelize");
  done:;
    (void)hipStreamSynchronize(st);
  }
  (void)hipStreamDestroy(st);
  return status;
}


extern "C" void pw_voxel_grid_free(pw_voxel_grid* g) {
  if (!g) return;
  std::free(g->geometry);
  std::free(g->boundary_cells);
  std::memset(g, 0, sizeof *g);
}
