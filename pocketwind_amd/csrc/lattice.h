// lattice.h -- D3Q19 constants and population storage codecs for gfx950 kernels.
//
// Velocity order and weights follow src/lbm/lattice.ts:2-52 (opposite pairs adjacent,
// i odd <-> i + 1); the codecs follow src/lbm/shaders/prelude-f16.wgsl:7-22 (f16 deviation
// storage scaled by 2^15 with lowbias32 proportional rounding). bf16 storage is this
// library's own format: the same hash decides the 16 dropped mantissa bits, no scale.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pw {

constexpr int kQ = 19;

__host__ __device__ constexpr int cx(int i) {
  constexpr int v[kQ] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
  return v[i];
}
__host__ __device__ constexpr int cy(int i) {
  constexpr int v[kQ] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1};
  return v[i];
}
__host__ __device__ constexpr int cz(int i) {
  constexpr int v[kQ] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1};
  return v[i];
}
__host__ __device__ constexpr float wf(int i) {
  return i == 0 ? 1.0f / 3.0f : (i < 7 ? 1.0f / 18.0f : 1.0f / 36.0f);
}
__host__ __device__ constexpr int opp(int i) { return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1); }

__host__ __device__ __forceinline__ float as_float(uint32_t u) { return __builtin_bit_cast(float, u); }
__host__ __device__ __forceinline__ uint32_t as_uint(float f) { return __builtin_bit_cast(uint32_t, f); }

enum : uint32_t { kFluid = 0, kSolid = 1, kSupport = 2, kFloor = 3, kEq = 4, kTypeMask = 7 };
constexpr uint32_t kNoQ = 0xffffffffu;

__host__ __device__ __forceinline__ bool is_wall(uint32_t kind) { return kind - 1u < 3u; }
__host__ __device__ __forceinline__ bool has_wall(uint32_t flag, int i) { return (flag >> (3 + i)) & 1u; }

// lowbias32 over (value bits, cell, direction) (prelude-f16.wgsl:8-13).
__host__ __device__ __forceinline__ uint32_t storage_hash(uint32_t bits, uint32_t i, uint32_t n) {
  uint32_t r = bits ^ (n * 0x9e3779b9u) ^ (i * 0x85ebca6bu);
  r = (r ^ (r >> 16)) * 0x7feb352du;
  r = (r ^ (r >> 15)) * 0x846ca68bu;
  return r ^ (r >> 16);
}

// Storage codecs: load() returns the f32 deviation g = f - w; store() takes g.
struct StoreF32 {
  using T = float;
  static constexpr int kBytes = 4;
  __host__ __device__ __forceinline__ static float decode(T v) { return v; }
  __host__ __device__ __forceinline__ static float load(const T* p) { return *p; }
  __host__ __device__ __forceinline__ static T encode(float g, uint32_t, uint32_t) { return g; }
};

struct StoreF16 {
  using T = _Float16;
  static constexpr int kBytes = 2;
  static constexpr float kScale = 32768.0f;
  __host__ __device__ __forceinline__ static float decode(T v) { return static_cast<float>(v) * (1.0f / kScale); }
  __host__ __device__ __forceinline__ static float load(const T* p) { return decode(*p); }
  __host__ __device__ __forceinline__ static T encode(float g, uint32_t i, uint32_t n) {
    const float value = g * kScale;
    const uint32_t bits = as_uint(value);
    const uint32_t mag = bits & 0x7fffffffu;
    // Below 2^-14 the half is subnormal: keep the native round-to-nearest cast.
    if (mag < 0x38800000u) return static_cast<T>(value);
    const uint32_t r = storage_hash(bits, i, n);
    const uint32_t inc = (r & 0x1fffu) < (mag & 0x1fffu) ? 0x2000u : 0u;
    const uint32_t rounded = (mag & 0xffffe000u) + inc;
    return static_cast<T>(as_float((bits & 0x80000000u) | rounded));
  }
};

struct StoreBF16 {
  using T = uint16_t;
  static constexpr int kBytes = 2;
  __host__ __device__ __forceinline__ static float decode(T v) { return as_float(static_cast<uint32_t>(v) << 16); }
  __host__ __device__ __forceinline__ static float load(const T* p) { return decode(*p); }
  __host__ __device__ __forceinline__ static T encode(float g, uint32_t i, uint32_t n) {
    const uint32_t bits = as_uint(g);
    const uint32_t mag = bits & 0x7fffffffu;
    if (mag >= 0x7f800000u) return static_cast<T>((bits >> 16) | (mag > 0x7f800000u ? 0x40u : 0u));
    const uint32_t r = storage_hash(bits, i, n);
    const uint32_t inc = (r & 0xffffu) < (mag & 0xffffu) ? 0x10000u : 0u;
    return static_cast<T>((((bits & 0x80000000u) | (mag & 0xffff0000u)) + inc) >> 16);
  }
};

// Population streams are read once and written once per step. NT selects non-temporal
// accesses for them (no reuse to keep in the caches); the flag and wall-path loads stay cached.
template <bool NT, class T>
__host__ __device__ __forceinline__ T ld(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

template <bool NT, class T>
__host__ __device__ __forceinline__ void st(T* p, T v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

template <class S, bool NT>
__host__ __device__ __forceinline__ float load_pop(const typename S::T* p) {
  return S::decode(ld<NT>(p));
}

// Two consecutive storage elements moved by one access: one dword for 16-bit storage, two
// for f32. The address of an x-shifted pull is only element-aligned; gfx950 serves such
// loads as single unaligned dword accesses.
typedef uint32_t u32_align2 __attribute__((aligned(2)));
typedef uint64_t u64_align4 __attribute__((aligned(4)));

template <class T>
__host__ __device__ __forceinline__ T lo_of(uint32_t v) { return __builtin_bit_cast(T, (uint16_t)(v & 0xffffu)); }
template <class T>
__host__ __device__ __forceinline__ T hi_of(uint32_t v) { return __builtin_bit_cast(T, (uint16_t)(v >> 16)); }

template <class S, bool NT>
__host__ __device__ __forceinline__ void load_pair(const typename S::T* p, float& a, float& b) {
  using T = typename S::T;
  if constexpr (sizeof(T) == 4) {
    const uint64_t v = ld<NT>(reinterpret_cast<const u64_align4*>(p));
    a = S::decode(__builtin_bit_cast(T, (uint32_t)(v & 0xffffffffu)));
    b = S::decode(__builtin_bit_cast(T, (uint32_t)(v >> 32)));
  } else {
    const uint32_t v = ld<NT>(reinterpret_cast<const u32_align2*>(p));
    a = S::decode(lo_of<T>(v));
    b = S::decode(hi_of<T>(v));
  }
}

template <class S, bool NT>
__host__ __device__ __forceinline__ void store_pair(typename S::T* p, typename S::T a, typename S::T b) {
  using T = typename S::T;
  if constexpr (sizeof(T) == 4) {
    const uint64_t v = (uint64_t)__builtin_bit_cast(uint32_t, a) | ((uint64_t)__builtin_bit_cast(uint32_t, b) << 32);
    st<NT>(reinterpret_cast<uint64_t*>(p), v);
  } else {
    const uint32_t v = (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
    st<NT>(reinterpret_cast<uint32_t*>(p), v);
  }
}

}  // namespace pw
