// lattice.h -- D3Q19 constants and population storage codecs for gfx950 kernels.
//
// Velocity order and weights follow src/lbm/lattice.ts:2-52 (opposite pairs adjacent,
// i odd <-> i + 1); the f16 codec keeps the format of src/lbm/shaders/prelude-f16.wgsl:5-22
// (deviation scaled by 2^15, proportional stochastic rounding of the dropped bits). bf16
// storage is this library's own format (deviation unscaled, the same rounding).
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

// Index of the velocity (x, y, z) (every component in {-1, 0, 1}; 19 of the 27 exist).
__host__ __device__ constexpr int dir_of(int x, int y, int z) {
  for (int i = 0; i < 19; ++i)
    if (cx(i) == x && cy(i) == y && cz(i) == z) return i;
  return -1;
}
// Specular image of velocity i in a free-slip side: its y (MY) and/or z (MZ) component negated.
__host__ __device__ constexpr int mirror_dir(int i, bool my, bool mz) {
  return dir_of(cx(i), my ? -cy(i) : cy(i), mz ? -cz(i) : cz(i));
}

__host__ __device__ __forceinline__ float as_float(uint32_t u) { return __builtin_bit_cast(float, u); }
__host__ __device__ __forceinline__ uint32_t as_uint(float f) { return __builtin_bit_cast(uint32_t, f); }

enum : uint32_t { kFluid = 0, kSolid = 1, kSupport = 2, kFloor = 3, kEq = 4, kTypeMask = 7 };
constexpr uint32_t kNoQ = 0xffffffffu;

__host__ __device__ __forceinline__ bool is_wall(uint32_t kind) { return kind - 1u < 3u; }
__host__ __device__ __forceinline__ bool has_wall(uint32_t flag, int i) { return (flag >> (3 + i)) & 1u; }

// lowbias32 (public-domain hash-prospector constants; the same mixer as prelude-f16.wgsl:8-13).
__host__ __device__ __forceinline__ uint32_t hash32(uint32_t r) {
  r = (r ^ (r >> 16)) * 0x7feb352du;
  r = (r ^ (r >> 15)) * 0x846ca68bu;
  return r ^ (r >> 16);
}

// Random words for the stochastic rounding of one cell's 19 stores at one step: one hash per
// cell and step, then a Weyl sequence over the directions, so every direction's word is
// uniformly distributed and a cell draws fresh words each step.
__host__ __device__ __forceinline__ uint32_t cell_random(uint32_t n, uint32_t seed) {
  return hash32(n ^ (seed * 0x85ebca6bu));
}
__host__ __device__ __forceinline__ uint32_t direction_random(uint32_t h, int i) {
  return h + (uint32_t)i * 0x9e3779b9u;
}

// Seed of the initial (reset) state; step k uses seed k.
constexpr uint32_t kInitSeed = 0xffffffffu;

// Stochastic rounding of an f32 to the format's dropped mantissa bits: round away from zero
// with probability equal to the dropped fraction, chosen by the random word r. On the device
// this is one v_cvt_sr_* instruction; the host pass (CPU checks of this kernel source) adds
// r's low bits to the dropped bits and truncates, which has the same distribution.
__host__ __device__ __forceinline__ _Float16 round_f16_sr(float v, uint32_t r) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
  half2_t h = __builtin_nondeterministic_value(h);
  h = __builtin_amdgcn_cvt_sr_f16_f32(h, v, r, false);
  return h[0];
#else
  const uint32_t bits = as_uint(v), mag = bits & 0x7fffffffu;
  if (mag >= 0x7f800000u) return static_cast<_Float16>(v);
  if (mag < 0x38800000u) {  // f16 subnormal range: steps of 2^-24
    const float steps = as_float(mag) * 16777216.0f;
    const float low = __builtin_floorf(steps);
    const bool up = (float)(r & 0xffffffu) * (1.0f / 16777216.0f) < steps - low;
    const float m = (low + (up ? 1.0f : 0.0f)) * (1.0f / 16777216.0f);
    return static_cast<_Float16>((bits >> 31) ? -m : m);
  }
  const uint32_t rounded = (mag + (r & 0x1fffu)) & 0xffffe000u;
  return static_cast<_Float16>(as_float((bits & 0x80000000u) | rounded));
#endif
}

__host__ __device__ __forceinline__ uint16_t round_bf16_sr(float v, uint32_t r) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef short short2_t __attribute__((ext_vector_type(2)));
  short2_t h = __builtin_nondeterministic_value(h);
  h = __builtin_amdgcn_cvt_sr_bf16_f32(h, v, r, false);
  return static_cast<uint16_t>(h[0]);
#else
  const uint32_t bits = as_uint(v), mag = bits & 0x7fffffffu;
  if (mag >= 0x7f800000u) return static_cast<uint16_t>((bits >> 16) | (mag > 0x7f800000u ? 0x40u : 0u));
  const uint32_t rounded = (mag + (r & 0xffffu)) & 0xffff0000u;
  return static_cast<uint16_t>(((bits & 0x80000000u) | rounded) >> 16);
#endif
}

// The random words of a codec's stores: cell_word(n, seed) once per cell and step, then
// dir_word(h, i) per direction.
struct HashedWords {
  __host__ __device__ __forceinline__ static uint32_t cell_word(uint32_t n, uint32_t seed) { return cell_random(n, seed); }
  __host__ __device__ __forceinline__ static uint32_t dir_word(uint32_t h, int i) { return direction_random(h, i); }
};

// Storage codecs: decode() returns the f32 deviation g = f - w; encode() takes g and the
// store's random word.
struct StoreF32 : HashedWords {
  using T = float;
  static constexpr int kBytes = 4;
  static constexpr int kShift = 2;
  static constexpr int kScaleLog2 = 0;  // stored value = g * 2^kScaleLog2
  // The REG collision in its regrouped form (step_kernels.h collide_reg_regrouped) rather than
  // through the per-direction equilibrium.
  static constexpr bool kRegroupedReg = false;
  __host__ __device__ __forceinline__ static float decode(T v) { return v; }
  __host__ __device__ __forceinline__ static float load(const T* p) { return *p; }
  __host__ __device__ __forceinline__ static T encode(float g, uint32_t) { return g; }
  // The same from a value already multiplied by 2^kScaleLog2.
  __host__ __device__ __forceinline__ static T encode_scaled(float gs, uint32_t r) { return encode(gs, r); }
};

// f16 deviations scaled by 2^15 (prelude-f16.wgsl:5-22 fixes the format; the rounding's
// random source is this library's own).
struct StoreF16 : HashedWords {
  using T = _Float16;
  static constexpr int kBytes = 2;
  static constexpr int kShift = 1;
  static constexpr int kScaleLog2 = 15;
  static constexpr bool kRegroupedReg = true;
  static constexpr float kScale = 32768.0f;
  // An fma with a zero addend (one v_fma_mix_f32 from the half operand) rather than a conversion
  // and a multiply; the scale is a power of two, so the result is the same exact product.
  __host__ __device__ __forceinline__ static float decode(T v) {
    return __builtin_fmaf(static_cast<float>(v), 1.0f / kScale, 0.0f);
  }
  __host__ __device__ __forceinline__ static float load(const T* p) { return decode(*p); }
  __host__ __device__ __forceinline__ static T encode(float g, uint32_t r) { return round_f16_sr(g * kScale, r); }
  __host__ __device__ __forceinline__ static T encode_scaled(float gs, uint32_t r) { return round_f16_sr(gs, r); }
  // Two already-scaled values rounded into the low and high halves of one register; `old`
  // supplies the register (both halves are overwritten).
  __host__ __device__ __forceinline__ static uint32_t encode_pair_scaled(uint32_t old, float as, uint32_t ra, float bs,
                                                                         uint32_t rb) {
#if defined(__HIP_DEVICE_COMPILE__)
    typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
    half2_t h = __builtin_bit_cast(half2_t, old);
    h = __builtin_amdgcn_cvt_sr_f16_f32(h, as, ra, false);
    h = __builtin_amdgcn_cvt_sr_f16_f32(h, bs, rb, true);
    return __builtin_bit_cast(uint32_t, h);
#else
    (void)old;
    return (uint32_t)__builtin_bit_cast(uint16_t, encode_scaled(as, ra)) |
           ((uint32_t)__builtin_bit_cast(uint16_t, encode_scaled(bs, rb)) << 16);
#endif
  }
  __host__ __device__ __forceinline__ static uint32_t encode_pair(float a, uint32_t ra, float b, uint32_t rb) {
    return encode_pair_scaled(0u, a * kScale, ra, b * kScale, rb);
  }
};

// The f16 format with the reference's rounding (prelude-f16.wgsl:7-22): the random word of a
// store is lowbias32 of the value's f32 bits, the cell index n and the direction i, and the 13
// dropped bits round up when the word's low 13 bits are below them; values under the smallest
// normal half convert to nearest. Written out in integer operations (about 20 per value, where
// the hardware rounding takes one); selected per solver to compare stored bits with the reference.
__host__ __device__ __forceinline__ _Float16 round_f16_reference(float v, uint32_t cell_dir) {
  const uint32_t bits = as_uint(v), mag = bits & 0x7fffffffu;
  if (mag < 0x38800000u) return static_cast<_Float16>(v);
  const uint32_t random = hash32(bits ^ cell_dir);
  const uint32_t rounded = (mag & 0xffffe000u) + ((random & 0x1fffu) < (mag & 0x1fffu) ? 0x2000u : 0u);
  return static_cast<_Float16>(as_float((bits & 0x80000000u) | rounded));
}

struct StoreF16Reference : StoreF16 {
  // The same arithmetic as f32 storage, so its stored bits compare with the reference's encoding
  // of the f32 solver's values.
  static constexpr bool kRegroupedReg = false;
  // The words carry the cell and direction only (prelude-f16.wgsl:10); the value's bits join at
  // the rounding, so the seed (the step) takes no part.
  __host__ __device__ __forceinline__ static uint32_t cell_word(uint32_t n, uint32_t) { return n * 0x9e3779b9u; }
  __host__ __device__ __forceinline__ static uint32_t dir_word(uint32_t h, int i) { return h ^ ((uint32_t)i * 0x85ebca6bu); }
  __host__ __device__ __forceinline__ static T encode(float g, uint32_t r) { return round_f16_reference(g * kScale, r); }
  __host__ __device__ __forceinline__ static T encode_scaled(float gs, uint32_t r) { return round_f16_reference(gs, r); }
  __host__ __device__ __forceinline__ static uint32_t encode_pair_scaled(uint32_t, float as, uint32_t ra, float bs,
                                                                         uint32_t rb) {
    return (uint32_t)__builtin_bit_cast(uint16_t, encode_scaled(as, ra)) |
           ((uint32_t)__builtin_bit_cast(uint16_t, encode_scaled(bs, rb)) << 16);
  }
  __host__ __device__ __forceinline__ static uint32_t encode_pair(float a, uint32_t ra, float b, uint32_t rb) {
    return encode_pair_scaled(0u, a * kScale, ra, b * kScale, rb);
  }
};

// bf16 deviations, unscaled (this library's own format).
struct StoreBF16 : HashedWords {
  using T = uint16_t;
  static constexpr int kBytes = 2;
  static constexpr int kShift = 1;
  static constexpr int kScaleLog2 = 0;
  static constexpr bool kRegroupedReg = true;
  __host__ __device__ __forceinline__ static float decode(T v) { return as_float(static_cast<uint32_t>(v) << 16); }
  __host__ __device__ __forceinline__ static float load(const T* p) { return decode(*p); }
  __host__ __device__ __forceinline__ static T encode(float g, uint32_t r) { return round_bf16_sr(g, r); }
  __host__ __device__ __forceinline__ static T encode_scaled(float gs, uint32_t r) { return encode(gs, r); }
  __host__ __device__ __forceinline__ static uint32_t encode_pair_scaled(uint32_t old, float a, uint32_t ra, float b,
                                                                         uint32_t rb) {
#if defined(__HIP_DEVICE_COMPILE__)
    typedef short short2_t __attribute__((ext_vector_type(2)));
    short2_t h = __builtin_bit_cast(short2_t, old);
    h = __builtin_amdgcn_cvt_sr_bf16_f32(h, a, ra, false);
    h = __builtin_amdgcn_cvt_sr_bf16_f32(h, b, rb, true);
    return __builtin_bit_cast(uint32_t, h);
#else
    (void)old;
    return (uint32_t)encode(a, ra) | ((uint32_t)encode(b, rb) << 16);
#endif
  }
  __host__ __device__ __forceinline__ static uint32_t encode_pair(float a, uint32_t ra, float b, uint32_t rb) {
#if defined(__HIP_DEVICE_COMPILE__)
    typedef short short2_t __attribute__((ext_vector_type(2)));
    short2_t h = __builtin_nondeterministic_value(h);
    h = __builtin_amdgcn_cvt_sr_bf16_f32(h, a, ra, false);
    h = __builtin_amdgcn_cvt_sr_bf16_f32(h, b, rb, true);
    return __builtin_bit_cast(uint32_t, h);
#else
    return (uint32_t)encode(a, ra) | ((uint32_t)encode(b, rb) << 16);
#endif
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

}  // namespace pw
