// Device-side helpers shared by the EP kernels (gfx950 / CDNA4, wave64).
//
// Row layout: a feature row of D values is read by one whole wavefront; lane l
// owns the PL = ceil(D/64) consecutive values [l*PL, l*PL+PL).  For D = 64,
// 128, 256 this gives 4/8/16-byte per-lane accesses and one fully coalesced
// 256 B / 512 B / 1 KiB transaction per wave instruction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define EP_WAVE 64

namespace ep {

// ------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy as
// 1, 2, 3", SC'11).  Counter-based: every (key, counter) pair is an independent draw,
// so a kernel can draw for (pass, example, purpose, index) without carrying state.
// ------------------------------------------------------------------------------------
struct u32x4 { uint32_t x, y, z, w; };

__host__ __device__ inline void philox_round(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3,
                                             uint32_t k0, uint32_t k1) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  c0 = hi1 ^ c1 ^ k0;
  c1 = lo1;
  c2 = hi0 ^ c3 ^ k1;
  c3 = lo0;
}

__host__ __device__ inline u32x4 philox4x32(u32x4 ctr, uint32_t k0, uint32_t k1) {
  uint32_t c0 = ctr.x, c1 = ctr.y, c2 = ctr.z, c3 = ctr.w;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    philox_round(c0, c1, c2, c3, k0, k1);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return u32x4{c0, c1, c2, c3};
}

// Purposes (third counter word) keep streams of different uses disjoint.
enum RngPurpose : uint32_t {
  RNG_NEIGHBOURS = 1,
  RNG_NEGATIVES = 2,
  RNG_ROUNDING = 3,
  RNG_NOISE = 4,
  RNG_INIT = 5,
  RNG_WALK = 6,
  RNG_HARD_WALK = 7,
  RNG_FEATURES = 8,  // max_features: purpose RNG_FEATURES | (entry << 4), one stream per node of an example
  RNG_NBR_KEEP = 9,  // max_nodes Probability: one keep draw per edge of the anchor
  RNG_ATT_COUNT = 10,  // Random attention, Probability size: | (block << 8) | (head << 4)
  RNG_ATT_COLS = 11,   // Random attention, columns of each row: | (block << 8) | (head << 4)
};

// Uniform integer in [0, n) from a 32-bit draw (multiply-shift; bias <= n / 2^32).
__host__ __device__ inline uint32_t uniform_below(uint32_t r, uint32_t n) {
  return (uint32_t)(((uint64_t)r * (uint64_t)n) >> 32);
}

// Uniform float in [0, 1) with 24 random bits.
__host__ __device__ inline float uniform01(uint32_t r) {
  return (float)(r >> 8) * (1.0f / 16777216.0f);
}

// 32-bit integer hash (C. Wellons' "lowbias32" finalizer: xorshift-multiply, two rounds).
// Used as a cheap counter-based generator for stochastic-rounding bits, keyed by
// (row, column block, step); Philox stays on the sampling paths.
__host__ __device__ inline uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

__host__ __device__ inline uint32_t sr_key(uint32_t row, uint32_t col4, uint32_t step, uint32_t seed) {
  return hash32(row ^ hash32(col4 * 0x9E3779B9u ^ step * 0x85EBCA77u ^ seed));
}

// ------------------------------------------------------------------------------------
// bfloat16 storage
// ------------------------------------------------------------------------------------
__device__ inline float bf16_to_f32(uint16_t h) {
  return __uint_as_float(((uint32_t)h) << 16);
}

// Round-to-nearest-even (used for the initial copy of a table).
__host__ __device__ inline uint16_t f32_to_bf16_rne(float f) {
#ifdef __HIP_DEVICE_COMPILE__
  uint32_t u = __float_as_uint(f);
#else
  uint32_t u; __builtin_memcpy(&u, &f, 4);
#endif
  if ((u & 0x7f800000u) == 0x7f800000u) {  // inf / nan: keep, quiet nan
    return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  }
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// Stochastic rounding: add 16 uniform random bits below the kept mantissa, then
// truncate.  P(round up) equals the discarded fraction, so E[bf16(x)] = x.
__device__ inline uint16_t f32_to_bf16_sr(float f, uint32_t rnd16) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += (rnd16 & 0xffffu);
  return (uint16_t)(u >> 16);
}

// ------------------------------------------------------------------------------------
// Row slices
// ------------------------------------------------------------------------------------
template <int D>
struct RowShape {
  static constexpr int PL = (D + EP_WAVE - 1) / EP_WAVE;   // values per lane
  static constexpr int ACTIVE = (D + PL - 1) / PL;         // lanes that own values
};

// Load lane's slice of an FP32 row (zero outside the row).
template <int D>
__device__ inline void load_slice(const float* __restrict__ row, int lane, float (&v)[RowShape<D>::PL]) {
  constexpr int PL = RowShape<D>::PL;
  const int base = lane * PL;
  if constexpr (PL > 4 && PL % 4 == 0 && D % PL == 0) {  // wide rows: 16-byte chunks
    if (base < D) {
#pragma unroll
      for (int k = 0; k < PL; k += 4) {
        const float4 t = *reinterpret_cast<const float4*>(row + base + k);
        v[k] = t.x; v[k + 1] = t.y; v[k + 2] = t.z; v[k + 3] = t.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < PL; ++k) v[k] = 0.f;
    }
  } else if constexpr (PL == 4 && D % 4 == 0) {
    if (base < D) {
      const float4 t = *reinterpret_cast<const float4*>(row + base);
      v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    } else { v[0] = v[1] = v[2] = v[3] = 0.f; }
  } else if constexpr (PL == 2 && D % 2 == 0) {
    if (base < D) {
      const float2 t = *reinterpret_cast<const float2*>(row + base);
      v[0] = t.x; v[1] = t.y;
    } else { v[0] = v[1] = 0.f; }
  } else {
#pragma unroll
    for (int j = 0; j < PL; ++j) v[j] = (base + j < D) ? row[base + j] : 0.f;
  }
}

template <int D>
__device__ inline void load_slice(const uint16_t* __restrict__ row, int lane, float (&v)[RowShape<D>::PL]) {
  constexpr int PL = RowShape<D>::PL;
  const int base = lane * PL;
  if constexpr (PL > 4 && PL % 4 == 0 && D % PL == 0) {  // wide rows: 8-byte chunks
    if (base < D) {
#pragma unroll
      for (int k = 0; k < PL; k += 4) {
        const uint2 t = *reinterpret_cast<const uint2*>(row + base + k);
        v[k] = __uint_as_float(t.x << 16); v[k + 1] = __uint_as_float(t.x & 0xffff0000u);
        v[k + 2] = __uint_as_float(t.y << 16); v[k + 3] = __uint_as_float(t.y & 0xffff0000u);
      }
    } else {
#pragma unroll
      for (int k = 0; k < PL; ++k) v[k] = 0.f;
    }
  } else if constexpr (PL == 4 && D % 4 == 0) {
    if (base < D) {
      const uint2 t = *reinterpret_cast<const uint2*>(row + base);
      v[0] = __uint_as_float(t.x << 16); v[1] = __uint_as_float(t.x & 0xffff0000u);
      v[2] = __uint_as_float(t.y << 16); v[3] = __uint_as_float(t.y & 0xffff0000u);
    } else { v[0] = v[1] = v[2] = v[3] = 0.f; }
  } else if constexpr (PL == 2 && D % 2 == 0) {
    if (base < D) {
      const uint32_t t = *reinterpret_cast<const uint32_t*>(row + base);
      v[0] = __uint_as_float(t << 16); v[1] = __uint_as_float(t & 0xffff0000u);
    } else { v[0] = v[1] = 0.f; }
  } else {
#pragma unroll
    for (int j = 0; j < PL; ++j) v[j] = (base + j < D) ? bf16_to_f32(row[base + j]) : 0.f;
  }
}

template <int D>
__device__ inline void store_slice(float* __restrict__ row, int lane, const float (&v)[RowShape<D>::PL]) {
  constexpr int PL = RowShape<D>::PL;
  const int base = lane * PL;
  if constexpr (PL > 4 && PL % 4 == 0 && D % PL == 0) {  // wide rows: 16-byte chunks
    if (base < D) {
#pragma unroll
      for (int k = 0; k < PL; k += 4)
        *reinterpret_cast<float4*>(row + base + k) = make_float4(v[k], v[k + 1], v[k + 2], v[k + 3]);
    }
  } else if constexpr (PL == 4 && D % 4 == 0) {
    if (base < D) *reinterpret_cast<float4*>(row + base) = make_float4(v[0], v[1], v[2], v[3]);
  } else if constexpr (PL == 2 && D % 2 == 0) {
    if (base < D) *reinterpret_cast<float2*>(row + base) = make_float2(v[0], v[1]);
  } else {
#pragma unroll
    for (int j = 0; j < PL; ++j) if (base + j < D) row[base + j] = v[j];
  }
}

// ------------------------------------------------------------------------------------
// FP16 row slices (IEEE binary16, round to nearest even on store), same lane layout as
// the FP32 slices: the values are widened exactly on load.
// ------------------------------------------------------------------------------------
__device__ inline uint16_t f32_to_f16_rne(float x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }
__device__ inline float f16_to_f32(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }

template <int D>
__device__ inline void store_slice_f16(uint16_t* __restrict__ row, int lane, const float (&v)[RowShape<D>::PL],
                                       float mul) {
  constexpr int PL = RowShape<D>::PL;
  const int base = lane * PL;
  if constexpr (PL % 4 == 0 && D % PL == 0) {  // 8-byte chunks
    if (base < D) {
#pragma unroll
      for (int k = 0; k < PL; k += 4) {
        uint2 t;
        t.x = (uint32_t)f32_to_f16_rne(v[k] * mul) | ((uint32_t)f32_to_f16_rne(v[k + 1] * mul) << 16);
        t.y = (uint32_t)f32_to_f16_rne(v[k + 2] * mul) | ((uint32_t)f32_to_f16_rne(v[k + 3] * mul) << 16);
        *reinterpret_cast<uint2*>(row + base + k) = t;
      }
    }
  } else if constexpr (PL == 2 && D % 2 == 0) {
    if (base < D)
      *reinterpret_cast<uint32_t*>(row + base) =
          (uint32_t)f32_to_f16_rne(v[0] * mul) | ((uint32_t)f32_to_f16_rne(v[1] * mul) << 16);
  } else {
#pragma unroll
    for (int j = 0; j < PL; ++j) if (base + j < D) row[base + j] = f32_to_f16_rne(v[j] * mul);
  }
}

// v = mul * row slice (zero outside the row)
template <int D>
__device__ inline void load_slice_f16(const uint16_t* __restrict__ row, int lane, float (&v)[RowShape<D>::PL],
                                      float mul) {
  constexpr int PL = RowShape<D>::PL;
  const int base = lane * PL;
  if constexpr (PL % 4 == 0 && D % PL == 0) {
    if (base < D) {
#pragma unroll
      for (int k = 0; k < PL; k += 4) {
        const uint2 t = *reinterpret_cast<const uint2*>(row + base + k);
        v[k] = mul * f16_to_f32((uint16_t)(t.x & 0xffffu));
        v[k + 1] = mul * f16_to_f32((uint16_t)(t.x >> 16));
        v[k + 2] = mul * f16_to_f32((uint16_t)(t.y & 0xffffu));
        v[k + 3] = mul * f16_to_f32((uint16_t)(t.y >> 16));
      }
    } else {
#pragma unroll
      for (int k = 0; k < PL; ++k) v[k] = 0.f;
    }
  } else if constexpr (PL == 2 && D % 2 == 0) {
    if (base < D) {
      const uint32_t t = *reinterpret_cast<const uint32_t*>(row + base);
      v[0] = mul * f16_to_f32((uint16_t)(t & 0xffffu));
      v[1] = mul * f16_to_f32((uint16_t)(t >> 16));
    } else { v[0] = v[1] = 0.f; }
  } else {
#pragma unroll
    for (int j = 0; j < PL; ++j) v[j] = (base + j < D) ? mul * f16_to_f32(row[base + j]) : 0.f;
  }
}

// ------------------------------------------------------------------------------------
// Wave reductions (wave64)
// ------------------------------------------------------------------------------------
__device__ inline float wave_max(float x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = fmaxf(x, __shfl_xor(x, off, EP_WAVE));
  return x;
}

__device__ inline float wave_sum(float x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, EP_WAVE);
  return x;
}

// DPP reductions: two quad permutes, the half-row and row mirrors (every lane then holds
// its 16-lane row's total), then the row broadcasts of lanes 15 and 31 (rows 1 and 3,
// then rows 2 and 3), which leave the wave's total in lane 63.  No LDS round trips, where
// a shuffle butterfly takes six.  Lanes a DPP step leaves unwritten read `old` = 0.
namespace dpp {
constexpr int kQuad1032 = 0xB1;      // quad_perm [1, 0, 3, 2]
constexpr int kQuad2301 = 0x4E;      // quad_perm [2, 3, 0, 1]
constexpr int kRowHalfMirror = 0x141;
constexpr int kRowMirror = 0x140;
constexpr int kRowBcast15 = 0x142;
constexpr int kRowBcast31 = 0x143;
template <int CTRL, int ROW_MASK>
__device__ inline float mov(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROW_MASK, 0xf, true));
}
template <int CTRL, int ROW_MASK>
__device__ inline uint32_t movu(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROW_MASK, 0xf, true);
}
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
}  // namespace dpp

// Inclusive prefix sum over the wave: shifts by 1, 2, 4 and 8 within each 16-lane row
// (lanes shifted in from outside the row read 0), then the row broadcasts of lanes 15 and
// 31 carry the rows' totals up.  Lane 63 ends with the wave's total.
__device__ inline uint32_t wave_inclusive_scan_u32(uint32_t x) {
  x += dpp::movu<dpp::kRowShr1, 0xf>(x);
  x += dpp::movu<dpp::kRowShr2, 0xf>(x);
  x += dpp::movu<dpp::kRowShr4, 0xf>(x);
  x += dpp::movu<dpp::kRowShr8, 0xf>(x);
  x += dpp::movu<dpp::kRowBcast15, 0xa>(x);
  x += dpp::movu<dpp::kRowBcast31, 0xc>(x);
  return x;
}

// Sum over the wave, returned to every lane (the order differs from wave_sum's).
__device__ inline float wave_sum_dpp(float x) {
  x += dpp::mov<dpp::kQuad1032, 0xf>(x);
  x += dpp::mov<dpp::kQuad2301, 0xf>(x);
  x += dpp::mov<dpp::kRowHalfMirror, 0xf>(x);
  x += dpp::mov<dpp::kRowMirror, 0xf>(x);
  x += dpp::mov<dpp::kRowBcast15, 0xa>(x);
  x += dpp::mov<dpp::kRowBcast31, 0xc>(x);
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
}

// Maximum over the wave of values >= 0 (the zero of unwritten lanes is then neutral),
// returned to every lane; a maximum is exact, so it equals wave_max's.
__device__ inline float wave_max_nonneg(float x) {
  x = fmaxf(x, dpp::mov<dpp::kQuad1032, 0xf>(x));
  x = fmaxf(x, dpp::mov<dpp::kQuad2301, 0xf>(x));
  x = fmaxf(x, dpp::mov<dpp::kRowHalfMirror, 0xf>(x));
  x = fmaxf(x, dpp::mov<dpp::kRowMirror, 0xf>(x));
  x = fmaxf(x, dpp::mov<dpp::kRowBcast15, 0xa>(x));
  x = fmaxf(x, dpp::mov<dpp::kRowBcast31, 0xc>(x));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
}

__device__ inline double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, EP_WAVE);
  return x;
}

__device__ inline uint32_t wave_sum_u32(uint32_t x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, EP_WAVE);
  return x;
}

__device__ inline bool wave_any(bool p) { return __ballot(p) != 0ull; }

__device__ inline uint32_t lane_id() { return __lane_id(); }

}  // namespace ep
