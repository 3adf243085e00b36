// Device helpers shared by the mftx kernels (twiddles, BF16 splitting).
#pragma once
#include "mftx_internal.h"

namespace mftx {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));
typedef float v4f32 __attribute__((ext_vector_type(4)));
typedef uint32_t v2u32 __attribute__((ext_vector_type(2)));
typedef float v2f32 __attribute__((ext_vector_type(2)));

// Global-memory views: pointers that arrive inside descriptors are generic,
// and generic (flat) accesses also count on the LDS/scalar counter, so every
// wait for a scalar load would drain them too.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
// Read-only for the whole launch (descriptor tables): a wave-uniform read
// through it is a scalar load into SGPRs.
template <typename T>
using cptr = __attribute__((address_space(4))) T*;

// --------------------------------------------------------------------------
// scalar helpers
// --------------------------------------------------------------------------

// cis(t) = (cos 2*pi*t, sin 2*pi*t) for t in turns, taking q = 4t (quarter
// turns).  q is reduced in FP64 to a quarter-turn index n and a remainder f in
// [-1/2, 1/2]; the remainder angle (|a| <= pi/4) goes through Taylor series to
// degree 9 (sin) and 10 (cos), whose truncation error is below 3e-8 there.
// The quarter-turn rotation by i^n is branch-free (swap + sign bits).
// |q| must stay below 2^31.
__device__ __forceinline__ float2 cis_quarter(double q) {
  const double n = __builtin_rint(q);
  const float f = (float)(q - n);
  const int quad = (int)n;
  const float a = f * 1.57079632679489662f;
  const float a2 = a * a;
  float s = 2.75573192239858907e-6f;           // 1/9!
  s = __builtin_fmaf(s, a2, -1.98412698412698413e-4f);  // -1/7!
  s = __builtin_fmaf(s, a2, 8.33333333333333333e-3f);   // 1/5!
  s = __builtin_fmaf(s, a2, -1.66666666666666667e-1f);  // -1/3!
  s = a + a * a2 * s;
  float c = -2.75573192239858907e-7f;          // -1/10!
  c = __builtin_fmaf(c, a2, 2.48015873015873016e-5f);   // 1/8!
  c = __builtin_fmaf(c, a2, -1.38888888888888889e-3f);  // -1/6!
  c = __builtin_fmaf(c, a2, 4.16666666666666667e-2f);   // 1/4!
  c = __builtin_fmaf(c, a2, -0.5f);
  c = __builtin_fmaf(c, a2, 1.0f);
  // multiply by i^quad: (c, s), (-s, c), (-c, -s), (s, -c) for quad & 3 = 0..3
  const bool swap = quad & 1;
  const uint32_t sre = (uint32_t)((quad + 1) & 2) << 30;
  const uint32_t sim = (uint32_t)(quad & 2) << 30;
  const float re = __uint_as_float(__float_as_uint(swap ? s : c) ^ sre);
  const float im = __uint_as_float(__float_as_uint(swap ? c : s) ^ sim);
  return make_float2(re, im);
}

__device__ __forceinline__ float2 cis_turns(double t) { return cis_quarter(t * 4.0); }

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(__builtin_fmaf(a.x, b.x, -a.y * b.y), __builtin_fmaf(a.x, b.y, a.y * b.x));
}

// Round-to-nearest-even pack of two floats into two bf16 (v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint32_t pack_bf16(float a, float b) {
  bf16x2_t v = {(__bf16)a, (__bf16)b};
  return __builtin_bit_cast(uint32_t, v);
}
__device__ __forceinline__ float bf16_lo_to_f32(uint32_t u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float bf16_hi_to_f32(uint32_t u) { return __uint_as_float(u & 0xffff0000u); }

// Split two floats into (hi pair, lo pair) with x = hi + lo + O(2^-17 |x|).
// (The empty asm keeps the compiler from looking through the pack, where it
// would convert a to bf16 a second time instead of shifting the pair.)
__device__ __forceinline__ void split2(float a, float b, uint32_t& hi, uint32_t& lo) {
  hi = pack_bf16(a, b);
  asm("" : "+v"(hi));
  lo = pack_bf16(a - bf16_lo_to_f32(hi), b - bf16_hi_to_f32(hi));
}

}  // namespace mftx
