// Correctly rounded fp32 square root and division in short forms for the value ranges a
// neighbour walk meets.
//
// With -fhip-fp32-correctly-rounded-divide-sqrt the compiler lowers sqrtf and '/' to
// general sequences that handle every input: a square root pre-scales arguments below
// 2^-96 by 2^32 and patches +-0 / +inf afterwards (v_cmp_class); a division range-reduces
// both operands with v_div_scale, post-scales in v_div_fmas and patches special cases in
// v_div_fixup.  Inside the ranges below those steps are identities, and what is left is
// the same instruction sequence written out here, so these functions return the same bits
// as sqrtf / '/' (checked on the GPU against the compiled general forms: every float for
// the square root, every pair of significands for the division; rtp_selfcheck_*):
//
//   square root (result s0 = v_sqrt_f32(x) corrected by one ulp either way from the
//   signs of the residuals x - s*(s -+ ulp)), for x = +-0, x >= 2^-96, +inf or NaN;
//
//   division n / d for d > 0 (Newton-Raphson reciprocal refinement followed by two
//   remainder corrections, the FMA division of Markstein, IBM J. Res. Dev. 34(1), 1990):
//     y0 = rcp(d), y = y0 + (1 - d*y0)*y0,
//     q0 = n*y, q1 = q0 + (n - d*q0)*y, q = q1 + (n - d*q1)*y,
//   the sign taken from n as v_div_fixup does, for d in [2^-40, 2^40] and n = +-0 or
//   |n| in [2^-90, 2^50] with |n|/d in [2^-120, 2^90]: operands and quotient normal, the
//   remainders multiples of 2^(e(n) - 46) >= 2^-136 and so exact even where they are
//   denormal, and exponent differences below the 96 at which v_div_scale starts scaling.
//
// The reciprocal y depends on d alone, so divisions by one divisor share it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtp {

constexpr uint32_t kSqrtShortMin = 0x0f800000u;  // 2^-96
constexpr uint32_t kDivNumMin = 0x12800000u;     // 2^-90
constexpr float kDivDenMin = 0x1p-40f, kDivDenMax = 0x1p40f;

// Marks a wave-uniform branch taken only when some lane needs a general form: the empty
// volatile statement keeps the compiler from if-converting (always executing) the block.
__device__ __forceinline__ void rare_path()
{
  asm volatile("");
}

// sqrt_short(x) == sqrtf(x) for x = +-0, x >= 2^-96, +inf, NaN.
__device__ __forceinline__ float sqrt_short(float x)
{
  float s = __builtin_amdgcn_sqrtf(x);
  const uint32_t sb = __float_as_uint(s);
  const float sd = __uint_as_float(sb - 1u), su = __uint_as_float(sb + 1u);
  const float rd = __builtin_fmaf(-sd, s, x), ru = __builtin_fmaf(-su, s, x);
  s = rd <= 0.0f ? sd : s;
  s = ru > 0.0f ? su : s;
  return s;
}

// sqrtf(x) for x >= 0: the short form, and the general one for lanes with 0 < x < 2^-96.
__device__ __forceinline__ float sqrt_cr(float x)
{
  float s = sqrt_short(x);
  const bool general = __float_as_uint(x) - 1u < kSqrtShortMin - 1u;
  if (__builtin_amdgcn_ballot_w64(general) != 0) {
    rare_path();
    if (general)
      s = __builtin_sqrtf(x);
  }
  return s;
}

// Refined reciprocal of a divisor d in [2^-40, 2^40].
__device__ __forceinline__ float rcp_refined(float d)
{
  const float y0 = __builtin_amdgcn_rcpf(d);
  const float e = __builtin_fmaf(-d, y0, 1.0f);
  return __builtin_fmaf(e, y0, y0);
}

// n / d for d in [2^-40, 2^40] (y = rcp_refined(d)) and n = +-0 or |n| in [2^-90, 2^50]
// with |n|/d in [2^-120, 2^90].
__device__ __forceinline__ float div_short(float n, float d, float y)
{
  const float q0 = n * y;
  const float r0 = __builtin_fmaf(-d, q0, n);
  const float q1 = __builtin_fmaf(r0, y, q0);
  const float r1 = __builtin_fmaf(-d, q1, n);
  const float q = __builtin_fmaf(r1, y, q1);
  return __uint_as_float((__float_as_uint(q) & 0x7fffffffu) | (__float_as_uint(n) & 0x80000000u));
}

// (bits << 1) - 2 orders |n| as unsigned with +-0 mapped to the top: the smallest of three
// keys is below kDivNumMin's key exactly when one of the numerators is a nonzero value
// under 2^-90.
__device__ __forceinline__ uint32_t num_key(float n)
{
  return (__float_as_uint(n) << 1) - 2u;
}

__device__ __forceinline__ bool tiny_numerator3(float a, float b, float c)
{
  const uint32_t k = min(min(num_key(a), num_key(b)), num_key(c));
  return k < (kDivNumMin << 1) - 2u;
}

}  // namespace rtp
