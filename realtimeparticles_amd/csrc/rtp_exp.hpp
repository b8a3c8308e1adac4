// The exponential of the Clouds model (clouds.cl:67-96, 942-952).
//
// OpenCL leaves exp's precision to the vendor (within a few ulp).  The Clouds model
// evaluates it in float (externalHeatSource, the wind of cld_updatePosition) and in double
// (saturationVaporDensity, whose literals are double), so bit-exact parity with the CPU
// oracle needs one fixed algorithm evaluated the same way on both sides.  This is
// Cody-Waite argument reduction followed by a Taylor polynomial (a published textbook
// method; e.g. J.-M. Muller, "Elementary Functions", ch. 11):
//   k = rint(x log2 e),  r = (x - k ln2_hi) - k ln2_lo   (ln2_hi has 31 significant bits,
//                                                         so k ln2_hi is exact)
//   e^r = sum_{i=0}^{13} r^i / i!   in Horner form with fused multiply-adds, |r| <= ln2/2
//   e^x = (e^r 2^k1) 2^k2,  k1 = k/2, k2 = k - k1   (two exact power-of-two factors, so a
//                                                     subnormal result rounds once)
// Every operation is an IEEE double operation (fma, rint, multiplication) with one correctly
// rounded result on the CPU and on the GPU, so both give the same bits; the result is within
// 1 ulp of the correctly rounded e^x.  The float form rounds the double result once.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rtp {

__host__ __device__ inline double exp_pow2(int k)  // 2^k, -1022 <= k <= 1023
{
  return __builtin_bit_cast(double, (uint64_t)(k + 1023) << 52);
}

__host__ __device__ inline double exp64(double x)
{
  if (x != x)
    return x + x;
  if (x > 709.782712893384)
    return __builtin_huge_val();
  if (x < -745.1332191019412)
    return 0.0;
  const double k = __builtin_rint(x * 0x1.71547652b82fep+0);
  double r = __builtin_fma(-k, 0x1.62e42fec00000p-1, x);
  r = __builtin_fma(-k, 0x1.d1cf79abc9e3bp-32, r);
  // 1/i!, i = 13 .. 0, rounded to double
  double p = 0x1.6124613a86d09p-33;
  p = __builtin_fma(p, r, 0x1.1eed8eff8d898p-29);
  p = __builtin_fma(p, r, 0x1.ae64567f544e4p-26);
  p = __builtin_fma(p, r, 0x1.27e4fb7789f5cp-22);
  p = __builtin_fma(p, r, 0x1.71de3a556c734p-19);
  p = __builtin_fma(p, r, 0x1.a01a01a01a01ap-16);
  p = __builtin_fma(p, r, 0x1.a01a01a01a01ap-13);
  p = __builtin_fma(p, r, 0x1.6c16c16c16c17p-10);
  p = __builtin_fma(p, r, 0x1.1111111111111p-7);
  p = __builtin_fma(p, r, 0x1.5555555555555p-5);
  p = __builtin_fma(p, r, 0x1.5555555555555p-3);
  p = __builtin_fma(p, r, 0x1.0000000000000p-1);
  p = __builtin_fma(p, r, 1.0);
  p = __builtin_fma(p, r, 1.0);
  const int ki = (int)k;
  const int k1 = ki / 2;
  const int k2 = ki - k1;
  return (p * exp_pow2(k1)) * exp_pow2(k2);
}

__host__ __device__ inline float exp32(float x) { return (float)exp64((double)x); }

}  // namespace rtp
