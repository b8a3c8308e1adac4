// On-device checks of the short correctly rounded forms of rtp_crmath.hpp against the
// compiler's general sequences (sqrtf and '/' under -fhip-fp32-correctly-rounded-divide-sqrt),
// which the neighbour passes rely on returning the same bits.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "rtp_crmath.hpp"
#include "rtp_host.hpp"

namespace rtp {
namespace {

// Results are equal when their bits are, or when both are NaN.
__device__ __forceinline__ bool same(float a, float b)
{
  return __float_as_uint(a) == __float_as_uint(b) || (a != a && b != b);
}

// The general forms, kept out of line so the compiler cannot fold them into the short ones.
__device__ __noinline__ float sqrt_general(float x)
{
  return __builtin_sqrtf(x);
}

__device__ __noinline__ float div_general(float n, float d)
{
  return n / d;
}

// Every bit pattern x in [lo, hi): sqrt_cr(x) against sqrtf(x).  first[0] keeps the smallest
// mismatching pattern.
__global__ __launch_bounds__(256) void k_check_sqrt(uint32_t lo, uint32_t hi,
                                                    unsigned long long* bad, uint32_t* first)
{
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t nbad = 0;
  for (uint64_t b = (uint64_t)lo + blockIdx.x * blockDim.x + threadIdx.x; b < hi; b += stride) {
    const float x = __uint_as_float((uint32_t)b);
    if (!same(sqrt_cr(x), sqrt_general(x))) {
      ++nbad;
      atomicMin(first, (uint32_t)b);
    }
  }
  if (nbad)
    atomicAdd(bad, (unsigned long long)nbad);
}

// Divisors d = 2^dExp * (1 + s/2^23) for s in [dSigLo, dSigHi), numerators
// n = +-2^nExp * (1 + t/2^23) for t = 0, tStep, 2 tStep, ... < 2^23 (both signs) and +-0:
// div_short(n, d, rcp_refined(d)) against n / d.
__global__ __launch_bounds__(256) void k_check_div(int32_t nExp, int32_t dExp, uint32_t dSigLo,
                                                   uint32_t dSigHi, uint32_t tStep,
                                                   unsigned long long* bad, uint32_t* first)
{
  const uint32_t s = dSigLo + blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= dSigHi)
    return;
  const float d = __uint_as_float(((uint32_t)(dExp + 127) << 23) | s);
  const float y = rcp_refined(d);
  const uint32_t nBase = (uint32_t)(nExp + 127) << 23;
  uint32_t nbad = 0;
  for (uint32_t t = 0; t < (1u << 23); t += tStep) {
    const float n = __uint_as_float(nBase | t);
    if (!same(div_short(n, d, y), div_general(n, d)))
      ++nbad;
    if (!same(div_short(-n, d, y), div_general(-n, d)))
      ++nbad;
  }
  if (!same(div_short(0.0f, d, y), div_general(0.0f, d)))
    ++nbad;
  if (!same(div_short(-0.0f, d, y), div_general(-0.0f, d)))
    ++nbad;
  if (nbad) {
    atomicAdd(bad, (unsigned long long)nbad);
    atomicMin(first, s);
  }
}

// Mismatch count (64-bit) and smallest failing input, zeroed / set to ~0 on the device.
struct CheckBuf {
  unsigned long long* bad = nullptr;
  uint32_t* first = nullptr;
  ~CheckBuf()
  {
    if (bad)
      (void)hipFree(bad);
  }
  int init()
  {
    RTP_HIP_TRY(hipMalloc(reinterpret_cast<void**>(&bad), 16));
    first = reinterpret_cast<uint32_t*>(bad + 1);
    const uint32_t init[4] = { 0u, 0u, 0xffffffffu, 0u };
    RTP_HIP_TRY(hipMemcpy(bad, init, 16, hipMemcpyHostToDevice));
    return RTP_OK;
  }
  int fetch(uint64_t* mismatches, uint32_t* firstBad)
  {
    RTP_HIP_TRY(hipGetLastError());
    RTP_HIP_TRY(hipDeviceSynchronize());
    uint32_t out[4];
    RTP_HIP_TRY(hipMemcpy(out, bad, 16, hipMemcpyDeviceToHost));
    *mismatches = (uint64_t)out[0] | ((uint64_t)out[1] << 32);
    if (firstBad)
      *firstBad = out[2];
    return RTP_OK;
  }
};

}  // namespace
}  // namespace rtp

using namespace rtp;

extern "C" int rtp_selfcheck_sqrt(int device, uint32_t lo, uint32_t hi, uint64_t* mismatches,
                                  uint32_t* firstBad)
{
  RTP_REQUIRE(mismatches != nullptr && lo <= hi, "rtp_selfcheck_sqrt: bad arguments");
  const int st = check_device(device);
  if (st != RTP_OK)
    return st;
  CheckBuf b;
  const int ok = b.init();
  if (ok != RTP_OK)
    return ok;
  if (hi > lo)
    k_check_sqrt<<<4096, 256>>>(lo, hi, b.bad, b.first);
  return b.fetch(mismatches, firstBad);
}

extern "C" int rtp_selfcheck_division(int device, int32_t nExp, int32_t dExp, uint32_t dSigLo,
                                      uint32_t dSigHi, uint32_t nSigStep, uint64_t* mismatches,
                                      uint32_t* firstBad)
{
  RTP_REQUIRE(mismatches != nullptr && nSigStep >= 1 && dSigLo <= dSigHi && dSigHi <= (1u << 23) &&
                  nExp >= -126 && nExp <= 127 && dExp >= -126 && dExp <= 127,
              "rtp_selfcheck_division: bad arguments");
  const int st = check_device(device);
  if (st != RTP_OK)
    return st;
  CheckBuf b;
  const int ok = b.init();
  if (ok != RTP_OK)
    return ok;
  const uint32_t count = dSigHi - dSigLo;
  if (count)
    k_check_div<<<(count + 255u) / 256u, 256>>>(nExp, dExp, dSigLo, dSigHi, nSigStep, b.bad, b.first);
  return b.fetch(mismatches, firstBad);
}
