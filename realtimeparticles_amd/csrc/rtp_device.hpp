// Device-side helpers shared by the gfx950 kernels.
//
// Floating-point contract of every kernel in this library (build flags: -ffp-contract=off,
// correctly rounded fp32 division and square root, IEEE denormals):
//   * each OpenCL expression of the reference is evaluated in its written order;
//   * dot(a,b)        = ((a.x*b.x + a.y*b.y) + a.z*b.z)   (the w lanes are 0, so the
//                       reference's "+ a.w*b.w" adds +0 and changes nothing);
//   * length(v)       = sqrtf(dot(v,v));  normalize(v) = v / length(v), 0 for v == 0;
//   * pow(x, n)       = ((x*x)*x)... with n-1 products for the integral exponents the
//                       reference uses (3 and 2 in sph.cl, artPressureExp in fluids.cl:56);
//   * clamp(x,lo,hi)  = fminf(fmaxf(x, lo), hi).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "rtp_internal.hpp"

namespace rtp {

constexpr float kFloatEps = 0.00000001f;  // FLOAT_EPS, define.cl:6

__device__ __forceinline__ float clampf(float x, float lo, float hi)
{
  return fminf(fmaxf(x, lo), hi);
}

__device__ __forceinline__ float pown(float x, uint32_t n)
{
  if (n == 0)
    return 1.0f;
  float r = x;
  for (uint32_t k = 1; k < n; ++k)
    r = r * x;
  return r;
}

__device__ __forceinline__ float cube(float x) { return (x * x) * x; }

// getCell3DIndexFromPos, grid.cl:14-23.
__device__ __forceinline__ void cell3d(float4 p, const GridConsts& g, uint32_t& cx, uint32_t& cy,
                                       uint32_t& cz)
{
  cx = (uint32_t)floorf((clampf(p.x, -g.absWall[0], g.absWall[0]) + g.absWall[0]) / g.cellSize);
  cy = (uint32_t)floorf((clampf(p.y, -g.absWall[1], g.absWall[1]) + g.absWall[1]) / g.cellSize);
  cz = (uint32_t)floorf((clampf(p.z, -g.absWall[2], g.absWall[2]) + g.absWall[2]) / g.cellSize);
}

// getCell1DIndexFromPos, grid.cl:28-38.
__device__ __forceinline__ uint32_t cell1d(float4 p, const GridConsts& g)
{
  uint32_t cx, cy, cz;
  cell3d(p, g, cx, cy, cz);
  const uint32_t ry = (uint32_t)g.res[1], rz = (uint32_t)g.res[2];
  return cx * rz * ry + cy * rz + cz;
}

// The thread-order key of the cell of q under RTP_LAUNCH_BRICKS: (y, z) tile of kBrick x kBrick
// cells (cell coordinates 0..res), then x, then the cell's place in its tile, so that the
// order runs along x through one pencil of tiles after another.
__device__ __forceinline__ uint32_t brick_key(float4 q, const GridConsts& g)
{
  uint32_t cx, cy, cz;
  cell3d(q, g, cx, cy, cz);
  const uint32_t nTz = ((uint32_t)g.res[2] + kBrick) / kBrick;
  const uint32_t tile = (cy / kBrick) * nTz + cz / kBrick;
  return ((tile * ((uint32_t)g.res[0] + 1u) + cx) * kBrick + cy % kBrick) * kBrick + cz % kBrick;
}

// (c + o + res) % res for c in [0, res] and o in {-1, 0, 1} (fluids.cl:107).
__device__ __forceinline__ int32_t wrap(int32_t v, int32_t res)
{
  v = v < 0 ? v + res : v;
  return v >= res ? v - res : v;
}

// Logical block index that hands every XCD a contiguous run of particle blocks, so the
// neighbour data of consecutive (spatially adjacent) blocks shares one L2.  Workgroups are
// dispatched to the 8 XCDs round-robin by blockIdx.
__device__ __forceinline__ uint32_t logical_block(int remap)
{
  const uint32_t b = blockIdx.x;
  if (!remap)
    return b;
  const uint32_t nb = gridDim.x;
  const uint32_t q = nb / 8u, r = nb % 8u;
  const uint32_t x = b % 8u, slot = b / 8u;
  return x * q + (x < r ? x : r) + slot;
}

__device__ __forceinline__ float4 ld4(const float4* p) { return *p; }

// ------------------------------------------------------------------- quads of lanes
// Walks that give a particle a quad of four lanes (lanes 4q .. 4q + 3 of a wave) exchange
// their terms within the quad and queue survivors in LDS that the quad's other lanes read.

// Orders the wave's LDS accesses on both sides (a quad's lanes read what the others wrote).
__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// v of lane K of the quad.
template <int K>
__device__ __forceinline__ float quad_lane(float v)
{
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                               K * 0x55, 0xF, 0xF, false));
}

// ((((s + t0) + t1) + t2) + t3) for t_k the t of quad lane k: four terms added to a sum
// in the order of the quad's lanes.
__device__ __forceinline__ float quad_add(float s, float t)
{
  s = s + quad_lane<0>(t);
  s = s + quad_lane<1>(t);
  s = s + quad_lane<2>(t);
  return s + quad_lane<3>(t);
}

// The quad's four bits of a wave ballot, lane 0 lowest.
__device__ __forceinline__ uint32_t quad_bits(bool x)
{
  return (uint32_t)(__builtin_amdgcn_ballot_w64(x) >> (threadIdx.x & 60u)) & 0xFu;
}

}  // namespace rtp
