// Boids steering pass for gfx950.
//
// Reference: physics/ocl/kernels/boids.cl:46-132 (bd_applyBoidsRulesWithGrid3D),
// :137-221 (bd_applyBoidsRulesWithGrid2D), :226-241 (bd_addTargetRule), :246-259
// (bd_updateVel), :264-284 (bd_updatePosAndApplyWallBC), :289-315
// (bd_updatePosAndApplyPeriodicBC), step order physics/ocl/Boids.cpp:323-384.  The kernels
// of one step are fused into one pass that reads the sorted (position, velocity) buffers and
// writes the next state into the other pair of buffers, so no boid reads a neighbour's
// already-updated state.
#include "rtp_crmath.hpp"
#include "rtp_device.hpp"

namespace rtp {

namespace {

// fillCellIDs on the current positions (Boids.cpp:335) + resetStartEndCell.
__global__ __launch_bounds__(256) void k_boids_hash(const float4* __restrict__ pos,
                                                    uint32_t* __restrict__ key,
                                                    uint2* __restrict__ startEnd, uint32_t n,
                                                    GridConsts g)
{
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n)
    key[t] = cell1d(pos[t], g);
  if (t < g.numCells)
    startEnd[t] = make_uint2(1u, 0u);
}

struct V3 {
  float x, y, z;
};

// fast_normalize pinned as v / sqrtf(dot(v, v)), 0 for the zero vector.
__device__ __forceinline__ V3 normalize3(V3 v)
{
  if (v.x == 0.0f && v.y == 0.0f && v.z == 0.0f)
    return v;
  const float l = __builtin_sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
  return V3{ v.x / l, v.y / l, v.z / l };
}

// fast_normalize(velocity[e]) of boids.cl:105, once per boid: the rules read a neighbour's
// unit velocity from here instead of normalising it again for every pair.
__global__ __launch_bounds__(256) void k_boids_unit_vel(const float4* __restrict__ velocity,
                                                        float4* __restrict__ unitVel, uint32_t n)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const float4 v = velocity[i];
  const V3 u = normalize3(V3{ v.x, v.y, v.z });
  unitVel[i] = make_float4(u.x, u.y, u.z, 0.0f);
}

template <int DIM2>
__global__ __launch_bounds__(256) void k_boids_step(const float4* __restrict__ position,
                                                    const float4* __restrict__ velocity,
                                                    const float4* __restrict__ unitVel,
                                                    const uint2* __restrict__ se,
                                                    float4* __restrict__ posOut,
                                                    float4* __restrict__ velOut,
                                                    float4* __restrict__ accOut, uint32_t n,
                                                    BoidsConsts c, int remap)
{
  const uint32_t i = logical_block(remap) * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const float4 p = position[i];
  uint32_t cx, cy, cz;
  cell3d(p, c.g, cx, cy, cz);
  const int32_t rx = c.g.res[0], ry = c.g.res[1], rz = c.g.res[2];
  int count = 0;
  V3 avgPos{ 0.0f, 0.0f, 0.0f }, avgVel{ 0.0f, 0.0f, 0.0f }, rep{ 0.0f, 0.0f, 0.0f };
  auto visit = [&](uint32_t cell) {
    const uint2 r = se[cell];
    for (uint32_t e = r.x; e <= r.y; ++e) {
      const float4 q = position[e];
      const float vx = p.x - q.x, vy = p.y - q.y, vz = p.z - q.z;
      const float sq = vx * vx + vy * vy + vz * vz;
      if (sq < c.radiusSq && sq > kFloatEps) {  // boids.cl:102-103
        avgPos.x = avgPos.x + q.x;
        avgPos.y = avgPos.y + q.y;
        avgPos.z = avgPos.z + q.z;
        const float4 nv = unitVel[e];
        avgVel.x = avgVel.x + nv.x;
        avgVel.y = avgVel.y + nv.y;
        avgVel.z = avgVel.z + nv.z;
        // vec / squaredDist: sq lies in (1e-8, radiusSq] with radiusSq <= 2^30 (checked on
        // the host), so the short division applies unless a numerator is a nonzero value
        // under 2^-90
        const float y = rcp_refined(sq);
        float qx = div_short(vx, sq, y), qy = div_short(vy, sq, y), qz = div_short(vz, sq, y);
        const bool general = !c.shortDiv || tiny_numerator3(vx, vy, vz);
        if (__builtin_amdgcn_ballot_w64(general) != 0) {
          rare_path();
          if (general) {
            qx = vx / sq;
            qy = vy / sq;
            qz = vz / sq;
          }
        }
        rep.x = rep.x + qx;
        rep.y = rep.y + qy;
        rep.z = rep.z + qz;
        ++count;
      }
    }
  };
  if (DIM2) {
    // boids.cl:170-182: the YZ neighbours in the x layer RX/2, every coordinate checked
    // against RX and the index formed as (RX/2 * RX + y) * RY + z.  On the cubic grids the
    // reference runs in 2-D these are the grid's own bounds and index; an index past the
    // table (a non-cubic grid, where the reference reads out of bounds) is skipped.
    const int32_t nx = (int32_t)cx;
    for (int iy = -1; iy <= 1; ++iy)
      for (int iz = -1; iz <= 1; ++iz) {
        const int32_t ny = (int32_t)cy + iy, nz = (int32_t)cz + iz;
        if (nx < 0 || ny < 0 || nz < 0 || nx >= rx || ny >= rx || nz >= rx)
          continue;
        const uint32_t cell = (uint32_t)((rx / 2 * rx + ny) * ry + nz);
        if (cell < c.g.numCells)
          visit(cell);
      }
  } else {
    for (int ix = -1; ix <= 1; ++ix) {
      const int32_t nx = (int32_t)cx + ix;
      if (nx < 0 || nx >= rx)
        continue;
      for (int iy = -1; iy <= 1; ++iy) {
        const int32_t ny = (int32_t)cy + iy;
        if (ny < 0 || ny >= ry)
          continue;
        for (int iz = -1; iz <= 1; ++iz) {
          const int32_t nz = (int32_t)cz + iz;
          if (nz < 0 || nz >= rz)
            continue;
          visit((uint32_t)((nx * ry + ny) * rz + nz));
        }
      }
    }
  }
  V3 acc{ 0.0f, 0.0f, 0.0f };
  if (count != 0) {  // boids.cl:115-129
    const float fc = (float)count;
    avgPos = V3{ avgPos.x / fc - p.x, avgPos.y / fc - p.y, avgPos.z / fc - p.z };
    const float s = c.p.velocityScale;
    const V3 coh = normalize3(avgPos), ali = normalize3(avgVel), sep = normalize3(rep);
    const V3 a{ coh.x * s, coh.y * s, coh.z * s };
    const V3 b{ ali.x * s, ali.y * s, ali.z * s };
    const V3 d{ sep.x * s, sep.y * s, sep.z * s };
    acc.x = (b.x * c.p.alignmentScale + d.x * c.p.separationScale) + a.x * c.p.cohesionScale;
    acc.y = (b.y * c.p.alignmentScale + d.y * c.p.separationScale) + a.y * c.p.cohesionScale;
    acc.z = (b.z * c.p.alignmentScale + d.z * c.p.separationScale) + a.z * c.p.cohesionScale;
  }
  if (c.targetOn) {  // bd_addTargetRule, boids.cl:226-241 (MAX_STEERING 0.5f, :10)
    const V3 tv{ c.target[0] - p.x, c.target[1] - p.y, c.target[2] - p.z };
    const float dist = __builtin_sqrtf(tv.x * tv.x + tv.y * tv.y + tv.z * tv.z);
    if (dist < c.targetRadius) {
      const float k = clampf(1.3f / dist, 0.0f, 1.4f * 0.5f);
      acc.x = acc.x + (c.targetSign * tv.x) * k;
      acc.y = acc.y + (c.targetSign * tv.y) * k;
      acc.z = acc.z + (c.targetSign * tv.z) * k;
    }
  }
  accOut[i] = make_float4(acc.x, acc.y, acc.z, 0.0f);

  // bd_updateVel, boids.cl:246-259 (maxVelocity = velocityScale, Boids.cpp:217)
  const float4 v = velocity[i];
  const V3 nv{ v.x + acc.x * c.dt, v.y + acc.y * c.dt, v.z + acc.z * c.dt };
  const float len = __builtin_sqrtf(nv.x * nv.x + nv.y * nv.y + nv.z * nv.z);
  const float norm = clampf(len, c.vmin, c.p.velocityScale);
  const V3 dir = normalize3(nv);
  V3 vel{ dir.x * norm, dir.y * norm, dir.z * norm };

  // bd_updatePosAndApplyWallBC, boids.cl:264-284, or bd_updatePosAndApplyPeriodicBC,
  // :289-315 (a clamped coordinate re-enters at the opposite wall, the velocity is kept)
  const V3 np{ p.x + vel.x * c.dt, p.y + vel.y * c.dt, p.z + vel.z * c.dt };
  V3 cp{ clampf(np.x, -c.g.absWall[0], c.g.absWall[0]),
         clampf(np.y, -c.g.absWall[1], c.g.absWall[1]),
         clampf(np.z, -c.g.absWall[2], c.g.absWall[2]) };
  if (c.cyclic) {
    if (!(cp.x == np.x))
      cp.x = cp.x * -1.0f;
    if (!(cp.y == np.y))
      cp.y = cp.y * -1.0f;
    if (!(cp.z == np.z))
      cp.z = cp.z * -1.0f;
  } else if (!(cp.x == np.x && cp.y == np.y && cp.z == np.z)) {
    vel = V3{ vel.x * -0.5f, vel.y * -0.5f, vel.z * -0.5f };
  }
  posOut[i] = make_float4(cp.x, cp.y, cp.z, 0.0f);
  velOut[i] = make_float4(vel.x, vel.y, vel.z, 0.0f);
}

inline uint32_t blocks_for(uint32_t n, uint32_t b) { return (n + b - 1u) / b; }

}  // namespace

void launch_boids_hash(const float4* pos, uint32_t* key, uint2* startEnd, uint32_t n,
                       const GridConsts& g, hipStream_t s)
{
  const uint32_t m = n > g.numCells ? n : g.numCells;
  if (m == 0)
    return;
  hipLaunchKernelGGL(k_boids_hash, dim3(blocks_for(m, 256)), dim3(256), 0, s, pos, key, startEnd,
                     n, g);
}

void launch_boids_step(const float4* pos, const float4* vel, float4* unitVel,
                       const uint2* startEnd, float4* posOut, float4* velOut, float4* acc,
                       uint32_t n, uint32_t sortN, const BoidsConsts& c, const Launch& L,
                       hipStream_t s)
{
  if (n == 0)
    return;
  // every sorted slot a cell range can reach: stale keys of a shrunk count sort among the
  // active ones
  hipLaunchKernelGGL(k_boids_unit_vel, dim3(blocks_for(sortN, 256)), dim3(256), 0, s, vel,
                     unitVel, sortN);
  if (c.dim2)
    hipLaunchKernelGGL(k_boids_step<1>, dim3(blocks_for(n, L.block)), dim3(L.block), 0, s, pos,
                       vel, unitVel, startEnd, posOut, velOut, acc, n, c, L.xcdRemap);
  else
    hipLaunchKernelGGL(k_boids_step<0>, dim3(blocks_for(n, L.block)), dim3(L.block), 0, s, pos,
                       vel, unitVel, startEnd, posOut, velOut, acc, n, c, L.xcdRemap);
}

}  // namespace rtp
