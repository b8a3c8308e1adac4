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
#include "rtp_grid_device.hpp"

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

// The neighbour sums of boids.cl:96-113, in the walk's order.
struct BoidsSums {
  int count = 0;
  V3 avgPos{ 0.0f, 0.0f, 0.0f }, avgVel{ 0.0f, 0.0f, 0.0f }, rep{ 0.0f, 0.0f, 0.0f };
};

// boids.cl:100-103: the squared distance of a candidate, and whether it is a neighbour.
__device__ __forceinline__ float boids_sq(const float4 p, const float4 q)
{
  const float vx = p.x - q.x, vy = p.y - q.y, vz = p.z - q.z;
  return vx * vx + vy * vy + vz * vz;
}
__device__ __forceinline__ bool boids_near(float sq, const BoidsConsts& c)
{
  return (sq < c.radiusSq) & (sq > kFloatEps);
}

// boids.cl:104-112 for a neighbour at q with unit velocity nv.
__device__ __forceinline__ void boids_pair(const float4 p, const float4 q, const float4 nv,
                                           BoidsSums& S, const BoidsConsts& c)
{
  const float vx = p.x - q.x, vy = p.y - q.y, vz = p.z - q.z;
  const float sq = vx * vx + vy * vy + vz * vz;
  S.avgPos.x = S.avgPos.x + q.x;
  S.avgPos.y = S.avgPos.y + q.y;
  S.avgPos.z = S.avgPos.z + q.z;
  S.avgVel.x = S.avgVel.x + nv.x;
  S.avgVel.y = S.avgVel.y + nv.y;
  S.avgVel.z = S.avgVel.z + nv.z;
  // vec / squaredDist: sq lies in (1e-8, radiusSq] with radiusSq <= 2^30 (checked on the
  // host), so the short division applies unless a numerator is a nonzero value under 2^-90
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
  S.rep.x = S.rep.x + qx;
  S.rep.y = S.rep.y + qy;
  S.rep.z = S.rep.z + qz;
  ++S.count;
}

// The cells of boids.cl:64-95 (3-D: the 27 cells inside the walls, x outer, z inner) or
// :170-182 (2-D: the YZ neighbours in the x layer RX/2, every coordinate checked against RX
// and the index formed as (RX/2 * RX + y) * RY + z; on the cubic grids the reference runs
// in 2-D these are the grid's own bounds and index; an index past the table, a non-cubic
// grid where the reference reads out of bounds, is skipped), as visit(cell, z-column, k):
// column `col` holds up to three cells k = 0..2 of consecutive indices.
template <int DIM2, class V>
__device__ __forceinline__ void boids_cells(const float4 p, const BoidsConsts& c, V&& visit)
{
  uint32_t cx, cy, cz;
  cell3d(p, c.g, cx, cy, cz);
  const int32_t rx = c.g.res[0], ry = c.g.res[1], rz = c.g.res[2];
  if (DIM2) {
    const int32_t nx = (int32_t)cx;
    for (int iy = -1; iy <= 1; ++iy)
      for (int iz = -1; iz <= 1; ++iz) {
        const int32_t ny = (int32_t)cy + iy, nz = (int32_t)cz + iz;
        if (nx < 0 || ny < 0 || nz < 0 || nx >= rx || ny >= rx || nz >= rx)
          continue;
        const uint32_t cell = (uint32_t)((rx / 2 * rx + ny) * ry + nz);
        if (cell < c.g.numCells)
          visit(cell, iy + 1, iz + 1);
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
          visit((uint32_t)((nx * ry + ny) * rz + nz), (ix + 1) * 3 + iy + 1, iz + 1);
        }
      }
    }
  }
}

// The neighbour sums with each pair's body run in place (the persistent update's walk).
template <int DIM2>
__device__ __forceinline__ void boids_sums_in_place(const float4 p, const float4* __restrict__ position,
                                                    const float4* __restrict__ unitVel,
                                                    const uint2* __restrict__ se,
                                                    const BoidsConsts& c, BoidsSums& S)
{
  boids_cells<DIM2>(p, c, [&](uint32_t cell, int, int) {
    const uint2 r = se[cell];
    for (uint32_t e = r.x; e <= r.y; ++e) {
      const float4 q = position[e];
      if (boids_near(boids_sq(p, q), c))
        boids_pair(p, q, unitVel[e], S, c);
    }
  });
}

// ---------------------------------------------------------- the walk with survivor queues
// A wave whose lanes ran each pair's body in place would run it for almost every candidate
// step with few lanes on (a boid keeps about one candidate in six).  Instead each lane lists
// its candidate ranges in walk order (the cells of a z-column merged into one range when
// their non-empty ranges are back to back, which visits the same candidates in the same
// order), the wave steps through every lane's list together, kBoidsStep candidates per lane
// a step, each lane queueing the neighbours it finds in its LDS column, and when any lane's
// queue is nearly full, and at the end, every lane runs its queued bodies in order.  Each
// lane's sums therefore add the same terms in the same order.
constexpr uint32_t kBoidsQueue = 32;   // queue entries per lane
constexpr uint32_t kBoidsRanges = 27;  // at most one range per cell
constexpr uint32_t kBoidsStep = 4;     // candidates per lane per scan step
constexpr uint32_t kBoidsDrain = 4;    // queued bodies per drain group

// Dynamic LDS of k_boids_step: lane l's queue entries at l, l + B, ... (B threads per block),
// then its ranges likewise.
inline size_t boids_step_lds(uint32_t block)
{
  return (size_t)block * (kBoidsQueue * sizeof(uint32_t) + kBoidsRanges * sizeof(uint2));
}

template <int DIM2>
__device__ __forceinline__ void boids_sums_queued(const float4 p, const float4* __restrict__ position,
                                                  const float4* __restrict__ unitVel,
                                                  const uint2* __restrict__ se,
                                                  const BoidsConsts& c, BoidsSums& S)
{
  extern __shared__ uint32_t boidsLds[];
  const uint32_t B = blockDim.x;
  uint32_t* const qslot = boidsLds + threadIdx.x;
  uint2* const rl = reinterpret_cast<uint2*>(boidsLds + kBoidsQueue * B) + threadIdx.x;
  // the lane's ranges
  uint32_t nr = 0, lo = 1u, hi = 0u;
  int lastCol = -1;
  boids_cells<DIM2>(p, c, [&](uint32_t cell, int col, int) {
    const uint2 r = se[cell];
    if (r.x > r.y)
      return;
    if (col == lastCol && lo <= hi && r.x == hi + 1u) {
      hi = r.y;  // back to back with the column's range so far
      return;
    }
    if (lo <= hi)
      rl[(nr++) * B] = make_uint2(lo, hi);
    lo = r.x;
    hi = r.y;
    lastCol = col;
  });
  if (lo <= hi)
    rl[(nr++) * B] = make_uint2(lo, hi);
  // the scan
  uint32_t* wp = qslot;
  uint32_t* const room = qslot + (kBoidsQueue - kBoidsStep) * B;
  uint32_t rk = 0, e = 1u;
  hi = 0u;
  auto next_range = [&]() {
    if (rk < nr) {
      const uint2 r = rl[(rk++) * B];
      e = r.x;
      hi = r.y;
    }
  };
  auto drain = [&]() {
    for (const uint32_t* rp = qslot; rp < wp; rp += kBoidsDrain * B) {
      uint32_t v[kBoidsDrain];
      float4 q[kBoidsDrain], nv[kBoidsDrain];
#pragma unroll
      for (uint32_t k = 0; k < kBoidsDrain; ++k) {
        const uint32_t* pk = rp + k * B;
        v[k] = *(pk < wp ? pk : qslot);  // in bounds either way
        q[k] = position[v[k]];
        nv[k] = unitVel[v[k]];
      }
#pragma unroll
      for (uint32_t k = 0; k < kBoidsDrain; ++k)
        if (rp + k * B < wp)
          boids_pair(p, q[k], nv[k], S, c);
    }
    wp = qslot;
  };
  next_range();
  for (;;) {
    const bool on = e <= hi;
    if (__builtin_amdgcn_ballot_w64(on) == 0)
      break;
    if (on) {
      float4 q[kBoidsStep];
#pragma unroll
      for (uint32_t k = 0; k < kBoidsStep; ++k)
        q[k] = position[min(e + k, hi)];  // the loads past hi repeat hi's
#pragma unroll
      for (uint32_t k = 0; k < kBoidsStep; ++k) {
        *wp = e + k;
        wp += (boids_near(boids_sq(p, q[k]), c) & (e + k <= hi)) ? B : 0u;
      }
      e += kBoidsStep;
      if (e > hi)
        next_range();
    }
    if (__builtin_amdgcn_ballot_w64(wp > room) != 0)
      drain();
  }
  drain();
}

// Boid i of the sorted state: the rules over its neighbours (sums S), the velocity update
// and the wall (or periodic) boundary, written to the other pair of buffers.
template <int DIM2>
__device__ __forceinline__ void boids_finish(uint32_t i, const float4 p, BoidsSums S,
                                             const float4* __restrict__ velocity,
                                             float4* __restrict__ posOut,
                                             float4* __restrict__ velOut,
                                             float4* __restrict__ accOut, const BoidsConsts& c)
{
  const int count = S.count;
  V3 avgPos = S.avgPos, avgVel = S.avgVel, rep = S.rep;
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

// The persistent update's per-boid step: sums with bodies in place, then the update.
template <int DIM2>
__device__ __forceinline__ void boids_step_body(uint32_t i, const float4* __restrict__ position,
                                                const float4* __restrict__ velocity,
                                                const float4* __restrict__ unitVel,
                                                const uint2* __restrict__ se,
                                                float4* __restrict__ posOut,
                                                float4* __restrict__ velOut,
                                                float4* __restrict__ accOut, const BoidsConsts& c)
{
  const float4 p = position[i];
  BoidsSums S;
  boids_sums_in_place<DIM2>(p, position, unitVel, se, c, S);
  boids_finish<DIM2>(i, p, S, velocity, posOut, velOut, accOut, c);
}

// QUEUED: the walk with survivor queues (its LDS is boids_step_lds(blockDim.x)), else
// the bodies in place.
template <int DIM2, int QUEUED>
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
  BoidsSums S;
  if (QUEUED)
    boids_sums_queued<DIM2>(p, position, unitVel, se, c, S);
  else
    boids_sums_in_place<DIM2>(p, position, unitVel, se, c, S);
  boids_finish<DIM2>(i, p, S, velocity, posOut, velOut, accOut, c);
}

// ------------------------------------------------------------------- the quad walk
// For flocks too small to fill the GPU with one lane per boid (RTP_LAUNCH_QUAD), each boid
// gets a quad of four lanes: the quad lists the boid's ranges as boids_sums_queued does (in
// LDS), scans them sixteen candidates per step (lane u: candidates u, u + 4, u + 8, u + 12),
// queues the neighbours in walk order, and runs four queued bodies at once, one per lane,
// adding their terms to every sum in walk order (quad_add).  The sums therefore add the same
// terms in the same order as one lane's walk.
constexpr uint32_t kBoidsQuadGroups = 16;  // boids per 64-lane workgroup
constexpr uint32_t kBoidsQuadQueue = 64;   // queued neighbours per boid

struct BoidsQuadLds {
  uint32_t queue[kBoidsQuadGroups][kBoidsQuadQueue];
  uint2 ranges[kBoidsQuadGroups][kBoidsRanges];
};

// boids_pair's terms for the quad's four neighbours (v: this lane has one), added in order.
__device__ __forceinline__ void boids_pair_quad(bool v, const float4 p, const float4 q,
                                                const float4 nv, BoidsSums& S,
                                                const BoidsConsts& c)
{
  const float vx = p.x - q.x, vy = p.y - q.y, vz = p.z - q.z;
  const float sq = vx * vx + vy * vy + vz * vz;
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
  S.avgPos.x = quad_add(S.avgPos.x, v ? q.x : 0.0f);
  S.avgPos.y = quad_add(S.avgPos.y, v ? q.y : 0.0f);
  S.avgPos.z = quad_add(S.avgPos.z, v ? q.z : 0.0f);
  S.avgVel.x = quad_add(S.avgVel.x, v ? nv.x : 0.0f);
  S.avgVel.y = quad_add(S.avgVel.y, v ? nv.y : 0.0f);
  S.avgVel.z = quad_add(S.avgVel.z, v ? nv.z : 0.0f);
  S.rep.x = quad_add(S.rep.x, v ? qx : 0.0f);
  S.rep.y = quad_add(S.rep.y, v ? qy : 0.0f);
  S.rep.z = quad_add(S.rep.z, v ? qz : 0.0f);
  S.count += __builtin_popcount(quad_bits(v));
}

template <int DIM2>
__device__ __forceinline__ void boids_sums_quad(bool on, uint32_t self, const float4 p,
                                                const float4* __restrict__ position,
                                                const float4* __restrict__ unitVel,
                                                const uint2* __restrict__ se,
                                                const BoidsConsts& c, BoidsQuadLds& lds,
                                                BoidsSums& S)
{
  const uint32_t gq = threadIdx.x >> 2, u = threadIdx.x & 3u, below = (1u << u) - 1u;
  uint2* const rl = lds.ranges[gq];
  uint32_t* const q = lds.queue[gq];
  // the boid's ranges (every lane of the quad lists the same ones)
  uint32_t nr = 0, e = 1u, hi = 0u;
  if (on) {
    uint32_t lo = 1u;
    int lastCol = -1;
    boids_cells<DIM2>(p, c, [&](uint32_t cell, int col, int) {
      const uint2 r = se[cell];
      if (r.x > r.y)
        return;
      if (col == lastCol && lo <= hi && r.x == hi + 1u) {
        hi = r.y;
        return;
      }
      if (lo <= hi)
        rl[nr++] = make_uint2(lo, hi);
      lo = r.x;
      hi = r.y;
      lastCol = col;
    });
    if (lo <= hi)
      rl[nr++] = make_uint2(lo, hi);
    hi = 0u;
  }
  uint32_t ri = 0, qn = 0;
  if (nr > 0) {
    e = rl[0].x;
    hi = rl[0].y;
  }
  const auto drain = [&]() {
    wave_sync();
    bool v = u < qn;
    uint32_t ev = v ? q[u] : self;
    float4 qq = position[ev], nv = unitVel[ev];
    for (uint32_t k = u; __builtin_amdgcn_ballot_w64(k - u < qn) != 0; k += 4u) {
      const uint32_t k1 = k + 4u;
      const bool v1 = k1 < qn;
      const uint32_t e1 = v1 ? q[k1] : self;
      const float4 q1 = position[e1], nv1 = unitVel[e1];
      boids_pair_quad(v, p, qq, nv, S, c);
      v = v1;
      qq = q1;
      nv = nv1;
    }
    qn = 0;
    wave_sync();
  };
  while (__builtin_amdgcn_ballot_w64(ri < nr) != 0) {
    const bool act = ri < nr;
    uint32_t ej[4];
    bool kept[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      ej[j] = e + u + 4u * (uint32_t)j;
      const bool valid = act & (ej[j] <= hi);
      kept[j] = valid & boids_near(boids_sq(p, position[valid ? ej[j] : self]), c);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t m = quad_bits(kept[j]);
      if (kept[j])
        q[qn + (uint32_t)__builtin_popcount(m & below)] = ej[j];
      qn += (uint32_t)__builtin_popcount(m);
    }
    e += 16u;
    if (act & (e > hi)) {
      ++ri;
      if (ri < nr) {
        e = rl[ri].x;
        hi = rl[ri].y;
      }
    }
    if (__builtin_amdgcn_ballot_w64(qn > kBoidsQuadQueue - 16u) != 0)
      drain();
  }
  drain();
}

// k_boids_step with a quad of lanes per boid.
template <int DIM2>
__global__ __launch_bounds__(64) void k_boids_step_q(const float4* __restrict__ position,
                                                     const float4* __restrict__ velocity,
                                                     const float4* __restrict__ unitVel,
                                                     const uint2* __restrict__ se,
                                                     float4* __restrict__ posOut,
                                                     float4* __restrict__ velOut,
                                                     float4* __restrict__ accOut, uint32_t n,
                                                     BoidsConsts c, int remap)
{
  __shared__ BoidsQuadLds lds;
  const uint32_t t = logical_block(remap) * kBoidsQuadGroups + (threadIdx.x >> 2);
  const bool on = t < n;
  const uint32_t i = on ? t : 0u;
  const float4 p = position[i];
  BoidsSums S;
  boids_sums_quad<DIM2>(on, i, p, position, unitVel, se, c, lds, S);
  if (on && (threadIdx.x & 3u) == 0u)
    boids_finish<DIM2>(i, p, S, velocity, posOut, velOut, accOut, c);
}

// Whether the current device gives one workgroup `bytes` of LDS (the limit is read once per
// device).
bool lds_fits(size_t bytes)
{
  static int limit[64] = {};  // 0 unknown
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
    return false;
  if (limit[dev] == 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
      v = -1;
    limit[dev] = v;
  }
  return limit[dev] > 0 && bytes <= (size_t)limit[dev];
}

// ----------------------------------------------------------------- persistent update
// Every step of an update in one cooperative launch: the five launches of a step (hash,
// sort, reorder, unit velocities, rules) become three phases of one loop separated by
// grid-wide barriers, with the same per-element bodies, so the results are those of the
// separate launches bit for bit.  Workgroup 0 runs the one-workgroup LDS sort; the element phases
// hand wave w of workgroup b the 64 elements from 64 * (b + G * w), so the waves of a phase
// spread over the workgroups.

// Grid-wide barrier: each workgroup's first thread adds to bar[0]; the last to arrive resets
// it and advances the generation bar[1], which the others poll (agent-scope release and
// acquire, so the writes of every workgroup before the barrier are seen by every workgroup
// after it).  The poll is bounded: after kBarrierPolls a workgroup sets bar[2] and goes on,
// and every later barrier returns at once, so a launch always ends (the host reports it).
constexpr uint32_t kBarrierPolls = 1u << 24;

__device__ __forceinline__ void grid_barrier(uint32_t* bar, uint32_t& gen)
{
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t arrived =
        __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (arrived + 1u == gridDim.x) {
      __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&bar[1], gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      for (uint32_t k = 0;; ++k) {
        if (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != gen)
          break;
        if (__hip_atomic_load(&bar[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
          break;
        if (k == kBarrierPolls) {
          __hip_atomic_store(&bar[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(4);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  ++gen;
  __syncthreads();
}

// Phases of one step, three barriers apart: (1) the cell table reset and, in workgroup 0,
// the sort; (2) the gather with the cell table, and each boid's unit velocity from the
// velocity its thread just gathered; (3) the rules, the velocity update and the wall, the
// sorted tail carried over, and each boid's cell key for the next step from the position
// its thread just wrote (not after the last step: the keys then stay the sorted ones, as
// after separate launches).  The first step's keys are hashed before the loop.
template <int DIM2>
__global__ __launch_bounds__(512) void k_boids_persistent(BoidsPersistentArgs a, BoidsConsts c)
{
  __shared__ SmallSortLds lds;
  uint32_t gen = 0;
  const uint32_t t0 = (threadIdx.x & 63u) + 64u * (blockIdx.x + gridDim.x * (threadIdx.x >> 6));
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t n = a.n, sn = a.sortN, nc = c.g.numCells;
  int kc = a.keyCur;
  const int res = a.passes & 1;
  float4* pos = a.pos[a.cur];
  float4* vel = a.vel[a.cur];
  float4* pos2 = a.pos[1 - a.cur];
  float4* vel2 = a.vel[1 - a.cur];
  for (uint32_t t = t0; t < n; t += stride)  // k_boids_hash of the first step
    a.key[kc][t] = cell1d(pos[t], c.g);
  grid_barrier(a.bar, gen);
  for (uint32_t step = 0; step < a.steps; ++step) {
    for (uint32_t t = t0; t < nc; t += stride)  // resetStartEndCell (k_boids_hash)
      a.startEnd[t] = make_uint2(1u, 0u);
    if (blockIdx.x == 0)  // k_sort_small: the sorted keys land in key[kc ^ res]
      small_sort_block<16>(a.key[kc], a.key[1 - kc], a.idx[0], a.idx[1], sn, a.passes, lds);
    kc = res ? 1 - kc : kc;
    grid_barrier(a.bar, gen);
    ReorderArgs r{};
    r.key = a.key[kc];
    r.idx = a.idx[res];
    r.sortN = sn;
    r.n = n;
    r.maxN = a.maxN;
    r.srcA = pos;
    r.dstA = pos2;
    r.srcB = vel;
    r.dstB = vel2;
    r.startEnd = a.startEnd;
    r.g = c.g;
    for (uint32_t t = t0; t < sn; t += stride) {  // k_reorder, k_boids_unit_vel
      reorder_body(t, r, 0u, 0u);
      const float4 v = vel2[t];
      const V3 u = normalize3(V3{ v.x, v.y, v.z });
      a.unitVel[t] = make_float4(u.x, u.y, u.z, 0.0f);
    }
    grid_barrier(a.bar, gen);
    const bool more = step + 1u < a.steps;
    for (uint32_t t = t0; t < sn; t += stride) {  // k_boids_step, the tail, the next hash
      if (t < n) {
        boids_step_body<DIM2>(t, pos2, vel2, a.unitVel, a.startEnd, pos, vel, a.acc, c);
        if (more)
          a.key[kc][t] = cell1d(pos[t], c.g);
      } else {
        pos[t] = pos2[t];
        vel[t] = vel2[t];
      }
    }
    grid_barrier(a.bar, gen);
  }
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
  if (L.quad) {
    const dim3 grid(blocks_for(n, kBoidsQuadGroups)), block(64);
    if (c.dim2)
      hipLaunchKernelGGL(k_boids_step_q<1>, grid, block, 0, s, pos, vel, unitVel, startEnd, posOut,
                         velOut, acc, n, c, L.xcdRemap);
    else
      hipLaunchKernelGGL(k_boids_step_q<0>, grid, block, 0, s, pos, vel, unitVel, startEnd, posOut,
                         velOut, acc, n, c, L.xcdRemap);
    return;
  }
  // the survivor-queue walk where its LDS fits in a workgroup, else bodies in place
  const size_t lds = boids_step_lds(L.block);
  const bool queued = lds_fits(lds);
  const dim3 grid(blocks_for(n, L.block)), block(L.block);
  if (c.dim2) {
    if (queued)
      hipLaunchKernelGGL((k_boids_step<1, 1>), grid, block, lds, s, pos, vel, unitVel, startEnd,
                         posOut, velOut, acc, n, c, L.xcdRemap);
    else
      hipLaunchKernelGGL((k_boids_step<1, 0>), grid, block, 0, s, pos, vel, unitVel, startEnd,
                         posOut, velOut, acc, n, c, L.xcdRemap);
  } else {
    if (queued)
      hipLaunchKernelGGL((k_boids_step<0, 1>), grid, block, lds, s, pos, vel, unitVel, startEnd,
                         posOut, velOut, acc, n, c, L.xcdRemap);
    else
      hipLaunchKernelGGL((k_boids_step<0, 0>), grid, block, 0, s, pos, vel, unitVel, startEnd,
                         posOut, velOut, acc, n, c, L.xcdRemap);
  }
}

int launch_boids_persistent(const BoidsPersistentArgs& a, const BoidsConsts& c, uint32_t groups,
                            hipStream_t s)
{
  const void* fn = c.dim2 ? reinterpret_cast<const void*>(&k_boids_persistent<1>)
                          : reinterpret_cast<const void*>(&k_boids_persistent<0>);
  BoidsPersistentArgs args = a;
  BoidsConsts consts = c;
  void* params[] = { &args, &consts };
  if (hipLaunchCooperativeKernel(fn, dim3(groups), dim3(kSmallThreads), params, 0, s) != hipSuccess)
    return fail(RTP_ERR_HIP, "cooperative launch of the persistent boids update failed");
  return RTP_OK;
}

uint32_t boids_persistent_max_groups(int dim2)
{
  const void* fn = dim2 ? reinterpret_cast<const void*>(&k_boids_persistent<1>)
                        : reinterpret_cast<const void*>(&k_boids_persistent<0>);
  int perCu = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, fn, (int)kSmallThreads, 0) != hipSuccess ||
      hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  return (uint32_t)(perCu > 0 && cus > 0 ? perCu * cus : 0);
}

}  // namespace rtp
