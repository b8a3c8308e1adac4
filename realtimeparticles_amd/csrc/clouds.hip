// The Clouds model (physics/ocl/Clouds.{hpp,cpp}, physics/ocl/kernels/clouds.cl): position
// based fluids coupled to a per-particle temperature, vapour and cloud density, with a
// temperature-smoothing constraint, on the uniform grid and radix sort of the fluids model.
//
// One Clouds::update (Clouds.cpp:503-610, without the rendering steps) is
//   k_cld_thermo        heat from the ground, buoyancy, adiabatic cooling, cloud generation,
//                       phase transition, latent heat, predicted position, the mixed
//                       boundary (periodic x/z, walls in y) and the cell key, fused
//   radix sort          of the keys, and k_reorder + k_cld_gather for the ten payloads
//   k_cld_temp_lambda   Laplacian of the temperature and its constraint factor (one walk)
//   k_cld_temp_correct  the temperature correction, applied at 0.3
//   iters x {k_cld_density_lambda, k_cld_correct}   (the last also updates the velocity)
//   k_cld_vorticity, k_cld_confinement, k_cld_xsph (+ cld_updatePosition), or
//   k_cld_update_pos without vorticity.
// The two temperature passes run on a second stream beside the density iterations (neither
// reads what the other writes); the step joins them before the vorticity pass.  The first
// walk over each set of positions (temperature, each density iteration, vorticity) records
// its survivors, and the passes after it over the same positions replay them.
// The neighbour walks visit the reference's 27 cells in its order (x outer, z inner), with x
// and z periodic -- a wrapped neighbour's position is shifted by +-2 ABS_WALL -- and y cells
// beyond the walls skipped (clouds.cl:322-353).  Every expression follows clouds.cl's order
// with the float4 lanes the reference evaluates (w included), so the results are the CPU
// oracle's bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "rtp_crmath.hpp"
#include "rtp_device.hpp"
#include "rtp_exp.hpp"
#include "rtp_host.hpp"
#include "rtp_internal.hpp"

namespace rtp {
namespace {

struct CloudConsts {
  PbfConsts f;         // grid, FluidParams and the SPH constants of the program
  rtp_cloud_params q;  // CloudParams (clouds.cl:18-32)
};

// ---------------------------------------------------------------- float4 as OpenCL sees it

__device__ __forceinline__ float4 sub4(float4 a, float4 b)
{
  return make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w);
}
__device__ __forceinline__ float4 add4(float4 a, float4 b)
{
  return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
__device__ __forceinline__ float4 mul4s(float4 a, float s)
{
  return make_float4(a.x * s, a.y * s, a.z * s, a.w * s);
}
__device__ __forceinline__ float4 muls4(float s, float4 a)
{
  return make_float4(s * a.x, s * a.y, s * a.z, s * a.w);
}
__device__ __forceinline__ float4 div4s(float4 a, float s)
{
  return make_float4(a.x / s, a.y / s, a.z / s, a.w / s);
}
__device__ __forceinline__ float dot4(float4 a, float4 b)
{
  return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w;
}
__device__ __forceinline__ float4 normalize4(float4 v)
{
  if (v.x == 0.0f && v.y == 0.0f && v.z == 0.0f && v.w == 0.0f)
    return v;
  return div4s(v, __builtin_sqrtf(dot4(v, v)));
}
__device__ __forceinline__ float4 cross4(float4 a, float4 b)
{
  return make_float4(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0.0f);
}
__device__ __forceinline__ float step1(float edge, float x) { return x < edge ? 0.0f : 1.0f; }

// sph.cl:10-14 given r = length(vec) = sqrtf(dot4(vec, vec))
__device__ __forceinline__ float poly6_r(float r, const PbfConsts& c)
{
  return ((1.0f - step1(c.h, r)) * c.poly6) * pown(c.h * c.h - r * r, 3);
}

// sph.cl:26-34 given r = length(vec)
__device__ __forceinline__ float4 grad_spiky_r(float4 vec, float r, const PbfConsts& c)
{
  if (r <= kFloatEps)
    return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float4 g = mul4s(vec, 1.0f - step1(c.h, r));
  g = mul4s(g, c.spiky);
  g = mul4s(g, -3.0f);
  g = mul4s(g, pown(c.h - r, 2));
  return div4s(g, r);
}

// length(vec) from d2 = dot4(vec, vec): sqrt_cr returns sqrtf's bits for every d2 >= 0.
__device__ __forceinline__ float len_of(float d2) { return sqrt_cr(d2); }

__device__ __forceinline__ bool tiny_numerator4(float a, float b, float c, float d)
{
  return tiny_numerator3(a, b, c) || num_key(d) < (kDivNumMin << 1) - 2u;
}

// grad_spiky_r with the short correctly rounded division where its operand ranges hold
// (the shortMath bit: kFloatEps < r < h, every numerator +-0 or not tiny): the four
// quotients by r share one refined reciprocal.  With t = 1 - step(h, r) = 1 the factor t
// is an exact identity.  Other lanes take the general expression.
__device__ __forceinline__ float4 grad_spiky_fast(float4 vec, float r, const PbfConsts& c)
{
  const bool in = (c.shortMath & kShortSqrtGrad) != 0 && r < c.h;
  float4 g = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float4 n = g;
  bool general = !in;
  if (in) {
    const float p2 = (c.h - r) * (c.h - r);
    n = make_float4(((vec.x * c.spiky) * -3.0f) * p2, ((vec.y * c.spiky) * -3.0f) * p2,
                    ((vec.z * c.spiky) * -3.0f) * p2, ((vec.w * c.spiky) * -3.0f) * p2);
    const float yr = rcp_refined(r);
    const bool on = r > kFloatEps;
    if (on)
      g = make_float4(div_short(n.x, r, yr), div_short(n.y, r, yr), div_short(n.z, r, yr),
                      div_short(n.w, r, yr));
    general = on && tiny_numerator4(n.x, n.y, n.z, n.w);
  }
  if (__builtin_amdgcn_ballot_w64(general) != 0) {
    rare_path();
    if (general)
      g = in ? div4s(n, r) : grad_spiky_r(vec, r, c);
  }
  return g;
}

// clouds.cl:106-114 given w = poly6(vec) (poly6L(artPressureRadius * h) is folded into
// c.wdq); ywdq = rcp_refined(c.wdq) when the short division applies (kShortArtDiv).
__device__ __forceinline__ float art_pressure_w(float w, float ywdq, const PbfConsts& c)
{
  if (c.p.isArtPressureEnabled == 0)
    return 0.0f;
  const float q = (c.shortMath & kShortArtDiv) ? div_short(w, c.wdq, ywdq) : w / c.wdq;
  return -c.p.artPressureCoeff * pown(q, c.p.artPressureExp);
}

// clouds.cl:67-70
__device__ __forceinline__ float ext_heat(float altitude, const GridConsts& g)
{
  return clampf(exp32(-(altitude + g.absWall[1]) / 3.0f), 0.0f, 1.0f);
}

// clouds.cl:75-86
__device__ __forceinline__ float env_temp(float altitude, const GridConsts& g)
{
  const float coeff = 3.5f;
  const float tempGround = 293.0f;
  return -coeff * (altitude + g.absWall[1]) + tempGround;
}

// clouds.cl:91-96 (double literals: evaluated in double, rounded to the float result)
__device__ __forceinline__ float svd(float temperature)
{
  const double t = (double)temperature;
  return (float)(217 * exp64(19.5 - 4303.4 / (t - 29.5)) / t);
}

// clouds.cl:279-300 on a predicted position
__device__ __forceinline__ float4 mixed_boundary(float4 p, const GridConsts& g)
{
  const float wx = g.absWall[0], wy = g.absWall[1], wz = g.absWall[2];
  const float cx = clampf(p.x, -wx, wx), cy = clampf(p.y, -wy, wy), cz = clampf(p.z, -wz, wz);
  float4 q = p;
  if (fabsf(p.x) > wx)
    q.x = p.x - 2 * cx;
  if (fabsf(p.y) > wy)
    q.y = cy;
  if (fabsf(p.z) > wz)
    q.z = p.z - 2 * cz;
  return q;
}

// ------------------------------------------------------------ the walk with survivor queues
// Most candidates of a walk contribute exactly +-0 and are skipped (negligible, below).  A
// wave whose lanes each ran the body in place would run it for almost every candidate step
// (some lane nearly always keeps its candidate) with few lanes on, and lanes whose cells hold
// different counts would leave the loops at different steps.  So the walk is split:
//  - each lane lists its non-empty candidate ranges in walk order (cld_ranges, in LDS);
//  - the wave steps through all lanes' lists together, kCldStep candidates per lane a step,
//    and each lane queues the candidates it keeps (index and wrap-shift code) in its own LDS
//    column;
//  - when any lane's queue is nearly full, and at the end, every lane runs the bodies of its
//    queued candidates, in order.
// Each lane's bodies run in its walk's order, so every sum is the same.
#ifndef RTP_CLD_STEP
#define RTP_CLD_STEP 4
#endif
constexpr uint32_t kCldLanes = 64;   // threads per block of a walk kernel: one wave
constexpr uint32_t kCldQueue = 32;   // queue entries per lane
constexpr uint32_t kCldStep = RTP_CLD_STEP;  // candidates per lane per scan step
#ifndef RTP_CLD_DRAIN
#define RTP_CLD_DRAIN 4
#endif
constexpr uint32_t kCldDrain = RTP_CLD_DRAIN;  // queued bodies per drain group
constexpr uint32_t kCldRanges = 27;  // at most one range per visited cell
constexpr uint32_t kCldIndexBits = 28;
constexpr uint32_t kCldIndexMask = (1u << kCldIndexBits) - 1u;

struct CldLds {
  uint32_t queue[kCldQueue * kCldLanes];  // lane l's entries at l, l + 64, ...
  uint2 ranges[kCldRanges * kCldLanes];   // lane l's ranges at l, l + 64, ...
};

struct CldQueue {
  uint32_t* slot;   // this lane's first entry
  uint32_t* wp;     // next free entry
  uint32_t* room;   // last position of wp that leaves kCldStep free entries
};

// The shift of a wrapped cell, absWall * sign with sign +-2 on x and z (clouds.cl:334-343),
// from its code: bit 0 / 1 x = +2 / -2, bit 2 / 3 z = +2 / -2.
__device__ __forceinline__ float4 cld_shift(uint32_t code, const GridConsts& g)
{
  const float sx = (code & 1u) ? 2.0f : ((code & 2u) ? -2.0f : 0.0f);
  const float sz = (code & 4u) ? 2.0f : ((code & 8u) ? -2.0f : 0.0f);
  return make_float4(g.absWall[0] * sx, g.absWall[1] * 0.0f, g.absWall[2] * sz, 0.0f * 0.0f);
}

// The non-empty candidate ranges of the 27-cell walk of clouds.cl:322-353 in its order (x
// outer, z inner), as {first | shift code << 28, last}, every STRIDE-th entry of out from
// out[0]; returns their count.  A wrapped neighbour's position is shifted by the code's
// +-2 absWall; y cells beyond the walls are skipped.  The three cells of a z-column are one
// range when none of them wraps in z and their non-empty ranges are back to back (cells are
// numbered z-minor), which lists the same candidates in the same order.
template <uint32_t STRIDE = kCldLanes>
__device__ __forceinline__ uint32_t cld_ranges(const float4 pos, const uint2* __restrict__ se,
                                               const GridConsts& g, uint2* __restrict__ out)
{
  uint32_t ux, uy, uz;
  cell3d(pos, g, ux, uy, uz);
  const int32_t cx = (int32_t)ux, cy = (int32_t)uy, cz = (int32_t)uz;
  const int32_t rx = g.res[0], ry = g.res[1], rz = g.res[2];
  const bool zInside = cz - 1 >= 0 && cz + 1 < rz;
  uint32_t nr = 0;
  const auto push = [&](uint32_t lo, uint32_t hi, uint32_t code) {
    out[nr * STRIDE] = make_uint2(lo | (code << kCldIndexBits), hi);
    ++nr;
  };
#pragma unroll
  for (int ix = -1; ix <= 1; ++ix) {
    const int32_t ax = cx + ix;
    const int32_t nx = (ax + rx) % rx;
    const uint32_t xcode = ax >= rx ? 1u : (ax < 0 ? 2u : 0u);
#pragma unroll
    for (int iy = -1; iy <= 1; ++iy) {
      const int32_t ay = cy + iy;
      if (ay >= ry || ay < 0)
        continue;
      const int32_t ny = (ay + ry) % ry;
      const int32_t row = (nx * ry + ny) * rz;
      if (zInside) {
        const uint2 cells[3] = { se[row + cz - 1], se[row + cz], se[row + cz + 1] };
        uint32_t lo = 1u, hi = 0u;
        bool merged = true;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const uint2 r = cells[k];
          if (r.x > r.y)
            continue;
          if (lo > hi) {
            lo = r.x;
            hi = r.y;
          } else if (r.x == hi + 1u) {
            hi = r.y;
          } else {
            merged = false;
          }
        }
        if (merged) {
          if (lo <= hi)
            push(lo, hi, xcode);
        } else {
#pragma unroll
          for (int k = 0; k < 3; ++k)
            if (cells[k].x <= cells[k].y)
              push(cells[k].x, cells[k].y, xcode);
        }
        continue;
      }
#pragma unroll
      for (int iz = -1; iz <= 1; ++iz) {
        const int32_t az = cz + iz;
        const int32_t nz = (az + rz) % rz;
        const uint2 r = se[row + nz];
        if (r.x <= r.y)
          push(r.x, r.y, xcode | (az >= rz ? 4u : (az < 0 ? 8u : 0u)));
      }
    }
  }
  return nr;
}

// Survivor lists: a walk over some positions records, per particle, the candidates it
// kept in walk order; a later walk over the same positions and cell table replays them
// instead of scanning (the candidates it would skip contribute exactly +-0; see negligible
// below).  Entry j of particle i is entry[j * stride + i]; count[i] is the list's length, or
// kCldListOverflow when the walk kept more than kCldListCap.
constexpr uint32_t kCldListCap = 128;
constexpr uint32_t kCldListOverflow = 0xffffffffu;
struct CldList {
  uint32_t* entry = nullptr;
  uint32_t* count = nullptr;
  uint32_t stride = 0;
};

struct CldNoRecord {
  __device__ __forceinline__ void put(uint32_t) {}
};
struct CldRecord {
  uint32_t* at;  // &entry[i]
  uint32_t stride;
  uint32_t n;
  __device__ __forceinline__ void put(uint32_t v)
  {
    if (n < kCldListCap)
      at[n * stride] = v;
    ++n;
  }
};

// Run the lane's queued bodies in order, kCldDrain entries at a time: their data loaded
// together, then their bodies; each entry run is also handed to rec.
template <class L, class F, class R>
__device__ __forceinline__ void cld_drain(CldQueue& q, const float4 pos, const float4* __restrict__ src,
                                          const GridConsts& g, L& ld, F& f, R& rec)
{
  for (const uint32_t* rp = q.slot; rp < q.wp; rp += kCldDrain * kCldLanes) {
    uint32_t v[kCldDrain];
    float4 o[kCldDrain];
    decltype(ld(0u)) x[kCldDrain];
#pragma unroll
    for (uint32_t k = 0; k < kCldDrain; ++k) {
      const uint32_t* pk = rp + k * kCldLanes;
      v[k] = *(pk < q.wp ? pk : q.slot);  // in bounds either way
      o[k] = src[v[k] & kCldIndexMask];
      x[k] = ld(v[k] & kCldIndexMask);
    }
#pragma unroll
    for (uint32_t k = 0; k < kCldDrain; ++k)
      if (rp + k * kCldLanes < q.wp) {
        f(v[k] & kCldIndexMask, x[k], sub4(sub4(pos, o[k]), cld_shift(v[k] >> kCldIndexBits, g)));
        rec.put(v[k]);
      }
  }
  q.wp = q.slot;
}

// The walk: f(e, ld(e), vec) for every candidate e with keep(ld(e), vec), in the order of
// clouds.cl:322-353, vec = pos - p_e - the wrap shift; each candidate run is handed to rec.
// Called by every thread of a block of kCldLanes threads that has a particle (the others
// return before); threads with active false walk no cells.
template <class L, class K, class F, class R>
__device__ __forceinline__ void cld_walk(const float4 pos, const float4* __restrict__ src,
                                         const uint2* __restrict__ se, const GridConsts& g,
                                         CldLds& lds, bool active, L&& ld, K&& keep, F&& f, R& rec)
{
  const uint32_t lane = threadIdx.x;
  const uint2* rl = &lds.ranges[lane];
  const uint32_t nr = active ? cld_ranges(pos, se, g, &lds.ranges[lane]) : 0u;
  CldQueue q{ &lds.queue[lane], &lds.queue[lane],
              &lds.queue[lane + (kCldQueue - kCldStep) * kCldLanes] };
  uint32_t rk = 0, e = 1u, hi = 0u, tag = 0u;
  float4 shift = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const auto next_range = [&]() {
    if (rk < nr) {
      const uint2 r = rl[rk * kCldLanes];
      ++rk;
      e = r.x & kCldIndexMask;
      tag = r.x & ~kCldIndexMask;
      hi = r.y;
      shift = cld_shift(r.x >> kCldIndexBits, g);
    }
  };
  next_range();
  for (;;) {
    const bool on = e <= hi;
    if (__builtin_amdgcn_ballot_w64(on) == 0)
      break;
    if (on) {
      // candidates e..e+kCldStep-1 that are <= hi; the loads of the others repeat hi's
      float4 o[kCldStep];
      decltype(ld(e)) x[kCldStep];
#pragma unroll
      for (uint32_t k = 0; k < kCldStep; ++k) {
        const uint32_t ek = min(e + k, hi);
        o[k] = src[ek];
        x[k] = ld(ek);
      }
#pragma unroll
      for (uint32_t k = 0; k < kCldStep; ++k) {
        const bool kept = keep(x[k], sub4(sub4(pos, o[k]), shift));
        *q.wp = (e + k) | tag;
        q.wp += (kept & (e + k <= hi)) ? kCldLanes : 0u;
      }
      e += kCldStep;
      if (e > hi)
        next_range();
    }
    if (__builtin_amdgcn_ballot_w64(q.wp > q.room) != 0)
      cld_drain(q, pos, src, g, ld, f, rec);
  }
  cld_drain(q, pos, src, g, ld, f, rec);
}

// A recording walk: the walk, with the lane's list written to L.
template <class L, class K, class F>
__device__ __forceinline__ void cld_walk_record(uint32_t i, const float4 pos, const float4* __restrict__ src,
                                                const uint2* __restrict__ se, const GridConsts& g,
                                                CldLds& lds, const CldList& list, L&& ld, K&& keep, F&& f)
{
  CldRecord rec{ list.entry + i, list.stride, 0u };
  cld_walk(pos, src, se, g, lds, true, ld, keep, f, rec);
  list.count[i] = rec.n <= kCldListCap ? rec.n : kCldListOverflow;
}

// A replaying walk: f over the candidates of particle i's list, in order, kCldDrain at a
// time.  Lanes whose list overflowed, and every lane when useList is false (some data the
// bodies multiply may be non-finite or huge, see cld_flag), walk instead.
template <class L, class K, class F>
__device__ __forceinline__ void cld_walk_replay(uint32_t i, const float4 pos, const float4* __restrict__ src,
                                                const uint2* __restrict__ se, const GridConsts& g,
                                                CldLds& lds, const CldList& list, bool useList, L&& ld,
                                                K&& keep, F&& f)
{
  const uint32_t cnt = useList ? list.count[i] : kCldListOverflow;
  const bool walk = cnt == kCldListOverflow;
  if (__builtin_amdgcn_ballot_w64(walk) != 0) {
    CldNoRecord none;
    cld_walk(pos, src, se, g, lds, walk, ld, keep, f, none);
  }
  if (walk)
    return;
  const uint32_t* at = list.entry + i;
  for (uint32_t j = 0; j < cnt; j += kCldDrain) {
    uint32_t v[kCldDrain];
    float4 o[kCldDrain];
    decltype(ld(0u)) x[kCldDrain];
#pragma unroll
    for (uint32_t k = 0; k < kCldDrain; ++k) {
      v[k] = at[min(j + k, cnt - 1u) * list.stride];
      o[k] = src[v[k] & kCldIndexMask];
      x[k] = ld(v[k] & kCldIndexMask);
    }
#pragma unroll
    for (uint32_t k = 0; k < kCldDrain; ++k)
      if (j + k < cnt)
        f(v[k] & kCldIndexMask, x[k], sub4(sub4(pos, o[k]), cld_shift(v[k] >> kCldIndexBits, g)));
  }
}

// The data a replaying walk's bodies multiply (lambdaT, lambda, vorticity, the confined
// velocity) come from the pass before it.  That pass sets bit `bit` of *flags when any of
// its outputs has a component that is not finite or is 2^60 or more in magnitude: then a
// candidate left out of a list may not contribute +-0 (for finite data below 2^60 the sums
// and differences the skip tests evaluate are finite, and so is the length of a 4-vector),
// and the replaying pass walks.  The flags are cleared at the start of every step.
__device__ __forceinline__ bool cld_safe(float x) { return fabsf(x) < 0x1p60f; }
__device__ __forceinline__ bool cld_safe(float4 v)
{
  return cld_safe(v.x) & cld_safe(v.y) & cld_safe(v.z) & cld_safe(v.w);
}
__device__ __forceinline__ void cld_flag(bool safe, uint32_t* flags, uint32_t bit)
{
  const uint64_t bad = __builtin_amdgcn_ballot_w64(!safe);
  if (bad != 0 && (threadIdx.x & 63u) == (uint32_t)__builtin_ctzll(bad))
    atomicOr(flags, bit);
}
constexpr uint32_t kFlagLambdaT = 1u, kFlagVort = 2u, kFlagVelTmp = 4u;
constexpr uint32_t kFlagLambda0 = 8u;   // iteration k: kFlagLambda0 << k for k < kFlagIters
constexpr uint32_t kFlagIters = 29u;

// A candidate at r >= h (d2 >= cullSq, so sqrt(d2) >= h) with a finite distance contributes
// exactly +-0 to every sum here, as long as the neighbour data its term multiplies are finite:
// its kernel weight and gradient are +-0 (the zero factor 1 - step(h, r) times finite
// values), and a sum that starts at +0 never holds -0, so adding +-0 leaves it unchanged.
// Such candidates are skipped; every other one (r < h, a NaN or huge distance, non-finite
// data) runs the reference's full expression.
__device__ __forceinline__ bool negligible(float4 vec, const PbfConsts& c)
{
  const float d2 = dot4(vec, vec);
  return (d2 >= c.cullSq) & (d2 < kFarD2);
}

__device__ __forceinline__ bool finite4(float4 v)
{
  return __builtin_isfinite(v.x) & __builtin_isfinite(v.y) & __builtin_isfinite(v.z) &
         __builtin_isfinite(v.w);
}

inline uint32_t blocks_for(uint32_t n, uint32_t b) { return (n + b - 1u) / b; }

// ------------------------------------------------------------------------ quad walks
// With quads (the default for scenes up to kQuadDefaultMaxN particles) each walk kernel gives a
// particle four lanes of a 64-lane workgroup.  The quad builds the particle's ranges
// (cld_ranges) in LDS and scans them sixteen candidates per step, lane u taking candidates
// u, u + 4, u + 8 and u + 12 with the pass's keep test; the kept ones go to a queue of
// kCqQueue entries per particle in LDS, in walk order.  A drain runs four queued bodies at
// once, one per lane, and adds the four terms into every sum in walk order (quad_add), so
// each sum adds the same terms in the same order as cld_walk's one-lane walk; lanes
// without an entry add +0, which leaves a sum started at +0 unchanged.  Survivor lists keep
// cld_walk's layout and codes, so either kind of walk replays the other's lists.
constexpr uint32_t kCqGroups = kCldLanes / 4u;  // particles per workgroup
constexpr uint32_t kCqQueue = 64;               // queued survivors per particle

struct CqLds {
  uint32_t queue[kCqGroups][kCqQueue];
  uint2 ranges[kCqGroups][kCldRanges];
};

// The body over entries 0..cnt-1 of the quad (entry(k): entry k, an index with its shift
// code), four at a time, loading each round's data during the previous round's bodies;
// rec(v, k, code) sees every round first.  The wave loops until its longest quad is done;
// lanes without an entry run the body on the particle itself and add +0.
template <class S, class R, class B>
__device__ __forceinline__ void cq_run(uint32_t cnt, uint32_t self, const float4 pos,
                                       const float4* __restrict__ src, const GridConsts& g,
                                       S&& entry, R&& rec, B& b)
{
  const uint32_t u = threadIdx.x & 3u;
  bool v = u < cnt;
  uint32_t code = v ? entry(u) : self;
  float4 o = src[code & kCldIndexMask];
  auto x = b.ld(code & kCldIndexMask);
  for (uint32_t k = u; __builtin_amdgcn_ballot_w64(k - u < cnt) != 0; k += 4u) {
    const uint32_t k1 = k + 4u;
    const bool v1 = k1 < cnt;
    const uint32_t code1 = v1 ? entry(k1) : self;
    const float4 o1 = src[code1 & kCldIndexMask];
    const auto x1 = b.ld(code1 & kCldIndexMask);
    rec(v, k, code);
    b.add(v, x, sub4(sub4(pos, o), cld_shift(code >> kCldIndexBits, g)));
    v = v1;
    code = code1;
    o = o1;
    x = x1;
  }
}

// The quad's walk of particle `self` over its nr ranges rl (in LDS), queued in q; with REC
// (and `record`) the kept candidates go to the particle's survivor list.
template <bool REC, class B>
__device__ __forceinline__ void cq_walk(uint32_t nr, const uint2* rl, uint32_t* q, uint32_t self,
                                        const float4 pos, const float4* __restrict__ src,
                                        const GridConsts& g, B& b, const CldList& list,
                                        bool record)
{
  const uint32_t u = threadIdx.x & 3u, below = (1u << u) - 1u;
  uint32_t ri = 0, e = 1u, hi = 0u, tag = 0u, qn = 0, tot = 0;
  float4 shift = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const auto range = [&](uint32_t k) {
    const uint2 r = rl[k];
    e = r.x & kCldIndexMask;
    tag = r.x & ~kCldIndexMask;
    hi = r.y;
    shift = cld_shift(r.x >> kCldIndexBits, g);
  };
  if (nr > 0)
    range(0);
  const auto drain = [&]() {
    wave_sync();
    cq_run(qn, self, pos, src, g, [&](uint32_t k) { return q[k]; },
           [&](bool v, uint32_t k, uint32_t code) {
             if (REC && v & record & (tot + k < kCldListCap))
               list.entry[(tot + k) * list.stride + self] = code;
           }, b);
    tot += qn;
    qn = 0;
    wave_sync();
  };
  while (__builtin_amdgcn_ballot_w64(ri < nr) != 0) {
    const bool act = ri < nr;
    uint32_t code[4];
    bool kept[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t ej = e + u + 4u * (uint32_t)j;
      const bool valid = act & (ej <= hi);
      const uint32_t at = valid ? ej : self;
      const float4 o = src[at];
      kept[j] = valid & b.keep(b.ld(at), sub4(sub4(pos, o), shift));
      code[j] = ej | tag;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t m = quad_bits(kept[j]);
      if (kept[j])
        q[qn + (uint32_t)__builtin_popcount(m & below)] = code[j];
      qn += (uint32_t)__builtin_popcount(m);
    }
    e += 16u;
    if (act & (e > hi)) {
      ++ri;
      if (ri < nr)
        range(ri);
    }
    if (__builtin_amdgcn_ballot_w64(qn > kCqQueue - 16u) != 0)
      drain();
  }
  drain();
  if (REC && record && u == 0u)
    list.count[self] = tot <= kCldListCap ? tot : kCldListOverflow;
}

// The candidates of particle i into the body b: a recording walk (MODE kCqRecord), a
// replay of its list when useList and the list is complete, else a walk (kCqReplay), or a
// walk (kCqWalk).
constexpr int kCqWalk = 0, kCqRecord = 1, kCqReplay = 2;
template <int MODE, class B>
__device__ __forceinline__ void cq_neighbours(bool on, uint32_t i, bool useList, const float4 pos,
                                              const float4* __restrict__ src,
                                              const uint2* __restrict__ se, const GridConsts& g,
                                              const CldList& list, CqLds& lds, B& b)
{
  const uint32_t gq = threadIdx.x >> 2;
  bool listed = false;
  if constexpr (MODE == kCqReplay) {
    const uint32_t cnt = (on & useList) ? list.count[i] : kCldListOverflow;
    listed = cnt != kCldListOverflow;
    cq_run(listed ? cnt : 0u, i, pos, src, g,
           [&](uint32_t m) { return list.entry[m * list.stride + i]; },
           [](bool, uint32_t, uint32_t) {}, b);
  }
  const uint32_t nr = (on & !listed) ? cld_ranges<1>(pos, se, g, lds.ranges[gq]) : 0u;
  // a recording walk writes every recorded count, also of a particle without candidates
  if (MODE == kCqRecord || __builtin_amdgcn_ballot_w64(nr > 0u) != 0)
    cq_walk<MODE == kCqRecord>(nr, lds.ranges[gq], lds.queue[gq], i, pos, src, g, b, list, on);
}

// --------------------------------------------------------------------------------- kernels

// cld_initTemperature (clouds.cl:116-121), cld_initVaporDensity (:127-133), and the host's
// zero cloud density and particle ids (Clouds.cpp:485-494), over all maxN slots.
__global__ __launch_bounds__(256) void k_cld_init(const float4* __restrict__ pos,
                                                  float* __restrict__ temp,
                                                  float* __restrict__ vapor,
                                                  float* __restrict__ cloud,
                                                  float* __restrict__ pid, uint32_t count,
                                                  CloudConsts c)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count)
    return;
  const float t = env_temp(pos[i].y, c.f.g);
  temp[i] = t;
  vapor[i] = c.q.initVaporDensityCoeff * svd(t);
  cloud[i] = 0.0f;
  pid[i] = (float)i;
}

// Clouds.cpp:516-541: the thermodynamics (clouds.cl:159-251), cld_predictPosition (:257-273),
// cld_applyMixedBoundaryConditions (:279-300) and fillCellIDs (grid.cl:76-86) of particle t;
// resetStartEndCell (grid.cl:91-96) over the cells.  Intermediate buffers of the reference
// (p_tempIn, p_cloudGen, p_vaporDensIn, p_cloudDensIn) stay in registers.
__global__ __launch_bounds__(256) void k_cld_thermo(const float4* __restrict__ pos,
                                                    const float4* __restrict__ vel,
                                                    float* __restrict__ temp,
                                                    float* __restrict__ vapor,
                                                    float* __restrict__ cloud,
                                                    float* __restrict__ buoyancy,
                                                    float4* __restrict__ pred,
                                                    float4* __restrict__ totCorr,
                                                    uint32_t* __restrict__ key,
                                                    uint2* __restrict__ startEnd, uint32_t n,
                                                    CloudConsts c, uint32_t* __restrict__ flags)
{
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < c.f.g.numCells)
    startEnd[t] = make_uint2(1u, 0u);
  if (t == 0)
    *flags = 0u;  // the survivor-list flags of this step's passes
  if (t >= n)
    return;
  const rtp_cloud_params& q = c.q;
  const GridConsts& g = c.f.g;
  const float4 p = pos[t], v = vel[t];
  const float tempIn = temp[t];
  // cld_heatFromGround
  const float heated = fminf(tempIn + ext_heat(p.y, g) * q.groundHeatCoeff * q.timeStep, 313.0f);
  // cld_computeBuoyancy
  const float envTemp = env_temp(p.y, g);
  const float cd = cloud[t];
  const float buoy = q.buoyancyCoeff * (heated - envTemp) / envTemp - q.gravCoeff * c.f.gravity * cd;
  // cld_applyAdiabaticCooling
  const float cooled = fmaxf(heated - q.adiabaticLapseRate * v.y * q.timeStep, 223.0f);
  // cld_generateCloud
  const float vd = vapor[t];
  const float gen = q.phaseTransitionRate * (vd - svd(cooled));
  // cld_applyPhaseTransition
  cloud[t] = fmaxf(cd + gen * q.timeStep, 0.0f);
  vapor[t] = fmaxf(vd - gen * q.timeStep, 0.0f);
  // cld_applyLatentHeat
  temp[t] = cooled + fmaxf(q.latentHeatCoeff * gen * q.timeStep, 0.0f);
  buoyancy[t] = buoy;
  // cld_predictPosition
  const float4 predVel = add4(v, mul4s(make_float4(0.0f, buoy, 0.0f, 0.0f), q.timeStep));
  totCorr[t] = mul4s(predVel, q.timeStep);
  const float4 pp = mixed_boundary(add4(p, mul4s(predVel, q.timeStep)), g);
  pred[t] = pp;
  key[t] = cell1d(pp, g);
}

// The payloads the cell sort permutes besides p_pos, p_vel and p_predPos (k_reorder):
// p_totCorrPos and the float buffers p_temp, p_buoyancy, p_vaporDens, p_cloudDens, p_partID
// (Clouds.cpp:543), from [0] into [1] by the permutation, over [0, sortN).
struct CloudGather {
  const float4* tot[2];
  float4* totOut;
  const float* src[5];
  float* dst[5];
};
__global__ __launch_bounds__(256) void k_cld_gather(const uint32_t* __restrict__ idx, uint32_t sortN,
                                                    CloudGather a)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sortN)
    return;
  const uint32_t j = idx[i];
  a.totOut[i] = a.tot[0][j];
#pragma unroll
  for (int k = 0; k < 5; ++k)
    a.dst[k][i] = a.src[k][j];
}

// The outputs of the walk kernels from their sums (shared by the one-lane and quad kernels);
// those the next pass multiplies are returned for cld_flag.
__device__ __forceinline__ float cld_temp_lambda_out(uint32_t i, float laplacian, float sumGradCi,
                                                     float sumSqGradC, float* __restrict__ lap,
                                                     float* __restrict__ lambdaT,
                                                     const CloudConsts& c)
{
  const float l = laplacian / c.q.restDensity;
  lap[i] = l;
  sumSqGradC = sumSqGradC + sumGradCi * sumGradCi;
  const float lt = -l / (sumSqGradC + c.q.relaxCFM);
  lambdaT[i] = lt;
  return lt;
}

__device__ __forceinline__ float cld_density_lambda_out(uint32_t i, float fluidDensity,
                                                        float4 sumGradCi, float sumSqGradC,
                                                        float* __restrict__ density,
                                                        float* __restrict__ lambda,
                                                        const CloudConsts& c)
{
  density[i] = fluidDensity;
  const rtp_fluid_params& p = c.f.p;
  const float densityC = fluidDensity / p.restDensity - 1.0f;
  sumSqGradC = sumSqGradC + dot4(sumGradCi, sumGradCi);
  sumSqGradC = sumSqGradC / (p.restDensity * p.restDensity);
  const float li = -densityC / (sumSqGradC + p.relaxCFM);
  lambda[i] = li;
  return li;
}

__device__ __forceinline__ void cld_correct_out(uint32_t i, const float4 pos, const float4 corr,
                                                float4* __restrict__ predOut,
                                                float4* __restrict__ totCorr,
                                                float4* __restrict__ vel, const CloudConsts& c)
{
  const float4 corrPos = div4s(corr, c.f.p.restDensity);
  predOut[i] = mixed_boundary(add4(pos, corrPos), c.f.g);
  const float4 tc = add4(totCorr[i], corrPos);
  totCorr[i] = tc;
  if (vel) {
    const float m = c.f.maxVel;
    const float4 u = div4s(tc, c.f.p.timeStep + kFloatEps);
    vel[i] = make_float4(clampf(u.x, -m, m), clampf(u.y, -m, m), clampf(u.z, -m, m),
                         clampf(u.w, -m, m));
  }
}

__device__ __forceinline__ float4 cld_confinement_out(uint32_t i, const float4 nn,
                                                      const float4 vorticity,
                                                      const float4* __restrict__ vel,
                                                      float4* __restrict__ velOut,
                                                      const CloudConsts& c)
{
  const rtp_fluid_params& p = c.f.p;
  const float4 vo = add4(vel[i], mul4s(muls4(p.vorticityConfCoeff, cross4(normalize4(nn), vorticity)), p.timeStep));
  velOut[i] = vo;
  return vo;
}

// cld_computeLaplacianTemp (clouds.cl:508-569) and cld_computeConstraintFactorTemp
// (:575-648) in one walk over p_pos: the two sums are independent and each keeps the
// reference's order.
__global__ __launch_bounds__(kCldLanes) void k_cld_temp_lambda(const float4* __restrict__ posP,
                                                         const float* __restrict__ tempP,
                                                         const uint2* __restrict__ se,
                                                         float* __restrict__ lap,
                                                         float* __restrict__ lambdaT, uint32_t n,
                                                         CloudConsts c, CldList list,
                                                         uint32_t* __restrict__ flags)
{
  const uint32_t i = logical_block(1) * kCldLanes + threadIdx.x;
  if (i >= n)
    return;
  const float4 pos = posP[i];
  const float temp = tempP[i];
  float laplacian = 0.0f, sumGradCi = 0.0f, sumSqGradC = 0.0f;
  const float rho0 = c.q.restDensity;
  __shared__ CldLds lds;
  const auto keep = [&](float te, float4 vec) {
    return !(negligible(vec, c.f) & __builtin_isfinite(temp - te));
  };
  cld_walk_record(i, pos, posP, se, c.f.g, lds, list, [&](uint32_t e) { return tempP[e]; }, keep, [&](uint32_t, float te, float4 vec) {
    const float dTemp = temp - te;
    const float vv = dot4(vec, vec);
    const float4 grad = grad_spiky_fast(vec, len_of(vv), c.f);
    const float vg = dot4(vec, grad);
    laplacian = laplacian + dTemp * vg / (vv + kFloatEps);
    const float derivativeTemp = vg / (vv * rho0 + kFloatEps);
    sumGradCi = sumGradCi + derivativeTemp;
    sumSqGradC = sumSqGradC + derivativeTemp * derivativeTemp;
  });
  const float lt = cld_temp_lambda_out(i, laplacian, sumGradCi, sumSqGradC, lap, lambdaT, c);
  cld_flag(cld_safe(lt), flags, kFlagLambdaT);
}

// cld_computeConstraintCorrectionTemp (clouds.cl:654-721) and cld_correctTemperature
// (:931-935): temp += 0.3 * correction.
__global__ __launch_bounds__(kCldLanes) void k_cld_temp_correct(const float4* __restrict__ posP,
                                                          const float* __restrict__ lambdaT,
                                                          const uint2* __restrict__ se,
                                                          float* __restrict__ temp, uint32_t n,
                                                          CloudConsts c, CldList list,
                                                          const uint32_t* __restrict__ flags, int useList)
{
  const uint32_t i = logical_block(1) * kCldLanes + threadIdx.x;
  if (i >= n)
    return;
  const float4 pos = posP[i];
  const float lambdaI = lambdaT[i];
  const float rho0 = c.q.restDensity;
  float corr = 0.0f;
  __shared__ CldLds lds;
  const auto keep = [&](float le, float4 vec) {
    return !(negligible(vec, c.f) & __builtin_isfinite(lambdaI + le));
  };
  cld_walk_replay(i, pos, posP, se, c.f.g, lds, list, useList && !(*flags & kFlagLambdaT),
                  [&](uint32_t e) { return lambdaT[e]; }, keep, [&](uint32_t, float le, float4 vec) {
    const float l = lambdaI + le;
    const float vv = dot4(vec, vec);
    const float4 grad = grad_spiky_fast(vec, len_of(vv), c.f);
    const float derivativeTemp = dot4(vec, grad) / (vv * rho0 + kFloatEps);
    corr = corr + l * derivativeTemp;
  });
  temp[i] = temp[i] + 0.3f * corr;
}

// cld_computeDensity (clouds.cl:305-359) + cld_computeConstraintFactor (:365-435).
__global__ __launch_bounds__(kCldLanes) void k_cld_density_lambda(const float4* __restrict__ pred,
                                                            const uint2* __restrict__ se,
                                                            float* __restrict__ density,
                                                            float* __restrict__ lambda, uint32_t n,
                                                            CloudConsts c, CldList list,
                                                            uint32_t* __restrict__ flags, uint32_t bit)
{
  const uint32_t i = logical_block(1) * kCldLanes + threadIdx.x;
  if (i >= n)
    return;
  const float4 pos = pred[i];
  // the density sum and the gradient sums are independent: one walk, each in its order
  float fluidDensity = 0.0f;
  float4 sumGradCi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float sumSqGradC = 0.0f;
  __shared__ CldLds lds;
  const auto keep = [&](int, float4 vec) { return !negligible(vec, c.f); };
  cld_walk_record(i, pos, pred, se, c.f.g, lds, list, [](uint32_t) { return 0; }, keep, [&](uint32_t, int, float4 vec) {
    const float r = len_of(dot4(vec, vec));
    fluidDensity = fluidDensity + poly6_r(r, c.f);
    const float4 grad = grad_spiky_fast(vec, r, c.f);
    sumGradCi = add4(sumGradCi, grad);
    sumSqGradC = sumSqGradC + dot4(grad, grad);
  });
  const float li = cld_density_lambda_out(i, fluidDensity, sumGradCi, sumSqGradC, density, lambda, c);
  if (bit)
    cld_flag(cld_safe(li), flags, bit);
}

// cld_computeConstraintCorrection (clouds.cl:441-502), cld_correctPosition on p_predPos and
// p_totCorrPos (:920-925) and cld_applyMixedBoundaryConditions (Clouds.cpp:573-586); on the
// last iteration also cld_updateVel (clouds.cl:958-966).
__global__ __launch_bounds__(kCldLanes) void k_cld_correct(const float4* __restrict__ pred,
                                                     const float* __restrict__ lambda,
                                                     const uint2* __restrict__ se,
                                                     float4* __restrict__ predOut,
                                                     float4* __restrict__ totCorr,
                                                     float4* __restrict__ vel, uint32_t n,
                                                     CloudConsts c, CldList list,
                                                     const uint32_t* __restrict__ flags, uint32_t bit,
                                                     int useList)
{
  const uint32_t i = logical_block(1) * kCldLanes + threadIdx.x;
  if (i >= n)
    return;
  const float4 pos = pred[i];
  const float lambdaI = lambda[i];
  const float ywdq = (c.f.shortMath & kShortArtDiv) ? rcp_refined(c.f.wdq) : 0.0f;
  float4 corr = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  __shared__ CldLds lds;
  const auto keep = [&](float le, float4 vec) {
    return !(negligible(vec, c.f) & __builtin_isfinite(lambdaI + le));
  };
  cld_walk_replay(i, pos, pred, se, c.f.g, lds, list, useList && bit && !(*flags & bit),
                  [&](uint32_t e) { return lambda[e]; }, keep, [&](uint32_t, float le, float4 vec) {
    const float r = len_of(dot4(vec, vec));
    const float s = lambdaI + le + art_pressure_w(poly6_r(r, c.f), ywdq, c.f);
    corr = add4(corr, muls4(s, grad_spiky_fast(vec, r, c.f)));
  });
  cld_correct_out(i, pos, corr, predOut, totCorr, vel, c);
}

// cld_updateVel alone (a step without Jacobi iterations).
__global__ __launch_bounds__(256) void k_cld_update_vel(const float4* __restrict__ totCorr,
                                                        float4* __restrict__ vel, uint32_t n,
                                                        CloudConsts c)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const float m = c.f.maxVel;
  const float4 u = div4s(totCorr[i], c.f.p.timeStep + kFloatEps);
  vel[i] = make_float4(clampf(u.x, -m, m), clampf(u.y, -m, m), clampf(u.z, -m, m),
                       clampf(u.w, -m, m));
}

// cld_computeVorticity (clouds.cl:727-784).
__global__ __launch_bounds__(kCldLanes) void k_cld_vorticity(const float4* __restrict__ pred,
                                                       const uint2* __restrict__ se,
                                                       const float4* __restrict__ vel,
                                                       float4* __restrict__ vort, uint32_t n,
                                                       CloudConsts c, CldList list,
                                                       uint32_t* __restrict__ flags)
{
  const uint32_t i = logical_block(1) * kCldLanes + threadIdx.x;
  if (i >= n)
    return;
  const float4 pos = pred[i];
  const float4 velocity = vel[i];
  float4 w = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  __shared__ CldLds lds;
  const auto keep = [&](float4 ve, float4 vec) { return !(negligible(vec, c.f) & finite4(sub4(ve, velocity))); };
  cld_walk_record(i, pos, pred, se, c.f.g, lds, list, [&](uint32_t e) { return vel[e]; }, keep, [&](uint32_t, float4 ve, float4 vec) {
    const float4 dv = sub4(ve, velocity);
    w = add4(w, cross4(dv, grad_spiky_fast(vec, len_of(dot4(vec, vec)), c.f)));
  });
  vort[i] = w;
  cld_flag(cld_safe(w), flags, kFlagVort);
}

// cld_applyVorticityConfinement (clouds.cl:790-850), written to the copy XSPH reads
// (p_velInViscosity, Clouds.cpp:597-599).
__global__ __launch_bounds__(kCldLanes) void k_cld_confinement(const float4* __restrict__ pred,
                                                         const uint2* __restrict__ se,
                                                         const float4* __restrict__ vort,
                                                         const float4* __restrict__ vel,
                                                         float4* __restrict__ velOut, uint32_t n,
                                                         CloudConsts c, CldList list,
                                                         uint32_t* __restrict__ flags, int useList)
{
  const uint32_t i = logical_block(1) * kCldLanes + threadIdx.x;
  if (i >= n)
    return;
  const float4 pos = pred[i];
  const float4 vorticity = vort[i];
  float4 nn = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  __shared__ CldLds lds;
  const auto keep = [&](float4 we, float4 vec) {
    // sqrt keeps finiteness: isfinite(length(we)) == isfinite(dot4(we, we))
    return !(negligible(vec, c.f) & __builtin_isfinite(dot4(we, we)));
  };
  cld_walk_replay(i, pos, pred, se, c.f.g, lds, list, useList && !(*flags & kFlagVort),
                  [&](uint32_t e) { return vort[e]; }, keep, [&](uint32_t, float4 we, float4 vec) {
    const float lw = len_of(dot4(we, we));
    nn = add4(nn, muls4(lw, grad_spiky_fast(vec, len_of(dot4(vec, vec)), c.f)));
  });
  const float4 vo = cld_confinement_out(i, nn, vorticity, vel, velOut, c);
  cld_flag(cld_safe(vo), flags, kFlagVelTmp);
}

// cld_updatePosition (clouds.cl:942-952): the predicted position plus the wind.
__device__ __forceinline__ float4 cld_update_pos(float4 pp, const CloudConsts& c)
{
  const float wy = c.f.g.absWall[1];
  const rtp_cloud_params& q = c.q;
  float4 o = pp;
  o.x = o.x + (1 - exp32(-(pp.y + wy) * 0.2f)) * q.windCoeff * q.timeStep * (float)(q.dim - 2);
  o.z = o.z + (1 - exp32(-(pp.y + wy) * 0.3f)) * 0.7f * q.windCoeff * q.timeStep;
  return o;
}

// cld_applyXsphViscosityCorrection (clouds.cl:856-914) + cld_updatePosition.
__global__ __launch_bounds__(kCldLanes) void k_cld_xsph(const float4* __restrict__ pred,
                                                  const uint2* __restrict__ se,
                                                  const float4* __restrict__ velIn,
                                                  float4* __restrict__ velOut,
                                                  float4* __restrict__ pos, uint32_t n,
                                                  CloudConsts c, CldList list,
                                                  const uint32_t* __restrict__ flags, int useList)
{
  const uint32_t i = logical_block(1) * kCldLanes + threadIdx.x;
  if (i >= n)
    return;
  const float4 pp = pred[i];
  const float4 velocity = velIn[i];
  float4 viscosity = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  __shared__ CldLds lds;
  const auto keep = [&](float4 ve, float4 vec) { return !(negligible(vec, c.f) & finite4(sub4(ve, velocity))); };
  cld_walk_replay(i, pp, pred, se, c.f.g, lds, list, useList && !(*flags & kFlagVelTmp),
                  [&](uint32_t e) { return velIn[e]; }, keep, [&](uint32_t, float4 ve, float4 vec) {
    const float4 dv = sub4(ve, velocity);
    viscosity = add4(viscosity, mul4s(dv, poly6_r(len_of(dot4(vec, vec)), c.f)));
  });
  velOut[i] = add4(velocity, muls4(c.f.p.xsphViscosityCoeff, viscosity));
  pos[i] = cld_update_pos(pp, c);
}

__global__ __launch_bounds__(256) void k_cld_update_pos(const float4* __restrict__ pred,
                                                        float4* __restrict__ pos, uint32_t n,
                                                        CloudConsts c)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    pos[i] = cld_update_pos(pred[i], c);
}

// --------------------------------------------------------------- quad walk kernels
// The walk kernels above with a quad of lanes per particle (cq_neighbours): the same sums in
// the same order, the same survivor lists and flags.

// s + t component by component in quad order (t counts only when v).
__device__ __forceinline__ float4 quad_add4(float4 s, bool v, const float4 t)
{
  return make_float4(quad_add(s.x, v ? t.x : 0.0f), quad_add(s.y, v ? t.y : 0.0f),
                     quad_add(s.z, v ? t.z : 0.0f), quad_add(s.w, v ? t.w : 0.0f));
}

struct CqTempLambda {
  const CloudConsts& c;
  const float* __restrict__ tempP;
  float temp, rho0;
  float laplacian = 0.0f, sumGradCi = 0.0f, sumSqGradC = 0.0f;
  __device__ __forceinline__ float ld(uint32_t e) const { return tempP[e]; }
  __device__ __forceinline__ bool keep(float te, float4 vec) const
  {
    return !(negligible(vec, c.f) & __builtin_isfinite(temp - te));
  }
  __device__ __forceinline__ void add(bool v, float te, float4 vec)
  {
    const float dTemp = temp - te;
    const float vv = dot4(vec, vec);
    const float4 grad = grad_spiky_fast(vec, len_of(vv), c.f);
    const float vg = dot4(vec, grad);
    const float lapTerm = dTemp * vg / (vv + kFloatEps);
    const float derivativeTemp = vg / (vv * rho0 + kFloatEps);
    laplacian = quad_add(laplacian, v ? lapTerm : 0.0f);
    sumGradCi = quad_add(sumGradCi, v ? derivativeTemp : 0.0f);
    sumSqGradC = quad_add(sumSqGradC, v ? derivativeTemp * derivativeTemp : 0.0f);
  }
};

struct CqTempCorrect {
  const CloudConsts& c;
  const float* __restrict__ lambdaT;
  float lambdaI, rho0;
  float corr = 0.0f;
  __device__ __forceinline__ float ld(uint32_t e) const { return lambdaT[e]; }
  __device__ __forceinline__ bool keep(float le, float4 vec) const
  {
    return !(negligible(vec, c.f) & __builtin_isfinite(lambdaI + le));
  }
  __device__ __forceinline__ void add(bool v, float le, float4 vec)
  {
    const float l = lambdaI + le;
    const float vv = dot4(vec, vec);
    const float4 grad = grad_spiky_fast(vec, len_of(vv), c.f);
    const float derivativeTemp = dot4(vec, grad) / (vv * rho0 + kFloatEps);
    corr = quad_add(corr, v ? l * derivativeTemp : 0.0f);
  }
};

struct CqDensity {
  const CloudConsts& c;
  float fluidDensity = 0.0f;
  float4 sumGradCi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float sumSqGradC = 0.0f;
  __device__ __forceinline__ int ld(uint32_t) const { return 0; }
  __device__ __forceinline__ bool keep(int, float4 vec) const { return !negligible(vec, c.f); }
  __device__ __forceinline__ void add(bool v, int, float4 vec)
  {
    const float r = len_of(dot4(vec, vec));
    const float w = poly6_r(r, c.f);
    const float4 grad = grad_spiky_fast(vec, r, c.f);
    const float sq = dot4(grad, grad);
    fluidDensity = quad_add(fluidDensity, v ? w : 0.0f);
    sumGradCi = quad_add4(sumGradCi, v, grad);
    sumSqGradC = quad_add(sumSqGradC, v ? sq : 0.0f);
  }
};

struct CqCorrect {
  const CloudConsts& c;
  const float* __restrict__ lambda;
  float lambdaI, ywdq;
  float4 corr = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  __device__ __forceinline__ float ld(uint32_t e) const { return lambda[e]; }
  __device__ __forceinline__ bool keep(float le, float4 vec) const
  {
    return !(negligible(vec, c.f) & __builtin_isfinite(lambdaI + le));
  }
  __device__ __forceinline__ void add(bool v, float le, float4 vec)
  {
    const float r = len_of(dot4(vec, vec));
    const float s = lambdaI + le + art_pressure_w(poly6_r(r, c.f), ywdq, c.f);
    corr = quad_add4(corr, v, muls4(s, grad_spiky_fast(vec, r, c.f)));
  }
};

struct CqVorticity {
  const CloudConsts& c;
  const float4* __restrict__ vel;
  float4 velocity;
  float4 w = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  __device__ __forceinline__ float4 ld(uint32_t e) const { return vel[e]; }
  __device__ __forceinline__ bool keep(float4 ve, float4 vec) const
  {
    return !(negligible(vec, c.f) & finite4(sub4(ve, velocity)));
  }
  __device__ __forceinline__ void add(bool v, float4 ve, float4 vec)
  {
    const float4 dv = sub4(ve, velocity);
    w = quad_add4(w, v, cross4(dv, grad_spiky_fast(vec, len_of(dot4(vec, vec)), c.f)));
  }
};

struct CqConfinement {
  const CloudConsts& c;
  const float4* __restrict__ vort;
  float4 nn = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  __device__ __forceinline__ float4 ld(uint32_t e) const { return vort[e]; }
  __device__ __forceinline__ bool keep(float4 we, float4 vec) const
  {
    return !(negligible(vec, c.f) & __builtin_isfinite(dot4(we, we)));
  }
  __device__ __forceinline__ void add(bool v, float4 we, float4 vec)
  {
    const float lw = len_of(dot4(we, we));
    nn = quad_add4(nn, v, muls4(lw, grad_spiky_fast(vec, len_of(dot4(vec, vec)), c.f)));
  }
};

struct CqXsph {
  const CloudConsts& c;
  const float4* __restrict__ velIn;
  float4 velocity;
  float4 viscosity = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  __device__ __forceinline__ float4 ld(uint32_t e) const { return velIn[e]; }
  __device__ __forceinline__ bool keep(float4 ve, float4 vec) const
  {
    return !(negligible(vec, c.f) & finite4(sub4(ve, velocity)));
  }
  __device__ __forceinline__ void add(bool v, float4 ve, float4 vec)
  {
    const float4 dv = sub4(ve, velocity);
    viscosity = quad_add4(viscosity, v, mul4s(dv, poly6_r(len_of(dot4(vec, vec)), c.f)));
  }
};

// The quad's particle (quads past n stand in for particle 0 and write nothing).
#define RTP_CQ_PROLOGUE                                                      \
  __shared__ CqLds lds;                                                      \
  const uint32_t q0 = logical_block(1) * kCqGroups + (threadIdx.x >> 2);     \
  const bool on = q0 < n;                                                    \
  const uint32_t i = on ? q0 : 0u;                                           \
  const bool lead = on && (threadIdx.x & 3u) == 0u;

__global__ __launch_bounds__(kCldLanes) void k_cld_temp_lambda_q(const float4* __restrict__ posP,
                                                                 const float* __restrict__ tempP,
                                                                 const uint2* __restrict__ se,
                                                                 float* __restrict__ lap,
                                                                 float* __restrict__ lambdaT,
                                                                 uint32_t n, CloudConsts c,
                                                                 CldList list,
                                                                 uint32_t* __restrict__ flags)
{
  RTP_CQ_PROLOGUE
  const float4 pos = posP[i];
  CqTempLambda b{ c, tempP, tempP[i], c.q.restDensity };
  cq_neighbours<kCqRecord>(on, i, true, pos, posP, se, c.f.g, list, lds, b);
  float lt = 0.0f;
  if (lead)
    lt = cld_temp_lambda_out(i, b.laplacian, b.sumGradCi, b.sumSqGradC, lap, lambdaT, c);
  cld_flag(!lead | cld_safe(lt), flags, kFlagLambdaT);
}

__global__ __launch_bounds__(kCldLanes) void k_cld_temp_correct_q(const float4* __restrict__ posP,
                                                                  const float* __restrict__ lambdaT,
                                                                  const uint2* __restrict__ se,
                                                                  float* __restrict__ temp,
                                                                  uint32_t n, CloudConsts c,
                                                                  CldList list,
                                                                  const uint32_t* __restrict__ flags,
                                                                  int useList)
{
  RTP_CQ_PROLOGUE
  const float4 pos = posP[i];
  CqTempCorrect b{ c, lambdaT, lambdaT[i], c.q.restDensity };
  cq_neighbours<kCqReplay>(on, i, useList && !(*flags & kFlagLambdaT), pos, posP, se, c.f.g,
                           list, lds, b);
  if (lead)
    temp[i] = temp[i] + 0.3f * b.corr;
}

__global__ __launch_bounds__(kCldLanes) void k_cld_density_lambda_q(const float4* __restrict__ pred,
                                                                    const uint2* __restrict__ se,
                                                                    float* __restrict__ density,
                                                                    float* __restrict__ lambda,
                                                                    uint32_t n, CloudConsts c,
                                                                    CldList list,
                                                                    uint32_t* __restrict__ flags,
                                                                    uint32_t bit)
{
  RTP_CQ_PROLOGUE
  const float4 pos = pred[i];
  CqDensity b{ c };
  cq_neighbours<kCqRecord>(on, i, true, pos, pred, se, c.f.g, list, lds, b);
  float li = 0.0f;
  if (lead)
    li = cld_density_lambda_out(i, b.fluidDensity, b.sumGradCi, b.sumSqGradC, density, lambda, c);
  if (bit)
    cld_flag(!lead | cld_safe(li), flags, bit);
}

__global__ __launch_bounds__(kCldLanes) void k_cld_correct_q(const float4* __restrict__ pred,
                                                             const float* __restrict__ lambda,
                                                             const uint2* __restrict__ se,
                                                             float4* __restrict__ predOut,
                                                             float4* __restrict__ totCorr,
                                                             float4* __restrict__ vel, uint32_t n,
                                                             CloudConsts c, CldList list,
                                                             const uint32_t* __restrict__ flags,
                                                             uint32_t bit, int useList)
{
  RTP_CQ_PROLOGUE
  const float4 pos = pred[i];
  const float ywdq = (c.f.shortMath & kShortArtDiv) ? rcp_refined(c.f.wdq) : 0.0f;
  CqCorrect b{ c, lambda, lambda[i], ywdq };
  cq_neighbours<kCqReplay>(on, i, useList && bit && !(*flags & bit), pos, pred, se, c.f.g, list,
                           lds, b);
  if (lead)
    cld_correct_out(i, pos, b.corr, predOut, totCorr, vel, c);
}

__global__ __launch_bounds__(kCldLanes) void k_cld_vorticity_q(const float4* __restrict__ pred,
                                                               const uint2* __restrict__ se,
                                                               const float4* __restrict__ vel,
                                                               float4* __restrict__ vort, uint32_t n,
                                                               CloudConsts c, CldList list,
                                                               uint32_t* __restrict__ flags)
{
  RTP_CQ_PROLOGUE
  const float4 pos = pred[i];
  CqVorticity b{ c, vel, vel[i] };
  cq_neighbours<kCqRecord>(on, i, true, pos, pred, se, c.f.g, list, lds, b);
  if (lead)
    vort[i] = b.w;
  cld_flag(!lead | cld_safe(b.w), flags, kFlagVort);
}

__global__ __launch_bounds__(kCldLanes) void k_cld_confinement_q(const float4* __restrict__ pred,
                                                                 const uint2* __restrict__ se,
                                                                 const float4* __restrict__ vort,
                                                                 const float4* __restrict__ vel,
                                                                 float4* __restrict__ velOut,
                                                                 uint32_t n, CloudConsts c,
                                                                 CldList list,
                                                                 uint32_t* __restrict__ flags,
                                                                 int useList)
{
  RTP_CQ_PROLOGUE
  const float4 pos = pred[i];
  CqConfinement b{ c, vort };
  cq_neighbours<kCqReplay>(on, i, useList && !(*flags & kFlagVort), pos, pred, se, c.f.g, list,
                           lds, b);
  float4 vo = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (lead)
    vo = cld_confinement_out(i, b.nn, vort[i], vel, velOut, c);
  cld_flag(!lead | cld_safe(vo), flags, kFlagVelTmp);
}

__global__ __launch_bounds__(kCldLanes) void k_cld_xsph_q(const float4* __restrict__ pred,
                                                          const uint2* __restrict__ se,
                                                          const float4* __restrict__ velIn,
                                                          float4* __restrict__ velOut,
                                                          float4* __restrict__ pos, uint32_t n,
                                                          CloudConsts c, CldList list,
                                                          const uint32_t* __restrict__ flags,
                                                          int useList)
{
  RTP_CQ_PROLOGUE
  const float4 pp = pred[i];
  CqXsph b{ c, velIn, velIn[i] };
  cq_neighbours<kCqReplay>(on, i, useList && !(*flags & kFlagVelTmp), pp, pred, se, c.f.g, list,
                           lds, b);
  if (lead) {
    velOut[i] = add4(b.velocity, muls4(c.f.p.xsphViscosityCoeff, b.viscosity));
    pos[i] = cld_update_pos(pp, c);
  }
}
#undef RTP_CQ_PROLOGUE

__global__ __launch_bounds__(256) void k_cld_exp(const double* __restrict__ x64, double* __restrict__ y64,
                                                 const float* __restrict__ x32, float* __restrict__ y32,
                                                 float* __restrict__ s32, uint32_t n)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  y64[i] = exp64(x64[i]);
  y32[i] = exp32(x32[i]);
  s32[i] = svd(x32[i]);
}

}  // namespace
}  // namespace rtp

using namespace rtp;

// ------------------------------------------------------------------------------- the handle

struct rtp_clouds {
  int device = 0;
  hipStream_t stream = nullptr;
  // the temperature passes run on a second stream beside the density iterations: neither
  // reads what the other writes, until XSPH rewrites the positions the temperature walks
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  GridState grid;
  CloudConsts c{};
  uint32_t n = 0;
  uint32_t iters = 2;  // "Nb Jacobi Iterations" of initCloudsJson (Clouds.cpp:68)
  bool paused = false;
  // state: ping-pong partners of the cell sort's gather (the partners hold the same data
  // outside the sorted range); predicted positions also ping-pong between Jacobi iterations
  float4* pos[2] = {};
  float4* vel[2] = {};
  float4* tot[2] = {};
  float4* pred[2] = {};
  float* temp[2] = {};
  float* buoy[2] = {};
  float* vapor[2] = {};
  float* cloud[2] = {};
  float* pid[2] = {};
  float4* vort = nullptr;
  float4* velTmp = nullptr;
  float* density = nullptr;
  float* lambda = nullptr;
  float* lap = nullptr;
  float* lambdaT = nullptr;
  CldList list;              // survivor lists of the recording passes (kCldListCap per slot)
  CldList listT;             // ... of the temperature passes, which run beside the others
  uint32_t* flags = nullptr;  // cld_flag bits of the step
  bool quad = true;           // walk kernels: a quad of lanes per particle (cq_neighbours)
  int cur = 0;  // state in [cur]
  int pc = 0;   // predicted positions in pred[pc]
  Profiler prof;

  void release()
  {
    grid.release();
    for (int k = 0; k < 2; ++k) {
      dev_free(pos[k]);
      dev_free(vel[k]);
      dev_free(tot[k]);
      dev_free(pred[k]);
      dev_free(temp[k]);
      dev_free(buoy[k]);
      dev_free(vapor[k]);
      dev_free(cloud[k]);
      dev_free(pid[k]);
    }
    dev_free(vort);
    dev_free(velTmp);
    dev_free(density);
    dev_free(lambda);
    dev_free(lap);
    dev_free(lambdaT);
    dev_free(list.entry);
    dev_free(list.count);
    dev_free(listT.entry);
    dev_free(listT.count);
    dev_free(flags);
    prof.destroy();
    if (fork)
      (void)hipEventDestroy(fork);
    if (join)
      (void)hipEventDestroy(join);
    fork = join = nullptr;
    if (side)
      (void)hipStreamDestroy(side);
    side = nullptr;
    if (stream)
      (void)hipStreamDestroy(stream);
    stream = nullptr;
  }

  int step()
  {
    hipStream_t s = stream;
    GridState& G = grid;
    const uint32_t b256 = blocks_for(std::max(n, G.g.numCells), 256);
    RTP_PROF(prof, "cld_thermo_predict", s,
             hipLaunchKernelGGL(k_cld_thermo, dim3(b256), dim3(256), 0, s, pos[cur], vel[cur],
                                temp[cur], vapor[cur], cloud[cur], buoy[cur], pred[pc], tot[cur],
                                G.key[G.keyCur], G.startEnd, n, c, flags));
    int rc = G.sort(n, s, prof);
    if (rc)
      return rc;
    ReorderArgs a{};
    a.key = G.key[G.keyCur];
    a.idx = G.idx[G.idxCur];
    a.sortN = G.sortedN;
    a.n = n;
    a.maxN = G.maxN;
    a.srcA = pos[cur];
    a.dstA = pos[1 - cur];
    a.srcB = vel[cur];
    a.dstB = vel[1 - cur];
    a.srcC = pred[pc];
    a.dstC = pred[1 - pc];
    a.clampC = 0;
    a.startEnd = G.startEnd;
    a.g = c.f.g;
    RTP_PROF(prof, "reorder_cells", s, launch_reorder(a, s));
    if (G.sortedN) {
      CloudGather gth{};
      gth.tot[0] = tot[cur];
      gth.totOut = tot[1 - cur];
      float* const* srcs[5] = { temp, buoy, vapor, cloud, pid };
      for (int k = 0; k < 5; ++k) {
        gth.src[k] = srcs[k][cur];
        gth.dst[k] = srcs[k][1 - cur];
      }
      RTP_PROF(prof, "cld_gather", s,
               hipLaunchKernelGGL(k_cld_gather, dim3(blocks_for(G.sortedN, 256)), dim3(256), 0, s,
                                  G.idx[G.idxCur], G.sortedN, gth));
    }
    cur = 1 - cur;
    pc = 1 - pc;
    const int gathered = pc;
    const uint32_t bn = blocks_for(std::max(n, 1u), 256);
    // the walk kernels: 64 particles per workgroup, or 16 with quads
    const uint32_t bw = blocks_for(std::max(n, 1u), quad ? kCqGroups : kCldLanes);
    const bool qd = quad;
    // Replays need every candidate a walk may visit to have data from this step: not so
    // when the sorted range holds stale particles past n (after the count shrank).
    const int useList = G.sortedN <= n;
    const bool tempOn = n && c.q.isTempSmoothingEnabled;
    if (tempOn) {
      hipStream_t t = side;
      RTP_HIP_TRY(hipEventRecord(fork, s));
      RTP_HIP_TRY(hipStreamWaitEvent(t, fork, 0));
      RTP_PROF(prof, "cld_temp_lambda", t,
               hipLaunchKernelGGL(qd ? k_cld_temp_lambda_q : k_cld_temp_lambda, dim3(bw), dim3(kCldLanes), 0, t, pos[cur], temp[cur],
                                  G.startEnd, lap, lambdaT, n, c, listT, flags));
      RTP_PROF(prof, "cld_temp_correct", t,
               hipLaunchKernelGGL(qd ? k_cld_temp_correct_q : k_cld_temp_correct, dim3(bw), dim3(kCldLanes), 0, t, pos[cur], lambdaT,
                                  G.startEnd, temp[cur], n, c, listT, flags, useList));
      RTP_HIP_TRY(hipEventRecord(join, t));
    }
    for (uint32_t k = 0; k < iters && n; ++k) {
      const uint32_t bit = k < kFlagIters ? kFlagLambda0 << k : 0u;
      RTP_PROF(prof, "cld_density_lambda", s,
               hipLaunchKernelGGL(qd ? k_cld_density_lambda_q : k_cld_density_lambda, dim3(bw), dim3(kCldLanes), 0, s, pred[pc],
                                  G.startEnd, density, lambda, n, c, list, flags, bit));
      RTP_PROF(prof, "cld_correct", s,
               hipLaunchKernelGGL(qd ? k_cld_correct_q : k_cld_correct, dim3(bw), dim3(kCldLanes), 0, s, pred[pc], lambda,
                                  G.startEnd, pred[1 - pc], tot[cur],
                                  k + 1 == iters ? vel[cur] : nullptr, n, c, list, flags, bit,
                                  useList));
      pc = 1 - pc;
    }
    if (iters == 0 && n)
      RTP_PROF(prof, "cld_update_vel", s,
               hipLaunchKernelGGL(k_cld_update_vel, dim3(bn), dim3(256), 0, s, tot[cur], vel[cur], n, c));
    // the sorted range's stale tail [n, sortedN) lives in the gather target
    if (pc != gathered && G.sortedN > n)
      launch_copy_f4(pred[pc], pred[gathered], n, G.sortedN, s);
    if (tempOn)  // the temperature walks read p_pos, which XSPH or the position update rewrites
      RTP_HIP_TRY(hipStreamWaitEvent(s, join, 0));
    if (!n)
      return RTP_OK;
    if (c.f.p.isVorticityConfEnabled) {
      RTP_PROF(prof, "cld_vorticity", s,
               hipLaunchKernelGGL(qd ? k_cld_vorticity_q : k_cld_vorticity, dim3(bw), dim3(kCldLanes), 0, s, pred[pc], G.startEnd,
                                  vel[cur], vort, n, c, list, flags));
      RTP_PROF(prof, "cld_confinement", s,
               hipLaunchKernelGGL(qd ? k_cld_confinement_q : k_cld_confinement, dim3(bw), dim3(kCldLanes), 0, s, pred[pc],
                                  G.startEnd, vort, vel[cur], velTmp, n, c, list, flags, useList));
      RTP_PROF(prof, "cld_xsph", s,
               hipLaunchKernelGGL(qd ? k_cld_xsph_q : k_cld_xsph, dim3(bw), dim3(kCldLanes), 0, s, pred[pc], G.startEnd,
                                  velTmp, vel[cur], pos[cur], n, c, list, flags, useList));
    } else {
      RTP_PROF(prof, "cld_update_pos", s,
               hipLaunchKernelGGL(k_cld_update_pos, dim3(bn), dim3(256), 0, s, pred[pc], pos[cur],
                                  n, c));
    }
    return RTP_OK;
  }
};

namespace {

void refresh(rtp_clouds* h)
{
  refresh_pbf_consts(h->c.f);
}

// fld_* defaults are replaced by initCloudsJson's (Clouds.cpp:65-97) at the reference's
// reset; these are CloudKernelInputs' own defaults (Clouds.hpp:14-48).
rtp_cloud_params default_cloud_params()
{
  rtp_cloud_params q{};
  q.dim = 3;
  q.timeStep = 0.01f;
  q.restDensity = 400.0f;
  q.groundHeatCoeff = 10.0f;
  q.buoyancyCoeff = 0.10f;
  q.gravCoeff = 0.0005f;
  q.adiabaticLapseRate = 5.0f;
  q.phaseTransitionRate = 0.3485f;
  q.latentHeatCoeff = 0.07f;
  q.isTempSmoothingEnabled = 1;
  q.relaxCFM = 600.0f;
  q.initVaporDensityCoeff = 0.75f;
  q.windCoeff = 1.0f;
  return q;
}

}  // namespace

extern "C" {

int rtp_clouds_create(const rtp_model_desc* d, rtp_clouds** out)
{
  RTP_REQUIRE(d && out, "NULL argument");
  *out = nullptr;
  RTP_REQUIRE(d->maxNbParticles > 0, "maxNbParticles must be positive");
  RTP_REQUIRE(d->currNbParticles <= d->maxNbParticles, "currNbParticles exceeds maxNbParticles");
  RTP_REQUIRE(d->maxNbParticles < (1u << kCldIndexBits),
              "maxNbParticles must be below 2^28 for the Clouds model");
  int rc = check_device(d->device);
  if (rc)
    return rc;
  auto* h = new rtp_clouds();
  h->device = d->device;
  if ((rc = make_grid(d, 100u, &h->grid.g))) {
    delete h;
    return rc;
  }
  init_pbf_consts(h->c.f, h->grid.g, d->boxSize[0], d->gridRes[0], d->gravity);
  h->c.q = default_cloud_params();
  refresh(h);
  h->n = d->currNbParticles;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return fail(RTP_ERR_HIP, "hipStreamCreate failed");
  }
  const uint32_t maxN = d->maxNbParticles;
  rc = h->grid.alloc(maxN);
  for (int k = 0; k < 2 && !rc; ++k) {
    if ((rc = dev_alloc(&h->pos[k], maxN)) || (rc = dev_alloc(&h->vel[k], maxN)) ||
        (rc = dev_alloc(&h->tot[k], maxN)) || (rc = dev_alloc(&h->pred[k], maxN)) ||
        (rc = dev_alloc(&h->temp[k], maxN)) || (rc = dev_alloc(&h->buoy[k], maxN)) ||
        (rc = dev_alloc(&h->vapor[k], maxN)) || (rc = dev_alloc(&h->cloud[k], maxN)) ||
        (rc = dev_alloc(&h->pid[k], maxN)))
      break;
  }
  if (!rc)
    (void)((rc = dev_alloc(&h->vort, maxN)) || (rc = dev_alloc(&h->velTmp, maxN)) ||
           (rc = dev_alloc(&h->density, maxN)) || (rc = dev_alloc(&h->lambda, maxN)) ||
           (rc = dev_alloc(&h->lap, maxN)) || (rc = dev_alloc(&h->lambdaT, maxN)) ||
           (rc = dev_alloc(&h->list.entry, (size_t)kCldListCap * maxN)) ||
           (rc = dev_alloc(&h->list.count, maxN)) || (rc = dev_alloc(&h->flags, 1)) ||
           (rc = dev_alloc(&h->listT.entry, (size_t)kCldListCap * maxN)) ||
           (rc = dev_alloc(&h->listT.count, maxN)));
  h->list.stride = maxN;
  h->listT.stride = maxN;
  h->quad = maxN <= kQuadDefaultMaxN;
  if (!rc && (hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking) != hipSuccess ||
              hipEventCreateWithFlags(&h->fork, hipEventDisableTiming) != hipSuccess ||
              hipEventCreateWithFlags(&h->join, hipEventDisableTiming) != hipSuccess))
    rc = fail(RTP_ERR_HIP, "stream or event creation failed");
  if (rc) {
    h->release();
    delete h;
    return rc;
  }
  std::vector<float> zeros(4, 0.0f);
  if ((rc = rtp_clouds_load_state(h, zeros.data(), zeros.data(), 0))) {
    h->release();
    delete h;
    return rc;
  }
  *out = h;
  return RTP_OK;
}

int rtp_clouds_destroy(rtp_clouds* h)
{
  if (!h)
    return RTP_OK;
  (void)hipSetDevice(h->device);
  h->release();
  delete h;
  return RTP_OK;
}

int rtp_clouds_set_params(rtp_clouds* h, const rtp_fluid_params* p, const rtp_cloud_params* q)
{
  RTP_REQUIRE(h && (p || q), "NULL argument");
  if (p) {
    RTP_REQUIRE(p->restDensity != 0.0f, "restDensity must be non-zero");
    RTP_REQUIRE(p->artPressureExp <= kMaxArtPressureExp, "artPressureExp must be at most 64");
  }
  if (q)
    RTP_REQUIRE(q->restDensity != 0.0f, "cloud restDensity must be non-zero");
  if (p)
    h->c.f.p = *p;
  if (q)
    h->c.q = *q;
  refresh(h);
  return RTP_OK;
}

int rtp_clouds_get_params(rtp_clouds* h, rtp_fluid_params* p, rtp_cloud_params* q)
{
  RTP_REQUIRE(h && (p || q), "NULL argument");
  if (p)
    *p = h->c.f.p;
  if (q)
    *q = h->c.q;
  return RTP_OK;
}

int rtp_clouds_set_tuning(rtp_clouds* h, const rtp_tuning* t)
{
  RTP_REQUIRE(h && t, "NULL argument");
  RTP_REQUIRE((t->flags & ~RTP_LAUNCH_QUAD) == 0, "launch flags not supported here");
  h->quad = (t->flags & RTP_LAUNCH_QUAD) != 0;
  return RTP_OK;
}

int rtp_clouds_set_nb_jacobi_iters(rtp_clouds* h, uint32_t iters)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_REQUIRE(iters <= 1000, "iteration count out of range");
  h->iters = iters;
  return RTP_OK;
}

int rtp_clouds_get_nb_jacobi_iters(rtp_clouds* h, uint32_t* iters)
{
  RTP_REQUIRE(h && iters, "NULL argument");
  *iters = h->iters;
  return RTP_OK;
}

int rtp_clouds_set_nb_particles(rtp_clouds* h, uint32_t n)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_REQUIRE(n <= h->grid.maxN, "nb particles exceeds maxNbParticles");
  // keys in [n, old n) stay in the whole-buffer sort (RadixSort.cpp:134-161)
  h->grid.dirtyEnd = std::max(h->grid.dirtyEnd, h->n);
  if (h->grid.dirtyEnd <= n)
    h->grid.dirtyEnd = std::max(h->grid.dirtyEnd, h->grid.sortedN);
  h->n = n;
  return RTP_OK;
}

int rtp_clouds_load_state(rtp_clouds* h, const float* pos4, const float* vel4, uint32_t count)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_REQUIRE(count <= h->grid.maxN, "count exceeds maxNbParticles");
  RTP_REQUIRE(count == 0 || (pos4 && vel4), "NULL state arrays");
  RTP_HIP_TRY(hipSetDevice(h->device));
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  const float inf = INFINITY;
  const uint32_t maxN = h->grid.maxN;
  const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  int rc = 0;
  for (int k = 0; k < 2; ++k) {
    if ((rc = upload_f4(h->pos[k], pos4, count, maxN, make_float4(inf, inf, inf, 0.0f), h->stream)) ||
        (rc = upload_f4(h->vel[k], vel4, count, maxN, zero4, h->stream)) ||
        (rc = upload_f4(h->pred[k], nullptr, 0, maxN, zero4, h->stream)) ||
        (rc = upload_f4(h->tot[k], nullptr, 0, maxN, zero4, h->stream)))
      return rc;
    // Clouds::initCloudsParticles (Clouds.cpp:485-499): cloud density 0, particle ids, then
    // the temperature and vapour density over all maxN slots
    hipLaunchKernelGGL(k_cld_init, dim3(blocks_for(maxN, 256)), dim3(256), 0, h->stream, h->pos[k],
                       h->temp[k], h->vapor[k], h->cloud[k], h->pid[k], maxN, h->c);
  }
  RTP_HIP_TRY(hipMemsetAsync(h->buoy[0], 0, sizeof(float) * maxN, h->stream));
  RTP_HIP_TRY(hipMemsetAsync(h->buoy[1], 0, sizeof(float) * maxN, h->stream));
  RTP_HIP_TRY(hipMemsetAsync(h->density, 0, sizeof(float) * maxN, h->stream));
  RTP_HIP_TRY(hipMemsetAsync(h->lambda, 0, sizeof(float) * maxN, h->stream));
  RTP_HIP_TRY(hipMemsetAsync(h->lap, 0, sizeof(float) * maxN, h->stream));
  RTP_HIP_TRY(hipMemsetAsync(h->lambdaT, 0, sizeof(float) * maxN, h->stream));
  RTP_HIP_TRY(hipMemsetAsync(h->vort, 0, sizeof(float4) * maxN, h->stream));
  h->cur = 0;
  h->pc = 0;
  h->n = count;
  h->grid.reset_keys(h->stream);  // resetCellIDs (Clouds.cpp:396)
  RTP_HIP_TRY(hipGetLastError());
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return RTP_OK;
}

int rtp_clouds_read_state(rtp_clouds* h, float* pos4, float* vel4, uint32_t count)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_REQUIRE(count <= h->grid.maxN, "count exceeds maxNbParticles");
  RTP_HIP_TRY(hipSetDevice(h->device));
  if (pos4)
    RTP_HIP_TRY(hipMemcpyAsync(pos4, h->pos[h->cur], sizeof(float4) * count, hipMemcpyDeviceToHost, h->stream));
  if (vel4)
    RTP_HIP_TRY(hipMemcpyAsync(vel4, h->vel[h->cur], sizeof(float4) * count, hipMemcpyDeviceToHost, h->stream));
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return RTP_OK;
}

int rtp_clouds_update(rtp_clouds* h, uint32_t nsteps)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_HIP_TRY(hipSetDevice(h->device));
  if (h->paused)
    return RTP_OK;
  for (uint32_t k = 0; k < nsteps; ++k) {
    const int rc = h->step();
    if (rc)
      return rc;
  }
  RTP_HIP_TRY(hipGetLastError());
  return RTP_OK;
}

int rtp_clouds_pause(rtp_clouds* h, int pause)
{
  RTP_REQUIRE(h, "NULL handle");
  h->paused = pause != 0;
  return RTP_OK;
}

int rtp_clouds_sync(rtp_clouds* h)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_HIP_TRY(hipSetDevice(h->device));
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return RTP_OK;
}

int rtp_clouds_read_grid(rtp_clouds* h, uint32_t* cellIDs, uint32_t* sortIdx, uint32_t* startEnd)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_HIP_TRY(hipSetDevice(h->device));
  return h->grid.read(cellIDs, sortIdx, startEnd, h->stream);
}

int rtp_clouds_read_field(rtp_clouds* h, int field, float* out, uint32_t count)
{
  RTP_REQUIRE(h && out, "NULL argument");
  RTP_REQUIRE(count <= h->grid.maxN, "count exceeds maxNbParticles");
  RTP_HIP_TRY(hipSetDevice(h->device));
  const void* src = nullptr;
  size_t elem = sizeof(float);
  const int c = h->cur;
  switch (field) {
  case RTP_FIELD_DENSITY: src = h->density; break;
  case RTP_FIELD_LAMBDA: src = h->lambda; break;
  case RTP_FIELD_PRED_POS: src = h->pred[h->pc]; elem = sizeof(float4); break;
  case RTP_FIELD_VORTICITY: src = h->vort; elem = sizeof(float4); break;
  case RTP_FIELD_TEMPERATURE: src = h->temp[c]; break;
  case RTP_FIELD_VAPOR_DENSITY: src = h->vapor[c]; break;
  case RTP_FIELD_CLOUD_DENSITY: src = h->cloud[c]; break;
  case RTP_FIELD_BUOYANCY: src = h->buoy[c]; break;
  case RTP_FIELD_PARTICLE_ID: src = h->pid[c]; break;
  case RTP_FIELD_TOT_CORR_POS: src = h->tot[c]; elem = sizeof(float4); break;
  case RTP_FIELD_LAPLACIAN_TEMP: src = h->lap; break;
  case RTP_FIELD_LAMBDA_TEMP: src = h->lambdaT; break;
  default: return fail(RTP_ERR_INVALID, "unknown field for clouds");
  }
  RTP_HIP_TRY(hipMemcpyAsync(out, src, elem * count, hipMemcpyDeviceToHost, h->stream));
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return RTP_OK;
}

int rtp_clouds_device_state(rtp_clouds* h, void** pos4, void** vel4)
{
  RTP_REQUIRE(h, "NULL handle");
  if (pos4)
    *pos4 = h->pos[h->cur];
  if (vel4)
    *vel4 = h->vel[h->cur];
  return RTP_OK;
}

int rtp_clouds_enable_profiling(rtp_clouds* h, int enable)
{
  RTP_REQUIRE(h, "NULL handle");
  h->prof.enabled = enable != 0;
  return RTP_OK;
}

int rtp_clouds_profile(rtp_clouds* h, rtp_kernel_time* out, int capacity, int* count)
{
  RTP_REQUIRE(h, "NULL handle");
  return h->prof.report(out, capacity, count);
}

int rtp_clouds_update_with_json(rtp_clouds* h, const char* json)
{
  RTP_REQUIRE(h, "NULL handle");
  rtp_fluid_params p = h->c.f.p;
  rtp_cloud_params q = h->c.q;
  uint32_t iters = 0;
  int rc = rtp_cloud_params_from_json(json, &p, &q, &iters);
  if (rc)
    return rc;
  // Clouds.cpp:260, 288: dim follows the model's dimension, not the tree
  p.dim = h->c.f.p.dim;
  q.dim = h->c.q.dim;
  if ((rc = rtp_clouds_set_params(h, &p, &q)))
    return rc;
  return rtp_clouds_set_nb_jacobi_iters(h, iters > 1000u ? 1000u : iters);
}

int rtp_clouds_params_json(rtp_clouds* h, char* buf, size_t cap, size_t* len)
{
  RTP_REQUIRE(h, "NULL handle");
  return rtp_cloud_params_to_json(&h->c.f.p, &h->c.q, h->iters, buf, cap, len);
}

int rtp_selfcheck_exp(int device, const double* x64, const float* x32, uint32_t n, double* y64,
                      float* y32, float* svd32)
{
  RTP_REQUIRE(n == 0 || (x64 && x32 && y64 && y32 && svd32), "NULL argument");
  if (n == 0)
    return RTP_OK;
  int rc = check_device(device);
  if (rc)
    return rc;
  double *dx = nullptr, *dy = nullptr;
  float *fx = nullptr, *fy = nullptr, *fs = nullptr;
  if ((rc = dev_alloc(&dx, n)) || (rc = dev_alloc(&dy, n)) || (rc = dev_alloc(&fx, n)) ||
      (rc = dev_alloc(&fy, n)) || (rc = dev_alloc(&fs, n))) {
    dev_free(dx);
    dev_free(dy);
    dev_free(fx);
    dev_free(fy);
    dev_free(fs);
    return rc;
  }
  if (hipMemcpy(dx, x64, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(fx, x32, sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(RTP_ERR_HIP, "hipMemcpy inputs");
  if (!rc) {
    hipLaunchKernelGGL(k_cld_exp, dim3(blocks_for(n, 256)), dim3(256), 0, nullptr, dx, dy, fx, fy,
                       fs, n);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
      rc = fail(RTP_ERR_HIP, "exp self-check launch failed");
  }
  if (!rc && (hipMemcpy(y64, dy, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(y32, fy, sizeof(float) * n, hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(svd32, fs, sizeof(float) * n, hipMemcpyDeviceToHost) != hipSuccess))
    rc = fail(RTP_ERR_HIP, "hipMemcpy results");
  dev_free(dx);
  dev_free(dy);
  dev_free(fx);
  dev_free(fy);
  dev_free(fs);
  return rc;
}

}  // extern "C"
