// Position based fluids solver kernels for gfx950.
//
// Reference: physics/ocl/kernels/fluids.cl and sph.cl, step order physics/ocl/Fluids.cpp:400-471.
// The reference runs 5 kernels per Jacobi iteration and 6 after them; here the work of one
// substep is fused into 2 neighbour passes per iteration plus 3 after them:
//
//   predict  : fld_predictPosition + fillCellIDs + resetStartEndCell          (streaming)
//   [sort + reorder: grid_sort.hip; the reorder applies the first boundary clamp]
//   A_k      : fld_computeDensity + fld_computeConstraintFactor               (neighbours)
//   B_k      : fld_computeConstraintCorrection + fld_correctPosition, then the next
//              iteration's fld_applyBoundaryCondition, or for the last iteration
//              fld_updateVel (+ fld_updatePosition when vorticity is off)    (neighbours)
//   C        : fld_computeVorticity (w lane carries |vorticity|)              (neighbours)
//   D        : fld_applyVorticityConfinement -> velocity copy                  (neighbours)
//   E        : fld_applyXsphViscosityCorrection + fld_updatePosition           (neighbours)
//
// Fusion keeps each particle's summation order: every pass walks the 27 cells in the
// reference's order (x outer, z inner, periodic wrap) and each cell from start to end.
// Candidates whose squared distance proves r >= h contribute +-0 in the reference and
// are skipped; candidates that are very far (d2 >= kFarD2, where the reference's
// zero-weighted poly6 overflows to NaN) or NaN take the reference's exact expressions.
#include "rtp_crmath.hpp"
#include "rtp_device.hpp"

namespace rtp {

namespace {

// Candidates that survive the distance test are queued per lane in LDS and handed to the
// pass body afterwards, so the expensive body runs densely over each lane's survivors
// instead of once per candidate slot of the whole wave.  Survivors keep their order.  The
// queue is the kernel's dynamic LDS, depth x block size words; its depth is a launch
// parameter (deeper queues densify more, shallower ones leave room for more waves).
constexpr uint32_t kMaxBlock = 256; // threads per block

// fld_applyBoundaryCondition (fluids.cl:435-439) applied to a loaded position.
__device__ __forceinline__ float4 clamp_pos(float4 q, const PbfConsts& c)
{
  return make_float4(clampf(q.x, c.lo[0], c.hi[0]), clampf(q.y, c.lo[1], c.hi[1]),
                     clampf(q.z, c.lo[2], c.hi[2]), 0.0f);
}

// Position e; with CL the boundary clamp is applied on the way in, which gives a pass the
// positions the reference clamps in place at the start of a Jacobi iteration.
template <bool CL>
__device__ __forceinline__ float4 ldp(const float4* __restrict__ pred, uint32_t e, const PbfConsts& c)
{
  const float4 q = pred[e];
  return CL ? clamp_pos(q, c) : q;
}

// The same for a neighbour, whose w lane (always 0) no distance reads: 12 bytes instead of 16.
template <bool CL>
__device__ __forceinline__ float4 ldn(const float4* __restrict__ pred, uint32_t e, const PbfConsts& c)
{
  const float* f = reinterpret_cast<const float*>(pred + e);
  const float4 q = make_float4(f[0], f[1], f[2], 0.0f);
  return CL ? clamp_pos(q, c) : q;
}

// A lane's survivor queue: entries slot[0], slot[stride], ... up to cap entries, the next
// free one at wp.  Candidates are written unconditionally and kept by advancing wp, and the
// queue is drained whenever fewer than four entries are free (wp > room4), so a group of
// four candidates always fits.
struct LaneQueue {
  uint32_t* slot;   // this lane's first entry
  uint32_t* wp;     // next free entry
  uint32_t* room4;  // last position of wp that leaves four free entries
  uint32_t stride;  // threads per block
};

// A pass's per-neighbour work: ld(e) loads what the body needs of neighbour e besides its
// position (issued one survivor ahead), f(e, x, dx, dy, dz, d2) is the body.
template <class L, class F>
struct Body {
  L& ld;
  F& f;
};

// One survivor's body: the reference's d2 (unfused) and the pass's work.
template <class B, class X>
__device__ __forceinline__ void body_at(const float4 p, uint32_t e, const float4 o, const X& x,
                                        B& b)
{
  const float dx = p.x - o.x, dy = p.y - o.y, dz = p.z - o.z;
  const float d2 = dx * dx + dy * dy + dz * dz;
  b.f(e, x, dx, dy, dz, d2);
}

// Process the lane's queued survivors in order, loading survivor k+1 while computing k.
// Unrolled by two over two register sets, so no survivor's data is copied from one set to
// the other between bodies.
template <bool CL, class B>
__device__ __forceinline__ void drain(LaneQueue& q, const float4 p, const float4* __restrict__ pred,
                                      const PbfConsts& c, B& b)
{
  const uint32_t* rp = q.slot;
  if (rp != q.wp) {
    uint32_t e0 = *rp;
    float4 o0 = ldn<CL>(pred, e0, c);
    auto x0 = b.ld(e0);
    for (;;) {
      rp += q.stride;
      bool more = rp != q.wp;
      const uint32_t e1 = *(more ? rp : q.slot);  // in bounds either way
      const float4 o1 = ldn<CL>(pred, e1, c);
      const auto x1 = b.ld(e1);
      body_at(p, e0, o0, x0, b);
      if (!more)
        break;
      rp += q.stride;
      more = rp != q.wp;
      e0 = *(more ? rp : q.slot);
      o0 = ldn<CL>(pred, e0, c);
      x0 = b.ld(e0);
      body_at(p, e1, o1, x1, b);
      if (!more)
        break;
    }
  }
  q.wp = q.slot;
}

// Queue candidate e unless its contribution is provably +-0 (r >= h from the squared
// distance); very far (d2 >= kFarD2) and NaN distances are queued for the exact
// expressions.  The test uses the fused squared distance against bounds widened by 2^-20
// relative, so it never skips what the unfused one keeps; the few extra candidates it
// keeps contribute exact zeros, which leave every sum (started at +0) unchanged.
template <bool CL>
__device__ __forceinline__ void offer(LaneQueue& q, const float4 p, const float4 o, uint32_t e,
                                      const PbfConsts& c)
{
  const float dx = p.x - o.x, dy = p.y - o.y, dz = p.z - o.z;
  const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
  const bool keep = !(d2 >= c.cullFused && d2 < c.farFused);
  *q.wp = e;
  q.wp += keep ? q.stride : 0u;
}

// When any lane of the wave is short of room, every active lane drains: a drain costs the
// wave one body per entry of its fullest lane, so draining the others with it is free
// progress.  Drain timing does not change any lane's order of survivors.
template <bool CL, class B>
__device__ __forceinline__ void make_room(LaneQueue& q, const float4 p,
                                          const float4* __restrict__ pred, const PbfConsts& c, B& b)
{
  if (__builtin_amdgcn_ballot_w64(q.wp > q.room4) != 0)
    drain<CL>(q, p, pred, c, b);
}

// Candidates lo..hi (inclusive) in index order, four loads in flight at a time.
template <bool CL, class B>
__device__ __forceinline__ void scan_range(uint32_t lo, uint32_t hi, LaneQueue& q, const float4 p,
                                           const float4* __restrict__ pred, const PbfConsts& c,
                                           B& b)
{
  uint32_t e = lo;
  for (; e + 3u <= hi; e += 4u) {
    const float4* g = pred + e;
    const float4 o0 = ldn<CL>(g, 0u, c), o1 = ldn<CL>(g, 1u, c), o2 = ldn<CL>(g, 2u, c),
                 o3 = ldn<CL>(g, 3u, c);
    offer<CL>(q, p, o0, e, c);
    offer<CL>(q, p, o1, e + 1u, c);
    offer<CL>(q, p, o2, e + 2u, c);
    offer<CL>(q, p, o3, e + 3u, c);
    make_room<CL>(q, p, pred, c, b);
  }
  for (; e <= hi; ++e)
    offer<CL>(q, p, ldn<CL>(pred, e, c), e, c);
  make_room<CL>(q, p, pred, c, b);
}

// ------------------------------------------------------------------ wave tiles in LDS
// A wave's 64 consecutive particles (sorted by cell) mostly share their neighbour cells.
// Before a neighbour pass each wave stages into LDS the candidate positions of up to
// kTileCols cell columns (x, y) around its lanes' cells, each over the z-range its lanes
// can reach, and the walk reads a range from LDS whenever that range lies inside its
// column's staged span -- otherwise, as without a tile, from global memory.  The
// candidates, their order and every result are the same either way.
constexpr uint32_t kTileCols = 16;
#ifndef RTP_TILE_CAP
#define RTP_TILE_CAP 768
#endif
constexpr uint32_t kTileCap = RTP_TILE_CAP;      // staged candidates per wave
constexpr uint32_t kWaves = kMaxBlock / 64;

// What a neighbour walk works with besides the particle arrays: its lane's survivor queue
// in LDS and, optionally, the wave's tile.  The kernel owns (and sizes) both.
struct WalkCtx {
  uint32_t* qslot = nullptr;               // this lane's queue column
  uint32_t qcap = 0;
  bool on = false;                         // a tile is staged
  int32_t x0 = 0, y0 = 0, nx = 0, ny = 0;  // staged columns: [x0, x0+nx) x [y0, y0+ny)
  const uint32_t* lo = nullptr;            // per column: first staged global index
  const uint32_t* hi = nullptr;            //             last (hi < lo: not staged)
  const uint32_t* off = nullptr;           //             offset in x, y, z
  const float* x = nullptr;
  const float* y = nullptr;
  const float* z = nullptr;
  NbList nb;                               // survivor lists of a kNbRecord / kNbReplay pass
  uint32_t slot = 0;                       // this thread's column in them
};

// LDS a block's waves stage their tiles in (positions as three float arrays).
struct TileLds {
  float x[kWaves][kTileCap], y[kWaves][kTileCap], z[kWaves][kTileCap];
  uint32_t meta[kWaves][3][kTileCols];
};

__device__ __forceinline__ int32_t wave_min(int32_t v)
{
  for (int o = 32; o > 0; o >>= 1)
    v = min(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ int32_t wave_max(int32_t v)
{
  for (int o = 32; o > 0; o >>= 1)
    v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// Stage the wave's tile.  Every lane of the wave must call this (with its own position p,
// clamped on load when CL, as the pass will see it); positions are staged the same way.
template <bool CL>
__device__ __forceinline__ void build_tile(const float4 p, const float4* __restrict__ pred,
                                           const uint2* __restrict__ se, const PbfConsts& c,
                                           TileLds& lds, WalkCtx& T)
{
  const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  uint32_t ux, uy, uz;
  cell3d(p, c.g, ux, uy, uz);
  const int32_t cx = (int32_t)ux, cy = (int32_t)uy, cz = (int32_t)uz;
  const int32_t xmin = wave_min(cx), xmax = wave_max(cx), ymin = wave_min(cy), ymax = wave_max(cy);
  if (xmax - xmin > 1 || ymax - ymin > 1)
    return;
  const int32_t rx = c.g.res[0], ry = c.g.res[1], rz = c.g.res[2];
  T.x0 = xmin - 1;
  T.y0 = ymin - 1;
  T.nx = xmax - xmin + 3;
  T.ny = ymax - ymin + 3;
  const int32_t ncol = T.nx * T.ny;  // <= kTileCols
  // z-range of every column: the cells cz-1..cz+1 of the lanes whose 3x3 columns include it
  int32_t myLo = 1, myHi = 0;
  for (int32_t k = 0; k < ncol; ++k) {
    const int32_t colx = T.x0 + k / T.ny, coly = T.y0 + k % T.ny;
    const bool in = abs(cx - colx) <= 1 && abs(cy - coly) <= 1;
    const int32_t zl = wave_min(in ? cz - 1 : 0x7fffffff);
    const int32_t zh = wave_max(in ? cz + 1 : -0x7fffffff);
    if ((int32_t)lane == k) {
      myLo = max(zl, 0);
      myHi = min(zh, rz - 1);
      if (colx < 0 || colx >= rx || coly < 0 || coly >= ry)
        myHi = myLo - 1;  // outside the grid: the walk reaches such columns only by wrapping
    }
  }
  // lane k < ncol: index span of its column's cells myLo..myHi
  uint32_t first = 1u, last = 0u;
  if ((int32_t)lane < ncol && myLo <= myHi) {
    const int32_t colx = T.x0 + (int32_t)lane / T.ny, coly = T.y0 + (int32_t)lane % T.ny;
    const uint2* col = se + (colx * ry + coly) * rz;
    first = 0xffffffffu;
    for (int32_t z = myLo; z <= myHi; ++z) {
      const uint2 r = col[z];
      if (r.x <= r.y) {
        first = min(first, r.x);
        last = max(last, r.y);
      }
    }
    if (first == 0xffffffffu) {
      first = 1u;
      last = 0u;
    }
  }
  uint32_t size = first <= last ? last - first + 1u : 0u;
  // exclusive prefix of the sizes over the lanes
  uint32_t incl = size;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(incl, o, 64);
    if ((int)lane >= o)
      incl += t;
  }
  const uint32_t off = incl - size;
  if (off + size > kTileCap) {  // does not fit: this column reads global memory
    size = 0u;
    first = 1u;
    last = 0u;
  }
  if ((int32_t)lane < ncol) {
    lds.meta[w][0][lane] = first;
    lds.meta[w][1][lane] = last;
    lds.meta[w][2][lane] = off;
  }
  for (int32_t k = 0; k < ncol; ++k) {
    const uint32_t g0 = __shfl(first, k, 64), sz = __shfl(size, k, 64), o = __shfl(off, k, 64);
    for (uint32_t t = lane; t < sz; t += 64u) {
      const float4 q = ldp<CL>(pred, g0 + t, c);
      lds.x[w][o + t] = q.x;
      lds.y[w][o + t] = q.y;
      lds.z[w][o + t] = q.z;
    }
  }
  T.on = true;
  T.lo = lds.meta[w][0];
  T.hi = lds.meta[w][1];
  T.off = lds.meta[w][2];
  T.x = lds.x[w];
  T.y = lds.y[w];
  T.z = lds.z[w];
}

// Candidates lo..hi read from the tile; the staged entry of candidate e is e - shift.
template <bool CL, class B>
__device__ __forceinline__ void scan_tile(uint32_t lo, uint32_t hi, uint32_t shift,
                                          const WalkCtx& T, LaneQueue& q, const float4 p,
                                          const float4* __restrict__ pred, const PbfConsts& c,
                                          B& b)
{
  uint32_t e = lo;
  for (; e + 3u <= hi; e += 4u) {
    const uint32_t t = e - shift;
    const float4 o0 = make_float4(T.x[t], T.y[t], T.z[t], 0.0f);
    const float4 o1 = make_float4(T.x[t + 1u], T.y[t + 1u], T.z[t + 1u], 0.0f);
    const float4 o2 = make_float4(T.x[t + 2u], T.y[t + 2u], T.z[t + 2u], 0.0f);
    const float4 o3 = make_float4(T.x[t + 3u], T.y[t + 3u], T.z[t + 3u], 0.0f);
    offer<CL>(q, p, o0, e, c);
    offer<CL>(q, p, o1, e + 1u, c);
    offer<CL>(q, p, o2, e + 2u, c);
    offer<CL>(q, p, o3, e + 3u, c);
    make_room<CL>(q, p, pred, c, b);
  }
  for (; e <= hi; ++e) {
    const uint32_t t = e - shift;
    offer<CL>(q, p, make_float4(T.x[t], T.y[t], T.z[t], 0.0f), e, c);
  }
  make_room<CL>(q, p, pred, c, b);
}

// One index range of column (nx, ny): from the tile when staged, else from global memory.
template <bool CL, class B>
__device__ __forceinline__ void scan_any(uint32_t lo, uint32_t hi, const WalkCtx& T, int32_t k,
                                         LaneQueue& q, const float4 p,
                                         const float4* __restrict__ pred, const PbfConsts& c,
                                         B& b)
{
  if (lo > hi)
    return;
  if (k >= 0 && lo >= T.lo[k] && hi <= T.hi[k]) {
    scan_tile<CL>(lo, hi, T.lo[k] - T.off[k], T, q, p, pred, c, b);
    return;
  }
  scan_range<CL>(lo, hi, q, p, pred, c, b);
}

// Visit the candidates of particle position p in the reference's order: 27 cells, x
// outer and z inner with periodic wrap (fluids.cl:101-123), each cell from start to end.
// The three cells of one z-column are scanned as a single index range whenever the cell
// table shows their non-empty ranges are back to back (always, unless a cell was capped or
// the column wraps) -- the same candidates in the same order with one loop instead of three.
// The queue is drained after every x-slab of nine cells.
template <bool CL, class L, class F>
__device__ __forceinline__ void walk_cells(const float4 p, const float4* __restrict__ pred,
                                           const uint2* __restrict__ se, const PbfConsts& c,
                                           L& ld, F& f, const WalkCtx& T)
{
  Body<L, F> b{ ld, f };
  LaneQueue q{ T.qslot, T.qslot, T.qslot + (T.qcap - 4u) * blockDim.x, blockDim.x };
  uint32_t cx, cy, cz;
  cell3d(p, c.g, cx, cy, cz);
  const int32_t rx = c.g.res[0], ry = c.g.res[1], rz = c.g.res[2];
  const int32_t z0 = wrap((int32_t)cz - 1, rz), z1 = wrap((int32_t)cz, rz),
                z2 = wrap((int32_t)cz + 1, rz);
  for (int ix = -1; ix <= 1; ++ix) {
    const int32_t nx = wrap((int32_t)cx + ix, rx);
    for (int iy = -1; iy <= 1; ++iy) {
      const int32_t row = (nx * ry + wrap((int32_t)cy + iy, ry)) * rz;
      const uint2 ca = se[row + z0], cb = se[row + z1], cd = se[row + z2];
      // Merge when every non-empty range starts right after the previous non-empty one.
      uint32_t lo = 1u, hi = 0u;
      bool ok = true;
      const uint2 cells[3] = { ca, cb, cd };
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
          ok = false;
        }
      }
      int32_t k = -1;  // staged column of (nx, ny), if any
      if (T.on) {
        const int32_t dxc = nx - T.x0, dyc = wrap((int32_t)cy + iy, ry) - T.y0;
        if (dxc >= 0 && dxc < T.nx && dyc >= 0 && dyc < T.ny)
          k = dxc * T.ny + dyc;
      }
      if (ok) {
        scan_any<CL>(lo, hi, T, k, q, p, pred, c, b);
      } else {
        scan_any<CL>(ca.x, ca.y, T, k, q, p, pred, c, b);
        scan_any<CL>(cb.x, cb.y, T, k, q, p, pred, c, b);
        scan_any<CL>(cd.x, cd.y, T, k, q, p, pred, c, b);
      }
    }
    drain<CL>(q, p, pred, c, b);
  }
}

// Run a recorded survivor list: the same body calls, in the same order, as the walk that
// recorded it.  Entry k + 1's granule, position and pass data are loaded while entry k is
// computed (a new granule every fourth entry).  Unrolled by two: odd entries are lanes y or
// w of the current granule, even ones lanes x (of the next granule) or z, and the two
// register sets alternate, so nothing is copied between bodies.
template <bool CL, class B>
__device__ __forceinline__ void replay(const uint4* __restrict__ gp, uint32_t stride, uint32_t cnt,
                                       const float4 p, const float4* __restrict__ pred,
                                       const PbfConsts& c, B& b)
{
  if (cnt == 0)
    return;
  uint4 v = *gp;
  uint32_t e0 = v.x;
  float4 o0 = ldn<CL>(pred, e0, c);
  auto x0 = b.ld(e0);
  for (uint32_t k = 1;; k += 2u) {
    bool more = k < cnt;
    const uint32_t e1 = !more ? e0 : ((k & 2u) ? v.w : v.y);
    const float4 o1 = ldn<CL>(pred, e1, c);
    const auto x1 = b.ld(e1);
    body_at(p, e0, o0, x0, b);
    if (!more)
      break;
    const uint32_t k2 = k + 1u;
    more = k2 < cnt;
    if (more && (k2 & 3u) == 0u) {
      gp += stride;
      v = *gp;
    }
    e0 = !more ? e1 : ((k2 & 2u) ? v.z : v.x);
    o0 = ldn<CL>(pred, e0, c);
    x0 = b.ld(e0);
    body_at(p, e1, o1, x1, b);
    if (!more)
      break;
  }
}

// The neighbour walk of one particle: walk_cells, or (NB == kNbRecord) walk_cells recording
// every survivor the body runs on into the thread's list, or (NB == kNbReplay) the thread's
// recorded list when it holds all of them.
template <bool CL = false, int NB = 0, class L, class F>
__device__ __forceinline__ void walk27(const float4 p, const float4* __restrict__ pred,
                                       const uint2* __restrict__ se, const PbfConsts& c, L&& ld,
                                       F&& f, const WalkCtx& T)
{
  if constexpr (NB == kNbReplay) {
    const uint32_t cnt = T.nb.count[T.slot];
    if (cnt <= kNbCap) {
      Body<L, F> b{ ld, f };
      replay<CL>(T.nb.list + T.slot, T.nb.stride, cnt, p, pred, c, b);
      return;
    }
    walk_cells<CL>(p, pred, se, c, ld, f, T);
  } else if constexpr (NB == kNbRecord) {
    uint4* gp = T.nb.list + T.slot;
    uint4* const end = gp + (size_t)(kNbCap / 4u) * T.nb.stride;
    const uint32_t stride = T.nb.stride;
    uint32_t cnt = 0, r0 = 0, r1 = 0, r2 = 0, r3 = 0;  // the last four survivors, oldest first
    auto fr = [&](uint32_t e, auto x, float dx, float dy, float dz, float d2) {
      r0 = r1;
      r1 = r2;
      r2 = r3;
      r3 = e;
      ++cnt;
      if ((cnt & 3u) == 0u && gp != end) {
        *gp = make_uint4(r0, r1, r2, r3);
        gp += stride;
      }
      f(e, x, dx, dy, dz, d2);
    };
    walk_cells<CL>(p, pred, se, c, ld, fr, T);
    const uint32_t m = cnt & 3u;  // survivors of the last, partial granule
    if (m != 0u && gp != end)
      *gp = m == 1u ? make_uint4(r3, 0u, 0u, 0u)
                    : (m == 2u ? make_uint4(r2, r3, 0u, 0u) : make_uint4(r1, r2, r3, 0u));
    T.nb.count[T.slot] = cnt;
  } else {
    walk_cells<CL>(p, pred, se, c, ld, f, T);
  }
}

// sph.cl:26-34 for r > FLOAT_EPS with (1 - step(h, r)) == t.
struct Grad {
  float x, y, z;
};
__device__ __forceinline__ Grad grad_spiky(float dx, float dy, float dz, float r, float t,
                                           const PbfConsts& c)
{
  const float p2 = (c.h - r) * (c.h - r);
  Grad g;
  g.x = (((dx * t) * c.spiky) * -3.0f) * p2 / r;
  g.y = (((dy * t) * c.spiky) * -3.0f) * p2 / r;
  g.z = (((dz * t) * c.spiky) * -3.0f) * p2 / r;
  return g;
}

// sph.cl:10-14 with (1 - step(h, r)) == t.
__device__ __forceinline__ float poly6(float r, float t, const PbfConsts& c)
{
  return (t * c.poly6) * cube(c.hh - r * r);
}

// fluids.cl:51-57 given poly6(vec).
__device__ __forceinline__ float art_pressure(float w, const PbfConsts& c)
{
  if (c.p.isArtPressureEnabled == 0)
    return 0.0f;
  return -c.p.artPressureCoeff * pown(w / c.wdq, c.p.artPressureExp);
}

// The same three functions for a neighbour with r < h (t == 1, so the factors t are exact
// identities and left out) with the short correctly rounded forms; shortMath's bits say
// whether the operand ranges of those forms hold for this parameter set.
__device__ __forceinline__ float poly6_in(float r, const PbfConsts& c)
{
  return c.poly6 * cube(c.hh - r * r);
}

// gradSpiky for kFloatEps < r < h, as grad_spiky(dx, dy, dz, r, 1, c); yr = rcp_refined(r).
// Lanes with a nonzero numerator under 2^-90 take the general division.
__device__ __forceinline__ Grad grad_spiky_in(float dx, float dy, float dz, float r, float yr,
                                              const PbfConsts& c)
{
  const float p2 = (c.h - r) * (c.h - r);
  const float nx = ((dx * c.spiky) * -3.0f) * p2;
  const float ny = ((dy * c.spiky) * -3.0f) * p2;
  const float nz = ((dz * c.spiky) * -3.0f) * p2;
  Grad g;
  g.x = div_short(nx, r, yr);
  g.y = div_short(ny, r, yr);
  g.z = div_short(nz, r, yr);
  const bool general = tiny_numerator3(nx, ny, nz);
  if (__builtin_amdgcn_ballot_w64(general) != 0) {
    rare_path();
    if (general) {
      g.x = nx / r;
      g.y = ny / r;
      g.z = nz / r;
    }
  }
  return g;
}

// pown for the exponent of the dam-break parameters without the loop.
__device__ __forceinline__ float pown_art(float x, uint32_t n)
{
  if (n == 4u)
    return ((x * x) * x) * x;
  return pown(x, n);
}

// art_pressure(w, c) for w = poly6_in(r): ywdq = rcp_refined(c.wdq) when kShortArtDiv.
__device__ __forceinline__ float art_pressure_in(float w, float ywdq, const PbfConsts& c)
{
  if (c.p.isArtPressureEnabled == 0)
    return 0.0f;
  const float q = (c.shortMath & kShortArtDiv) ? div_short(w, c.wdq, ywdq) : w / c.wdq;
  return -c.p.artPressureCoeff * pown_art(q, c.p.artPressureExp);
}

__device__ __forceinline__ uint32_t particle_index(int remap)
{
  return logical_block(remap) * blockDim.x + threadIdx.x;
}

// ---------------------------------------------------------------------------- kernels

// The survivor queue of an untiled neighbour pass.
#define RTP_PLAIN_CTX(DEPTH)                                 \
  extern __shared__ uint32_t rtpQueue[];                     \
  WalkCtx W;                                                 \
  W.qslot = &rtpQueue[threadIdx.x];                          \
  W.qcap = (DEPTH);
// ... and the survivor lists of thread t (a plain pass's t and po).
#define RTP_NB_CTX \
  W.nb = po.nb;    \
  W.slot = t;

// fluids.cl:71-73 (newVel = vel + GRAVITY_ACC * dt; predPos = pos + newVel * dt) and
// fillCellIDs (grid.cl:76-86) on predPos, the key offset by the scene's key base.
__device__ __forceinline__ void predict_body(uint32_t t, const float4* __restrict__ pos,
                                             const float4* __restrict__ vel,
                                             float4* __restrict__ pred, uint32_t* __restrict__ key,
                                             uint32_t kbase, const PbfConsts& c)
{
  const float4 p = pos[t], v = vel[t];
  const float dt = c.p.timeStep;
  const float zdt = 0.0f * dt;
  const float vx = v.x + zdt, vy = v.y + c.gdt, vz = v.z + zdt;
  float4 q;
  q.x = p.x + vx * dt;
  q.y = p.y + vy * dt;
  q.z = p.z + vz * dt;
  q.w = 0.0f;
  pred[t] = q;
  key[t] = kbase + cell1d(q, c.g);
}

__global__ __launch_bounds__(256) void k_predict(const float4* __restrict__ pos,
                                                 const float4* __restrict__ vel,
                                                 float4* __restrict__ pred,
                                                 uint32_t* __restrict__ key,
                                                 uint2* __restrict__ startEnd, uint32_t n,
                                                 PbfConsts c)
{
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n)
    predict_body(t, pos, vel, pred, key, 0u, c);
  if (t < c.g.numCells)
    startEnd[t] = make_uint2(1u, 0u);  // grid.cl:91-96
}

// fld_computeDensity (fluids.cl:80-126) + fld_computeConstraintFactor (:131-193).
// MODE bits (the plain solver uses 0):
//   kFeatures : also write the 8 warm-start features of the particle
//   kAccum    : add lambda into la.lamAcc (teacher runs), or start it when la.accumInit
//   kDensOnly : density only, into `density` (the post-correction residual pass)
//   kClampIn  : clamp positions on load (first fallback iteration)
template <int MODE>
__device__ __forceinline__ void density_lambda_body(uint32_t i, const float4* __restrict__ pred,
                                                    const uint2* __restrict__ se,
                                                    float* __restrict__ density,
                                                    float* __restrict__ lambda,
                                                    const PbfConsts& c, const LearnArgs& la,
                                                    const WalkCtx& T)
{
  constexpr bool CL = (MODE & kClampIn) != 0;
  const float4 p = ldp<CL>(pred, i, c);
  float dens = 0.0f, sgx = 0.0f, sgy = 0.0f, sgz = 0.0f, ssq = 0.0f, count = 0.0f;
  const bool fast = (c.shortMath & kShortSqrtGrad) != 0;
  walk27<CL, (MODE & (kNbRecord | kNbReplay))>(p, pred, se, c, [](uint32_t) { return 0; },
             [&](uint32_t, int, float dx, float dy, float dz, float d2) {
    const float r = fast ? sqrt_cr(d2) : __builtin_sqrtf(d2);
    if (fast && r < c.h) {
      // gradient zero (and nothing added) for r <= kFloatEps; adding +0 changes no sum
      dens = dens + poly6_in(r, c);
      if (MODE & kFeatures)
        count += 1.0f;
      if (!(MODE & kDensOnly)) {
        Grad g = grad_spiky_in(dx, dy, dz, r, rcp_refined(r), c);
        const bool on = r > kFloatEps;
        g.x = on ? g.x : 0.0f;
        g.y = on ? g.y : 0.0f;
        g.z = on ? g.z : 0.0f;
        sgx = sgx + g.x;
        sgy = sgy + g.y;
        sgz = sgz + g.z;
        ssq = ssq + (g.x * g.x + g.y * g.y + g.z * g.z);
      }
      return;
    }
    const float t = r < c.h ? 1.0f : 0.0f;
    dens = dens + poly6(r, t, c);
    if (MODE & kFeatures)
      count += t;
    if (!(MODE & kDensOnly) && r > kFloatEps) {
      const Grad g = grad_spiky(dx, dy, dz, r, t, c);
      sgx = sgx + g.x;
      sgy = sgy + g.y;
      sgz = sgz + g.z;
      ssq = ssq + (g.x * g.x + g.y * g.y + g.z * g.z);
    }
  }, T);
  density[i] = dens;
  if (MODE & kDensOnly)
    return;
  const float C = dens / c.p.restDensity - 1.0f;
  float sq = ssq + (sgx * sgx + sgy * sgy + sgz * sgz);
  sq = sq / c.rho0sq;
  const float lam = -C / (sq + c.p.relaxCFM);
  lambda[i] = lam;
  if (MODE & kAccum)
    la.lamAcc[i] = la.accumInit ? lam : la.lamAcc[i] + lam;
  if (MODE & kFeatures) {
    // constraint, plain lambda, neighbour count, sum |grad C|^2, the particle's own
    // |grad C|^2 (near zero inside the fluid, large at its surface), speed and the
    // distances to the two nearest walls of the clamp box
    const float4 v = la.vel[i];
    float w[6] = { p.x - c.lo[0], c.hi[0] - p.x, p.y - c.lo[1], c.hi[1] - p.y, p.z - c.lo[2], c.hi[2] - p.z };
    float m0 = 3.0e38f, m1 = 3.0e38f;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const float d = fmaxf(w[k], 0.0f);
      if (d < m0) {
        m1 = m0;
        m0 = d;
      } else if (d < m1) {
        m1 = d;
      }
    }
    float* f = la.feat + (size_t)i * 8u;
    f[0] = C;
    f[1] = lam;
    f[2] = count;
    f[3] = sq;
    f[4] = (sgx * sgx + sgy * sgy + sgz * sgz) / c.rho0sq;
    f[5] = __builtin_sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    f[6] = m0;
    f[7] = m1;
  }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_density_lambda(const float4* __restrict__ pred,
                                                        const uint2* __restrict__ se,
                                                        float* __restrict__ density,
                                                        float* __restrict__ lambda, uint32_t n,
                                                        PbfConsts c, int remap, LearnArgs la, uint32_t qdepth,
                     PassOrder po)
{
  const uint32_t t = particle_index(remap);
  RTP_PLAIN_CTX(qdepth)
  RTP_NB_CTX
  if (t < n)
    density_lambda_body<MODE>(po.particle(t), pred, se, density, lambda, c, la, W);
}

// fld_computeConstraintCorrection (fluids.cl:198-247) + fld_correctPosition (:252-258),
// then one of (`last`):
//   kNextIter : fld_applyBoundaryCondition of the next iteration
//   kLastVel  : fld_updateVel (:263-273) and, without vorticity, fld_updatePosition (:444-450)
//   kLastOnly : nothing more (the learned step updates the velocity after its fallback)
// MODE bit kClampIn as for k_density_lambda.
// The cell the next pass's walk starts from, the key of the next thread order.
__device__ __forceinline__ void record_cell(uint32_t i, const float4 q, const PbfConsts& c,
                                            const PassOrder& po)
{
  if (!po.keyOut)
    return;
  po.keyOut[i] = po.keyBase + (po.bricks ? brick_key(q, c.g) : cell1d(q, c.g));
}

// The corrected position of particle i (position p, correction sum (cx, cy, cz)) and what
// `last` adds to it.
__device__ __forceinline__ void correct_out(uint32_t i, const float4 p, float cx, float cy,
                                            float cz, float4* __restrict__ predOut,
                                            const float4* __restrict__ pos,
                                            float4* __restrict__ vel, float4* __restrict__ posOut,
                                            int last, const PbfConsts& c, const PassOrder& po)
{
  float4 q;
  q.x = p.x + cx / c.p.restDensity;
  q.y = p.y + cy / c.p.restDensity;
  q.z = p.z + cz / c.p.restDensity;
  q.w = 0.0f;
  if (last == kNextIter) {
    q.x = clampf(q.x, c.lo[0], c.hi[0]);
    q.y = clampf(q.y, c.lo[1], c.hi[1]);
    q.z = clampf(q.z, c.lo[2], c.hi[2]);
  }
  predOut[i] = q;
  record_cell(i, q, c, po);
  if (last == kNextIter)
    return;
  if (last == kLastOnly)
    return;
  const float4 o = pos[i];
  float4 v;
  v.x = clampf((q.x - o.x) / c.dtEps, -c.maxVel, c.maxVel);
  v.y = clampf((q.y - o.y) / c.dtEps, -c.maxVel, c.maxVel);
  v.z = clampf((q.z - o.z) / c.dtEps, -c.maxVel, c.maxVel);
  v.w = 0.0f;
  vel[i] = v;
  if (posOut)
    posOut[i] = q;
}

template <int MODE>
__device__ __forceinline__ void correct_body(uint32_t i, const float4* __restrict__ pred,
                                             const float* __restrict__ lambda,
                                             const uint2* __restrict__ se,
                                             float4* __restrict__ predOut,
                                             const float4* __restrict__ pos,
                                             float4* __restrict__ vel, float4* __restrict__ posOut,
                                             int last, const PbfConsts& c, const LearnArgs& la,
                                             const WalkCtx& T, const PassOrder& po)
{
  (void)la;
  constexpr bool CL = (MODE & kClampIn) != 0;
  const float4 p = ldp<CL>(pred, i, c);
  const float li = lambda[i];
  float cx = 0.0f, cy = 0.0f, cz = 0.0f;
  const bool fast = (c.shortMath & kShortSqrtGrad) != 0;
  const float ywdq = (c.shortMath & kShortArtDiv) ? rcp_refined(c.wdq) : 0.0f;
  walk27<CL, (MODE & (kNbRecord | kNbReplay))>(p, pred, se, c, [&](uint32_t e) { return lambda[e]; },
             [&](uint32_t, float le, float dx, float dy, float dz, float d2) {
    const float r = fast ? sqrt_cr(d2) : __builtin_sqrtf(d2);
    if (fast && r < c.h) {
      const float s = li + le + art_pressure_in(poly6_in(r, c), ywdq, c);
      // r <= kFloatEps: gradSpiky is the zero vector
      const Grad g = grad_spiky_in(dx, dy, dz, r, rcp_refined(r), c);
      const bool on = r > kFloatEps;
      cx = cx + s * (on ? g.x : 0.0f);
      cy = cy + s * (on ? g.y : 0.0f);
      cz = cz + s * (on ? g.z : 0.0f);
      return;
    }
    const float t = r < c.h ? 1.0f : 0.0f;
    const float s = li + le + art_pressure(poly6(r, t, c), c);
    if (r > kFloatEps) {
      const Grad g = grad_spiky(dx, dy, dz, r, t, c);
      cx = cx + s * g.x;
      cy = cy + s * g.y;
      cz = cz + s * g.z;
    } else {  // gradSpiky returns the zero vector
      cx = cx + s * 0.0f;
      cy = cy + s * 0.0f;
      cz = cz + s * 0.0f;
    }
  }, T);
  correct_out(i, p, cx, cy, cz, predOut, pos, vel, posOut, last, c, po);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_correct(const float4* __restrict__ pred,
                                                 const float* __restrict__ lambda,
                                                 const uint2* __restrict__ se,
                                                 float4* __restrict__ predOut,
                                                 const float4* __restrict__ pos,
                                                 float4* __restrict__ vel,
                                                 float4* __restrict__ posOut, uint32_t n,
                                                 int last, PbfConsts c, int remap, LearnArgs la, uint32_t qdepth,
                     PassOrder po)
{
  const uint32_t t = particle_index(remap);
  RTP_PLAIN_CTX(qdepth)
  RTP_NB_CTX
  if (t < n)
    correct_body<MODE>(po.particle(t), pred, lambda, se, predOut, pos, vel, posOut, last, c, la, W, po);
}

// ------------------------------------------------- learned substep: per-particle fallback
// No reference counterpart.  After the learned iterations, particles whose post-correction
// |rho/rho0 - 1| exceeds tau seed the fallback set; a seed and every particle within h of it
// in the seed's walk run the plain fallback iterations, over a compacted list.

// Append the lanes with `on` to list at count[0] (one atomic per wave).
__device__ __forceinline__ void wave_append(bool on, uint32_t value, uint32_t* __restrict__ list,
                                            uint32_t* __restrict__ count)
{
  const uint64_t m = __builtin_amdgcn_ballot_w64(on);
  if (m == 0)
    return;
  const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
  const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
  const uint32_t leader = (uint32_t)__builtin_ctzll(m);
  uint32_t base = 0;
  if (lane == leader)
    base = atomicAdd(count, (uint32_t)__builtin_popcountll(m));
  base = __shfl(base, (int)leader, 64);
  if (on)
    list[base + below] = value;
}

// Seeds: |density / rho0 - 1| > tau (the expression of the density-error metric); their
// flags in act, their indices in fs.seeds.
__global__ __launch_bounds__(256) void k_fb_mark(const float* __restrict__ resDensity, uint32_t n,
                                                 float rho0, float tau, FallbackSet fs)
{
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  bool seed = false;
  if (t < n) {
    seed = fabsf(resDensity[t] / rho0 - 1.0f) > tau;
    fs.act[t] = seed ? 1u : 0u;
  }
  wave_append(seed, t, fs.seeds, fs.count);
}

// Cell k (0..26, x outer and z inner, periodic wrap) of the 27 that walk_cells visits around
// position p (fluids.cl:101-123).
__device__ __forceinline__ uint32_t nb_cell(const float4 p, uint32_t k, const GridConsts& g)
{
  uint32_t cx, cy, cz;
  cell3d(p, g, cx, cy, cz);
  const int32_t ix = (int32_t)(k / 9u) - 1, iy = (int32_t)((k / 3u) % 3u) - 1,
                iz = (int32_t)(k % 3u) - 1;
  return (uint32_t)((wrap((int32_t)cx + ix, g.res[0]) * g.res[1] + wrap((int32_t)cy + iy, g.res[1])) *
                        g.res[2] +
                    wrap((int32_t)cz + iz, g.res[2]));
}

// The candidates of one neighbour cell of p, in index order: f(j, sqrt(d2) < h), d2 as the
// walks compute it, ((dx dx + dy dy) + dz dz).
template <class F>
__device__ __forceinline__ void scan_cell_within_h(const float4 p, uint2 r,
                                                   const float4* __restrict__ pred,
                                                   const PbfConsts& c, F&& f)
{
  for (uint32_t j = r.x; j <= r.y; ++j) {
    const float4 o = pred[j];
    const float dx = p.x - o.x, dy = p.y - o.y, dz = p.z - o.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    f(j, __builtin_sqrtf(d2) < c.h);
  }
}

// Grows the set (flag 1: from the seeds) or the halo (flag 2: from the set's members).  Item
// (k, cell) scans neighbour cell `cell` of particle src[k]; every particle it meets within h
// whose act is 0 takes the flag (plain stores of one value: concurrent writers agree).  The
// last 27 items are sorted element 0's own cells: element 0 is in no cell list (fillStartCell
// skips it), so no other particle's cells meet it; it takes the flag when it meets, within h,
// a seed (flag 1: |resDensity / rho0 - 1| > tau) or a set member (flag 2).  The same cells
// and distances as the walks, so the same particles join.
__global__ __launch_bounds__(256) void k_fb_grow(const float4* __restrict__ pred,
                                                 const uint2* __restrict__ se,
                                                 const uint32_t* __restrict__ src,
                                                 const uint32_t* __restrict__ srcCount,
                                                 uint32_t flag, const float* __restrict__ resDensity,
                                                 float tau, uint32_t n, PbfConsts c, uint32_t* act)
{
  const uint32_t items = *srcCount * 27u;
  const uint32_t total = n > 0 ? items + 27u : 0u;
  const float rho0 = c.p.restDensity;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += gridDim.x * blockDim.x) {
    if (t < items) {
      const float4 p = pred[src[t / 27u]];
      scan_cell_within_h(p, se[nb_cell(p, t % 27u, c.g)], pred, c, [&](uint32_t j, bool near) {
        if (near && act[j] == 0u)
          act[j] = flag;
      });
    } else {
      const float4 p = pred[0];
      bool hit = false;
      scan_cell_within_h(p, se[nb_cell(p, t - items, c.g)], pred, c, [&](uint32_t j, bool near) {
        if (near && (flag == 1u ? fabsf(resDensity[j] / rho0 - 1.0f) > tau : act[j] == 1u))
          hit = true;
      });
      if (hit && act[0] == 0u)
        act[0] = flag;
    }
  }
}

// The particles flagged `flag` in act, appended to out and counted in *count: per workgroup
// of 1024 one atomic on the counter (wave counts summed in LDS), so the atomics stay few.
__global__ __launch_bounds__(1024) void k_fb_compact(uint32_t n, uint32_t flag,
                                                     const uint32_t* __restrict__ act,
                                                     uint32_t* __restrict__ out,
                                                     uint32_t* __restrict__ count)
{
  __shared__ uint32_t waveBase[16];
  __shared__ uint32_t blockBase;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool on = t < n && act[t] == flag;
  const uint64_t m = __builtin_amdgcn_ballot_w64(on);
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  if (lane == 0u)
    waveBase[wave] = (uint32_t)__builtin_popcountll(m);
  __syncthreads();
  if (threadIdx.x == 0u) {
    uint32_t sum = 0;
    for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) {
      const uint32_t k = waveBase[w];
      waveBase[w] = sum;
      sum += k;
    }
    blockBase = sum ? atomicAdd(count, sum) : 0u;
  }
  __syncthreads();
  if (on) {
    const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    out[blockBase + waveBase[wave] + below] = t;
  }
}

// The start of the fallback iterations when the set is not empty: fld_applyBoundaryCondition
// on every particle, into pred and its partner (the iterations alternate between the two:
// later ones keep the outsiders' positions, and clamp the set's own corrections), and lambda
// 0 outside the set (the set's own is written by the density pass).  An empty set leaves the
// substep as it was.
__global__ __launch_bounds__(256) void k_fb_begin(float4* __restrict__ pred,
                                                  float4* __restrict__ partner,
                                                  float* __restrict__ lambda, uint32_t n,
                                                  PbfConsts c, FallbackSet fs)
{
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || fs.count[1] == 0u)
    return;
  const float4 q = clamp_pos(pred[t], c);
  pred[t] = q;
  partner[t] = q;
  lambda[t] = 0.0f;
}

// The set's corrected positions, left in the partner buffer by an odd number of iterations,
// back into place.
__global__ __launch_bounds__(256) void k_fb_scatter(float4* __restrict__ pred,
                                                    const float4* __restrict__ corrected,
                                                    FallbackSet fs)
{
  const uint32_t cnt = fs.count[1];
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += gridDim.x * blockDim.x) {
    const uint32_t i = fs.list[k];
    pred[i] = corrected[i];
  }
}

// ------------------------------------------------------- quad walks of the fallback passes
// A fallback pass covers a few thousand particles: one lane per particle would leave most
// SIMDs idle while each wave waits out its own long chain of loads.  Here a particle has a
// quad of four lanes.  The quad scans its candidates sixteen at a time (lane u takes
// candidates u, u + 4, u + 8 and u + 12 of each step) into a survivor queue in LDS, in walk
// order, and then runs the bodies of four survivors at once, one per lane; the four terms
// enter every sum in walk order (a DPP broadcast within the quad and one add each), so each
// sum adds the same terms in the same order as walk_cells' one-lane walk.  Lanes without a
// survivor add +0, which leaves every sum (started at +0) unchanged.
constexpr uint32_t kQuadBlock = 64;                // threads per block: one wave, 16 particles
constexpr uint32_t kQuadGroups = kQuadBlock / 4u;
constexpr uint32_t kQuadQueue = 64;                // queued survivors per particle
constexpr uint32_t kQuadRanges = 27;               // index ranges per particle (<= one per cell)

struct QuadLds {
  uint32_t queue[kQuadGroups][kQuadQueue];
  uint2 ranges[kQuadGroups][kQuadRanges];
};

// Survivor-list entry k of slot s: word k % 4 of the slot's granule k / 4.
__device__ __forceinline__ uint32_t* nb_word(const NbList& nb, uint32_t s, uint32_t k)
{
  return reinterpret_cast<uint32_t*>(nb.list + s + (size_t)(k >> 2) * nb.stride) + (k & 3u);
}

// walk_cells' index ranges of position p in its order, empty ones left out: per (x, y)
// column one range when its non-empty cells are back to back, else one per non-empty cell.
// Returns their number (at most 27).
__device__ __forceinline__ uint32_t quad_ranges(const float4 p, const uint2* __restrict__ se,
                                                const GridConsts& g, uint2* out)
{
  uint32_t cx, cy, cz;
  cell3d(p, g, cx, cy, cz);
  const int32_t rx = g.res[0], ry = g.res[1], rz = g.res[2];
  const int32_t z0 = wrap((int32_t)cz - 1, rz), z1 = wrap((int32_t)cz, rz),
                z2 = wrap((int32_t)cz + 1, rz);
  uint32_t nr = 0;
#pragma unroll
  for (int ix = -1; ix <= 1; ++ix) {
    const int32_t nx = wrap((int32_t)cx + ix, rx);
#pragma unroll
    for (int iy = -1; iy <= 1; ++iy) {
      const int32_t row = (nx * ry + wrap((int32_t)cy + iy, ry)) * rz;
      const uint2 cells[3] = { se[row + z0], se[row + z1], se[row + z2] };
      uint32_t lo = 1u, hi = 0u;
      bool ok = true;
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
          ok = false;
        }
      }
      if (ok) {
        if (lo <= hi)
          out[nr++] = make_uint2(lo, hi);
      } else {
#pragma unroll
        for (int k = 0; k < 3; ++k)
          if (cells[k].x <= cells[k].y)
            out[nr++] = cells[k];
      }
    }
  }
  return nr;
}

// Runs the body over survivors 0..cnt-1 of the quad (src(k): survivor k), four at a time
// (lane u: survivors u, u + 4, ...), loading each round's data during the previous round's
// bodies; rec(v, k, e) sees every round first.  The wave loops until its longest quad is
// done; empty lanes run the body on the particle itself and add +0.
template <class S, class R, class B>
__device__ __forceinline__ void quad_run(uint32_t cnt, uint32_t self, const float4 p,
                                         const float4* __restrict__ pred, const PbfConsts& c,
                                         S&& src, R&& rec, B& b)
{
  const uint32_t u = threadIdx.x & 3u;
  bool v = u < cnt;
  uint32_t e = v ? src(u) : self;
  float4 o = ldn<false>(pred, e, c);
  auto x = b.ld(e);
  for (uint32_t k = u; __builtin_amdgcn_ballot_w64(k - u < cnt) != 0; k += 4u) {
    const uint32_t k1 = k + 4u;
    const bool v1 = k1 < cnt;
    const uint32_t e1 = v1 ? src(k1) : self;
    const float4 o1 = ldn<false>(pred, e1, c);
    const auto x1 = b.ld(e1);
    rec(v, k, e);
    b.add(v, p, o, x);
    v = v1;
    e = e1;
    o = o1;
    x = x1;
  }
}

// The quad's walk over its nr ranges rl (in LDS): candidates are queued (the test of offer)
// and the queues drained by quad_run whenever one of the wave's quads has fewer than 16
// free entries, and at the end.  With REC the survivors go to the quad's survivor list
// (slot `slot`, the first kNbCap of them) and their number to its count.
template <bool REC, class B>
__device__ __forceinline__ void quad_walk(uint32_t nr, const uint2* rl, uint32_t* q,
                                          uint32_t self, const float4 p,
                                          const float4* __restrict__ pred, const PbfConsts& c,
                                          B& b, const NbList& nb, uint32_t slot, bool record)
{
  const uint32_t u = threadIdx.x & 3u, below = (1u << u) - 1u;
  uint32_t ri = 0, cur = 1u, hi = 0u, qn = 0, tot = 0;
  if (nr > 0) {
    const uint2 r = rl[0];
    cur = r.x;
    hi = r.y;
  }
  auto drain = [&]() {
    wave_sync();
    quad_run(qn, self, p, pred, c, [&](uint32_t k) { return q[k]; },
             [&](bool v, uint32_t k, uint32_t e) {
               if (REC && v & record & (tot + k < kNbCap))
                 *nb_word(nb, slot, tot + k) = e;
             }, b);
    tot += qn;
    qn = 0;
    wave_sync();
  };
  while (__builtin_amdgcn_ballot_w64(ri < nr) != 0) {
    const bool act = ri < nr;
    uint32_t e[4];
    bool keep[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      e[j] = cur + u + 4u * (uint32_t)j;
      const bool valid = act & (e[j] <= hi);
      const float4 o = ldn<false>(pred, valid ? e[j] : self, c);
      const float dx = p.x - o.x, dy = p.y - o.y, dz = p.z - o.z;
      const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
      keep[j] = valid & !((d2 >= c.cullFused) & (d2 < c.farFused));
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t m = quad_bits(keep[j]);
      if (keep[j])
        q[qn + (uint32_t)__builtin_popcount(m & below)] = e[j];
      qn += (uint32_t)__builtin_popcount(m);
    }
    cur += 16u;
    if (act & (cur > hi)) {
      ++ri;
      if (ri < nr) {
        const uint2 r = rl[ri];
        cur = r.x;
        hi = r.y;
      }
    }
    if (__builtin_amdgcn_ballot_w64(qn > kQuadQueue - 16u) != 0)
      drain();
  }
  drain();
  if (REC && record && u == 0u)
    nb.count[slot] = tot;
}

// density_lambda_body's sums (MODE 0) as quad terms.
struct QuadDensity {
  const PbfConsts& c;
  bool fast;
  float dens = 0.0f, sgx = 0.0f, sgy = 0.0f, sgz = 0.0f, ssq = 0.0f;
  __device__ __forceinline__ int ld(uint32_t) const { return 0; }
  __device__ __forceinline__ void add(bool v, const float4 p, const float4 o, int)
  {
    const float dx = p.x - o.x, dy = p.y - o.y, dz = p.z - o.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float r = fast ? sqrt_cr(d2) : __builtin_sqrtf(d2);
    float w, gx = 0.0f, gy = 0.0f, gz = 0.0f, sq = 0.0f;
    if (fast && r < c.h) {
      w = poly6_in(r, c);
      const Grad g = grad_spiky_in(dx, dy, dz, r, rcp_refined(r), c);
      const bool on = r > kFloatEps;
      gx = on ? g.x : 0.0f;
      gy = on ? g.y : 0.0f;
      gz = on ? g.z : 0.0f;
      sq = gx * gx + gy * gy + gz * gz;
    } else {
      const float t = r < c.h ? 1.0f : 0.0f;
      w = poly6(r, t, c);
      if (r > kFloatEps) {  // else the reference adds no gradient: +0 here
        const Grad g = grad_spiky(dx, dy, dz, r, t, c);
        gx = g.x;
        gy = g.y;
        gz = g.z;
        sq = gx * gx + gy * gy + gz * gz;
      }
    }
    dens = quad_add(dens, v ? w : 0.0f);
    sgx = quad_add(sgx, v ? gx : 0.0f);
    sgy = quad_add(sgy, v ? gy : 0.0f);
    sgz = quad_add(sgz, v ? gz : 0.0f);
    ssq = quad_add(ssq, v ? sq : 0.0f);
  }
};

// correct_body's sums as quad terms.
struct QuadCorrect {
  const PbfConsts& c;
  const float* __restrict__ lambda;
  bool fast;
  float li, ywdq;
  float cx = 0.0f, cy = 0.0f, cz = 0.0f;
  __device__ __forceinline__ float ld(uint32_t e) const { return lambda[e]; }
  __device__ __forceinline__ void add(bool v, const float4 p, const float4 o, float le)
  {
    const float dx = p.x - o.x, dy = p.y - o.y, dz = p.z - o.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float r = fast ? sqrt_cr(d2) : __builtin_sqrtf(d2);
    float tx, ty, tz;
    if (fast && r < c.h) {
      const float s = li + le + art_pressure_in(poly6_in(r, c), ywdq, c);
      const Grad g = grad_spiky_in(dx, dy, dz, r, rcp_refined(r), c);
      const bool on = r > kFloatEps;
      tx = s * (on ? g.x : 0.0f);
      ty = s * (on ? g.y : 0.0f);
      tz = s * (on ? g.z : 0.0f);
    } else {
      const float t = r < c.h ? 1.0f : 0.0f;
      const float s = li + le + art_pressure(poly6(r, t, c), c);
      if (r > kFloatEps) {
        const Grad g = grad_spiky(dx, dy, dz, r, t, c);
        tx = s * g.x;
        ty = s * g.y;
        tz = s * g.z;
      } else {  // gradSpiky returns the zero vector
        tx = s * 0.0f;
        ty = s * 0.0f;
        tz = s * 0.0f;
      }
    }
    cx = quad_add(cx, v ? tx : 0.0f);
    cy = quad_add(cy, v ? ty : 0.0f);
    cz = quad_add(cz, v ? tz : 0.0f);
  }
};

// density_lambda_body's outputs (MODE 0) from a quad's sums.
__device__ __forceinline__ void density_lambda_out(uint32_t i, const QuadDensity& b,
                                                   float* __restrict__ density,
                                                   float* __restrict__ lambda, const PbfConsts& c)
{
  density[i] = b.dens;
  const float C = b.dens / c.p.restDensity - 1.0f;
  float sq = b.ssq + (b.sgx * b.sgx + b.sgy * b.sgy + b.sgz * b.sgz);
  sq = sq / c.rho0sq;
  lambda[i] = -C / (sq + c.p.relaxCFM);
}

// The quad's survivors of particle i, into the body b: with NB == kNbReplay its survivor list
// (slot `slot`) when that holds all of them, else its walk, recorded (slot `slot`) with
// NB == kNbRecord when `record`.
template <int NB, class B>
__device__ __forceinline__ void quad_neighbours(bool on, uint32_t i, uint32_t slot, bool record,
                                                const float4 p, const float4* __restrict__ pred,
                                                const uint2* __restrict__ se, const PbfConsts& c,
                                                const NbList& nb, QuadLds& lds, B& b)
{
  const uint32_t g = threadIdx.x >> 2;
  bool listed = false;
  if constexpr (NB == kNbReplay) {
    const uint32_t lc = on ? nb.count[slot] : 0u;
    listed = on & (lc <= kNbCap);
    quad_run(listed ? lc : 0u, i, p, pred, c, [&](uint32_t m) { return *nb_word(nb, slot, m); },
             [](bool, uint32_t, uint32_t) {}, b);
  }
  const uint32_t nr = (on & !listed) ? quad_ranges(p, se, c.g, lds.ranges[g]) : 0u;
  // a recording walk writes every recorded count, also of a particle without candidates
  if (NB == kNbRecord || __builtin_amdgcn_ballot_w64(nr > 0u) != 0)
    quad_walk<NB == kNbRecord>(nr, lds.ranges[g], lds.queue[g], i, p, pred, c, b, nb, slot,
                               record);
}

// fld_computeDensity + fld_computeConstraintFactor over the set and then the halo, a quad
// per particle.  With REC the set's survivors go to survivor-list slot k (the position in
// the set) for the correction.
template <bool REC>
__global__ __launch_bounds__(kQuadBlock) void k_density_lambda_quad(
    const float4* __restrict__ pred, const uint2* __restrict__ se, float* __restrict__ density,
    float* __restrict__ lambda, PbfConsts c, FallbackSet fs, NbList nb)
{
  __shared__ QuadLds lds;
  const uint32_t cs = fs.count[1], cnt = cs + fs.count[2];
  const uint32_t g = threadIdx.x >> 2;
  const bool fast = (c.shortMath & kShortSqrtGrad) != 0;
  for (uint32_t base = blockIdx.x * kQuadGroups; base < cnt; base += gridDim.x * kQuadGroups) {
    const uint32_t k = base + g;
    const bool on = k < cnt;
    const uint32_t i = on ? (k < cs ? fs.list[k] : fs.halo[k - cs]) : 0u;
    const float4 p = ldp<false>(pred, i, c);
    QuadDensity b{ c, fast };
    quad_neighbours<REC ? kNbRecord : 0>(on, i, k, on & (k < cs), p, pred, se, c, nb, lds, b);
    if (on && (threadIdx.x & 3u) == 0u)
      density_lambda_out(i, b, density, lambda, c);
  }
}

// fld_computeConstraintCorrection + fld_correctPosition over the set, a quad per particle
// (and, unless `last`, the next iteration's boundary condition).  With REPLAY a particle
// whose survivor list holds all its survivors runs the list; the others walk.
template <bool REPLAY>
__global__ __launch_bounds__(kQuadBlock) void k_correct_quad(
    const float4* __restrict__ pred, const float* __restrict__ lambda,
    const uint2* __restrict__ se, float4* __restrict__ predOut, int last, PbfConsts c,
    FallbackSet fs, NbList nb)
{
  __shared__ QuadLds lds;
  const uint32_t cnt = fs.count[1];
  const uint32_t g = threadIdx.x >> 2;
  const bool fast = (c.shortMath & kShortSqrtGrad) != 0;
  const float ywdq = (c.shortMath & kShortArtDiv) ? rcp_refined(c.wdq) : 0.0f;
  const PassOrder po;  // no cell keys to record
  for (uint32_t base = blockIdx.x * kQuadGroups; base < cnt; base += gridDim.x * kQuadGroups) {
    const uint32_t k = base + g;
    const bool on = k < cnt;
    const uint32_t i = on ? fs.list[k] : 0u;
    const float4 p = ldp<false>(pred, i, c);
    QuadCorrect b{ c, lambda, fast, lambda[i], ywdq };
    quad_neighbours<REPLAY ? kNbReplay : 0>(on, i, k, false, p, pred, se, c, nb, lds, b);
    if (on && (threadIdx.x & 3u) == 0u)
      correct_out(i, p, b.cx, b.cy, b.cz, predOut, nullptr, nullptr, nullptr, last, c, po);
  }
}

// fld_updateVel (+ fld_updatePosition) for a step with zero Jacobi iterations.
__device__ __forceinline__ void update_vel_body(uint32_t i, const float4* __restrict__ pred,
                                                const float4* __restrict__ pos,
                                                float4* __restrict__ vel, float4* __restrict__ posOut,
                                                const PbfConsts& c)
{
  const float4 q = pred[i], o = pos[i];
  float4 v;
  v.x = clampf((q.x - o.x) / c.dtEps, -c.maxVel, c.maxVel);
  v.y = clampf((q.y - o.y) / c.dtEps, -c.maxVel, c.maxVel);
  v.z = clampf((q.z - o.z) / c.dtEps, -c.maxVel, c.maxVel);
  v.w = 0.0f;
  vel[i] = v;
  if (posOut)
    posOut[i] = make_float4(q.x, q.y, q.z, 0.0f);
}

__global__ __launch_bounds__(256) void k_update_vel(const float4* __restrict__ pred,
                                                    const float4* __restrict__ pos,
                                                    float4* __restrict__ vel,
                                                    float4* __restrict__ posOut, uint32_t n,
                                                    PbfConsts c)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    update_vel_body(i, pred, pos, vel, posOut, c);
}

// fld_computeVorticity (fluids.cl:278-324).  The w lane of the output carries
// fast_length(vort) for the confinement pass.
template <int NB = 0>
__device__ __forceinline__ void vorticity_body(uint32_t i, const float4* __restrict__ pred,
                                               const uint2* __restrict__ se,
                                               const float4* __restrict__ vel,
                                               float4* __restrict__ vort, const PbfConsts& c,
                                               const WalkCtx& T)
{
  const float4 p = pred[i];
  const float4 vi = vel[i];
  float wx = 0.0f, wy = 0.0f, wz = 0.0f;
  const bool fast = (c.shortMath & kShortSqrtGrad) != 0;
  walk27<false, NB>(p, pred, se, c, [&](uint32_t e) { return vel[e]; },
         [&](uint32_t, const float4 ve, float dx, float dy, float dz, float d2) {
    const float r = fast ? sqrt_cr(d2) : __builtin_sqrtf(d2);
    if (fast && r < c.h) {
      // nothing is added for r <= kFloatEps
      const Grad g = grad_spiky_in(dx, dy, dz, r, rcp_refined(r), c);
      const float ax = ve.x - vi.x, ay = ve.y - vi.y, az = ve.z - vi.z;
      const bool on = r > kFloatEps;
      wx = on ? wx + (ay * g.z - az * g.y) : wx;
      wy = on ? wy + (az * g.x - ax * g.z) : wy;
      wz = on ? wz + (ax * g.y - ay * g.x) : wz;
      return;
    }
    if (r > kFloatEps) {
      const float t = r < c.h ? 1.0f : 0.0f;
      const Grad g = grad_spiky(dx, dy, dz, r, t, c);
      const float ax = ve.x - vi.x, ay = ve.y - vi.y, az = ve.z - vi.z;
      wx = wx + (ay * g.z - az * g.y);
      wy = wy + (az * g.x - ax * g.z);
      wz = wz + (ax * g.y - ay * g.x);
    }
  }, T);
  vort[i] = make_float4(wx, wy, wz, __builtin_sqrtf(wx * wx + wy * wy + wz * wz));
}

template <int NB>
__global__ __launch_bounds__(256) void k_vorticity(const float4* __restrict__ pred,
                                                   const uint2* __restrict__ se,
                                                   const float4* __restrict__ vel,
                                                   float4* __restrict__ vort, uint32_t n,
                                                   PbfConsts c, int remap, uint32_t qdepth,
                     PassOrder po)
{
  const uint32_t t = particle_index(remap);
  RTP_PLAIN_CTX(qdepth)
  RTP_NB_CTX
  if (t < n)
    vorticity_body<NB>(po.particle(t), pred, se, vel, vort, c, W);
}

// The confined velocity of particle i from its sum (nx, ny, nz) of |vort_j| gradW.
__device__ __forceinline__ void confinement_out(uint32_t i, float nx, float ny, float nz,
                                                const float4* __restrict__ vort,
                                                const float4* __restrict__ vel,
                                                float4* __restrict__ velOut, const PbfConsts& c)
{
  // normalize(n), 0 for the zero vector
  if (!(nx == 0.0f && ny == 0.0f && nz == 0.0f)) {
    const float len = __builtin_sqrtf(nx * nx + ny * ny + nz * nz);
    nx = nx / len;
    ny = ny / len;
    nz = nz / len;
  }
  const float4 w = vort[i];
  const float crx = ny * w.z - nz * w.y;
  const float cry = nz * w.x - nx * w.z;
  const float crz = nx * w.y - ny * w.x;
  const float k = c.p.vorticityConfCoeff, dt = c.p.timeStep;
  const float4 v = vel[i];
  velOut[i] = make_float4(v.x + (k * crx) * dt, v.y + (k * cry) * dt, v.z + (k * crz) * dt, 0.0f);
}

// fld_applyVorticityConfinement (fluids.cl:329-377), writing the confined velocity to
// velOut (the copy fld_applyXsphViscosityCorrection reads, Fluids.cpp:451).
template <int NB = 0>
__device__ __forceinline__ void confinement_body(uint32_t i, const float4* __restrict__ pred,
                                                 const uint2* __restrict__ se,
                                                 const float4* __restrict__ vort,
                                                 const float4* __restrict__ vel,
                                                 float4* __restrict__ velOut, const PbfConsts& c,
                                                 const WalkCtx& T)
{
  const float4 p = pred[i];
  float nx = 0.0f, ny = 0.0f, nz = 0.0f;
  const bool fast = (c.shortMath & kShortSqrtGrad) != 0;
  walk27<false, NB>(p, pred, se, c, [&](uint32_t e) { return vort[e].w; },
         [&](uint32_t, float l, float dx, float dy, float dz, float d2) {
    const float r = fast ? sqrt_cr(d2) : __builtin_sqrtf(d2);
    if (fast && r < c.h) {
      // nothing is added for r <= kFloatEps
      const Grad g = grad_spiky_in(dx, dy, dz, r, rcp_refined(r), c);
      const bool on = r > kFloatEps;
      nx = on ? nx + l * g.x : nx;
      ny = on ? ny + l * g.y : ny;
      nz = on ? nz + l * g.z : nz;
      return;
    }
    if (r > kFloatEps) {
      const float t = r < c.h ? 1.0f : 0.0f;
      const Grad g = grad_spiky(dx, dy, dz, r, t, c);
      nx = nx + l * g.x;
      ny = ny + l * g.y;
      nz = nz + l * g.z;
    }
  }, T);
  confinement_out(i, nx, ny, nz, vort, vel, velOut, c);
}

template <int NB>
__global__ __launch_bounds__(256) void k_confinement(const float4* __restrict__ pred,
                                                     const uint2* __restrict__ se,
                                                     const float4* __restrict__ vort,
                                                     const float4* __restrict__ vel,
                                                     float4* __restrict__ velOut, uint32_t n,
                                                     PbfConsts c, int remap, uint32_t qdepth,
                     PassOrder po)
{
  const uint32_t t = particle_index(remap);
  RTP_PLAIN_CTX(qdepth)
  RTP_NB_CTX
  if (t < n)
    confinement_body<NB>(po.particle(t), pred, se, vort, vel, velOut, c, W);
}

// fld_applyXsphViscosityCorrection (fluids.cl:383-430) + fld_updatePosition (:444-450).
template <int NB = 0>
__device__ __forceinline__ void xsph_body(uint32_t i, const float4* __restrict__ pred,
                                          const uint2* __restrict__ se,
                                          const float4* __restrict__ velIn,
                                          float4* __restrict__ velOut, float4* __restrict__ posOut,
                                          const PbfConsts& c, const WalkCtx& T)
{
  const float4 p = pred[i];
  const float4 vi = velIn[i];
  float sx = 0.0f, sy = 0.0f, sz = 0.0f;
  const bool fast = (c.shortMath & kShortSqrtGrad) != 0;
  walk27<false, NB>(p, pred, se, c, [&](uint32_t e) { return velIn[e]; },
         [&](uint32_t, const float4 ve, float, float, float, float d2) {
    const float r = fast ? sqrt_cr(d2) : __builtin_sqrtf(d2);
    const float t = r < c.h ? 1.0f : 0.0f;
    const float w = (fast && r < c.h) ? poly6_in(r, c) : poly6(r, t, c);
    sx = sx + (ve.x - vi.x) * w;
    sy = sy + (ve.y - vi.y) * w;
    sz = sz + (ve.z - vi.z) * w;
  }, T);
  const float k = c.p.xsphViscosityCoeff;
  velOut[i] = make_float4(vi.x + k * sx, vi.y + k * sy, vi.z + k * sz, 0.0f);
  posOut[i] = make_float4(p.x, p.y, p.z, 0.0f);
}

template <int NB>
__global__ __launch_bounds__(256) void k_xsph(const float4* __restrict__ pred,
                                              const uint2* __restrict__ se,
                                              const float4* __restrict__ velIn,
                                              float4* __restrict__ velOut,
                                              float4* __restrict__ posOut, uint32_t n,
                                              PbfConsts c, int remap, uint32_t qdepth,
                     PassOrder po)
{
  const uint32_t t = particle_index(remap);
  RTP_PLAIN_CTX(qdepth)
  RTP_NB_CTX
  if (t < n)
    xsph_body<NB>(po.particle(t), pred, se, velIn, velOut, posOut, c, W);
}

// ------------------------------------------------------- quad passes of the plain solver
// With Launch::quad the neighbour passes of the plain solver run a quad of lanes per
// particle (the fallback passes' walks above): thread order t, survivor-list slot t, the
// same sums in the same order.  For scenes too small to give every SIMD several waves with
// one lane per particle.

// vorticity_body's sums as quad terms.
struct QuadVorticity {
  const PbfConsts& c;
  const float4* __restrict__ vel;
  bool fast;
  float4 vi;
  float wx = 0.0f, wy = 0.0f, wz = 0.0f;
  __device__ __forceinline__ float4 ld(uint32_t e) const { return vel[e]; }
  __device__ __forceinline__ void add(bool v, const float4 p, const float4 o, const float4 ve)
  {
    const float dx = p.x - o.x, dy = p.y - o.y, dz = p.z - o.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float r = fast ? sqrt_cr(d2) : __builtin_sqrtf(d2);
    float tx = 0.0f, ty = 0.0f, tz = 0.0f;
    if (fast && r < c.h) {
      // nothing is added for r <= kFloatEps
      const Grad g = grad_spiky_in(dx, dy, dz, r, rcp_refined(r), c);
      const float ax = ve.x - vi.x, ay = ve.y - vi.y, az = ve.z - vi.z;
      const bool on = r > kFloatEps;
      tx = on ? ay * g.z - az * g.y : 0.0f;
      ty = on ? az * g.x - ax * g.z : 0.0f;
      tz = on ? ax * g.y - ay * g.x : 0.0f;
    } else if (r > kFloatEps) {
      const float t = r < c.h ? 1.0f : 0.0f;
      const Grad g = grad_spiky(dx, dy, dz, r, t, c);
      const float ax = ve.x - vi.x, ay = ve.y - vi.y, az = ve.z - vi.z;
      tx = ay * g.z - az * g.y;
      ty = az * g.x - ax * g.z;
      tz = ax * g.y - ay * g.x;
    }
    wx = quad_add(wx, v ? tx : 0.0f);
    wy = quad_add(wy, v ? ty : 0.0f);
    wz = quad_add(wz, v ? tz : 0.0f);
  }
};

// confinement_body's sums as quad terms.
struct QuadConfinement {
  const PbfConsts& c;
  const float4* __restrict__ vort;
  bool fast;
  float nx = 0.0f, ny = 0.0f, nz = 0.0f;
  __device__ __forceinline__ float ld(uint32_t e) const { return vort[e].w; }
  __device__ __forceinline__ void add(bool v, const float4 p, const float4 o, float l)
  {
    const float dx = p.x - o.x, dy = p.y - o.y, dz = p.z - o.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float r = fast ? sqrt_cr(d2) : __builtin_sqrtf(d2);
    float tx = 0.0f, ty = 0.0f, tz = 0.0f;
    if (fast && r < c.h) {
      // nothing is added for r <= kFloatEps
      const Grad g = grad_spiky_in(dx, dy, dz, r, rcp_refined(r), c);
      const bool on = r > kFloatEps;
      tx = on ? l * g.x : 0.0f;
      ty = on ? l * g.y : 0.0f;
      tz = on ? l * g.z : 0.0f;
    } else if (r > kFloatEps) {
      const float t = r < c.h ? 1.0f : 0.0f;
      const Grad g = grad_spiky(dx, dy, dz, r, t, c);
      tx = l * g.x;
      ty = l * g.y;
      tz = l * g.z;
    }
    nx = quad_add(nx, v ? tx : 0.0f);
    ny = quad_add(ny, v ? ty : 0.0f);
    nz = quad_add(nz, v ? tz : 0.0f);
  }
};

// xsph_body's sums as quad terms.
struct QuadXsph {
  const PbfConsts& c;
  const float4* __restrict__ velIn;
  bool fast;
  float4 vi;
  float sx = 0.0f, sy = 0.0f, sz = 0.0f;
  __device__ __forceinline__ float4 ld(uint32_t e) const { return velIn[e]; }
  __device__ __forceinline__ void add(bool v, const float4 p, const float4 o, const float4 ve)
  {
    const float dx = p.x - o.x, dy = p.y - o.y, dz = p.z - o.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float r = fast ? sqrt_cr(d2) : __builtin_sqrtf(d2);
    const float t = r < c.h ? 1.0f : 0.0f;
    const float w = (fast && r < c.h) ? poly6_in(r, c) : poly6(r, t, c);
    sx = quad_add(sx, v ? (ve.x - vi.x) * w : 0.0f);
    sy = quad_add(sy, v ? (ve.y - vi.y) * w : 0.0f);
    sz = quad_add(sz, v ? (ve.z - vi.z) * w : 0.0f);
  }
};

// The quad's thread t, its particle and whether it has one (quads past n stand in for
// particle 0 and write nothing).
#define RTP_QUAD_PROLOGUE                                                     \
  __shared__ QuadLds lds;                                                     \
  const uint32_t t = blockIdx.x * kQuadGroups + (threadIdx.x >> 2);           \
  const bool on = t < n;                                                      \
  const uint32_t i = on ? po.particle(t) : 0u;                                \
  const bool lead = on && (threadIdx.x & 3u) == 0u;                           \
  const bool fast = (c.shortMath & kShortSqrtGrad) != 0;

// k_density_lambda (MODE 0 or kNbRecord) with a quad per particle.
template <int NB>
__global__ __launch_bounds__(kQuadBlock) void k_density_lambda_q(const float4* __restrict__ pred,
                                                                 const uint2* __restrict__ se,
                                                                 float* __restrict__ density,
                                                                 float* __restrict__ lambda,
                                                                 uint32_t n, PbfConsts c,
                                                                 PassOrder po)
{
  RTP_QUAD_PROLOGUE
  const float4 p = ldp<false>(pred, i, c);
  QuadDensity b{ c, fast };
  quad_neighbours<NB>(on, i, t, on, p, pred, se, c, po.nb, lds, b);
  if (lead)
    density_lambda_out(i, b, density, lambda, c);
}

// k_correct (MODE 0 or kNbReplay) with a quad per particle.
template <int NB>
__global__ __launch_bounds__(kQuadBlock) void k_correct_q(const float4* __restrict__ pred,
                                                          const float* __restrict__ lambda,
                                                          const uint2* __restrict__ se,
                                                          float4* __restrict__ predOut,
                                                          const float4* __restrict__ pos,
                                                          float4* __restrict__ vel,
                                                          float4* __restrict__ posOut, uint32_t n,
                                                          int last, PbfConsts c, PassOrder po)
{
  RTP_QUAD_PROLOGUE
  const float4 p = ldp<false>(pred, i, c);
  const float ywdq = (c.shortMath & kShortArtDiv) ? rcp_refined(c.wdq) : 0.0f;
  QuadCorrect b{ c, lambda, fast, lambda[i], ywdq };
  quad_neighbours<NB>(on, i, t, false, p, pred, se, c, po.nb, lds, b);
  if (lead)
    correct_out(i, p, b.cx, b.cy, b.cz, predOut, pos, vel, posOut, last, c, po);
}

// k_vorticity with a quad per particle.
template <int NB>
__global__ __launch_bounds__(kQuadBlock) void k_vorticity_q(const float4* __restrict__ pred,
                                                            const uint2* __restrict__ se,
                                                            const float4* __restrict__ vel,
                                                            float4* __restrict__ vort, uint32_t n,
                                                            PbfConsts c, PassOrder po)
{
  RTP_QUAD_PROLOGUE
  const float4 p = pred[i];
  QuadVorticity b{ c, vel, fast, vel[i] };
  quad_neighbours<NB>(on, i, t, on, p, pred, se, c, po.nb, lds, b);
  if (lead)
    vort[i] = make_float4(b.wx, b.wy, b.wz, __builtin_sqrtf(b.wx * b.wx + b.wy * b.wy + b.wz * b.wz));
}

// k_confinement with a quad per particle.
template <int NB>
__global__ __launch_bounds__(kQuadBlock) void k_confinement_q(const float4* __restrict__ pred,
                                                              const uint2* __restrict__ se,
                                                              const float4* __restrict__ vort,
                                                              const float4* __restrict__ vel,
                                                              float4* __restrict__ velOut,
                                                              uint32_t n, PbfConsts c, PassOrder po)
{
  RTP_QUAD_PROLOGUE
  const float4 p = pred[i];
  QuadConfinement b{ c, vort, fast };
  quad_neighbours<NB>(on, i, t, false, p, pred, se, c, po.nb, lds, b);
  if (lead)
    confinement_out(i, b.nx, b.ny, b.nz, vort, vel, velOut, c);
}

// k_xsph with a quad per particle.
template <int NB>
__global__ __launch_bounds__(kQuadBlock) void k_xsph_q(const float4* __restrict__ pred,
                                                       const uint2* __restrict__ se,
                                                       const float4* __restrict__ velIn,
                                                       float4* __restrict__ velOut,
                                                       float4* __restrict__ posOut, uint32_t n,
                                                       PbfConsts c, PassOrder po)
{
  RTP_QUAD_PROLOGUE
  const float4 p = pred[i];
  QuadXsph b{ c, velIn, fast, velIn[i] };
  quad_neighbours<NB>(on, i, t, false, p, pred, se, c, po.nb, lds, b);
  if (lead) {
    const float k = c.p.xsphViscosityCoeff;
    velOut[i] = make_float4(b.vi.x + k * b.sx, b.vi.y + k * b.sy, b.vi.z + k * b.sz, 0.0f);
    posOut[i] = make_float4(p.x, p.y, p.z, 0.0f);
  }
}
#undef RTP_QUAD_PROLOGUE

// --------------------------------------------------------------- LDS-tiled wrappers
// Same passes with the wave tiles staged first.  Every lane takes part in staging (lanes
// past n stand in for the last particle), then only real particles run the pass.
#define RTP_TILED_PROLOGUE(CLV)                                                \
  __shared__ TileLds tl;                                                       \
  extern __shared__ uint32_t rtpQueue[];                                       \
  const uint32_t t0 = particle_index(remap);                                   \
  const bool active = t0 < n;                                                  \
  const uint32_t i = po.particle(active ? t0 : n - 1u);                        \
  WalkCtx T;                                                                   \
  T.qslot = &rtpQueue[threadIdx.x];                                            \
  T.qcap = qdepth;                                                             \
  build_tile<CLV>(ldp<CLV>(pred, i, c), pred, se, c, tl, T);                   \
  __syncthreads();

template <int MODE>
__global__ __launch_bounds__(256) void k_density_lambda_t(const float4* __restrict__ pred,
                                                          const uint2* __restrict__ se,
                                                          float* __restrict__ density,
                                                          float* __restrict__ lambda, uint32_t n,
                                                          PbfConsts c, int remap, LearnArgs la, uint32_t qdepth,
                     PassOrder po)
{
  RTP_TILED_PROLOGUE((MODE & kClampIn) != 0)
  if (active)
    density_lambda_body<MODE>(i, pred, se, density, lambda, c, la, T);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_correct_t(const float4* __restrict__ pred,
                                                   const float* __restrict__ lambda,
                                                   const uint2* __restrict__ se,
                                                   float4* __restrict__ predOut,
                                                   const float4* __restrict__ pos,
                                                   float4* __restrict__ vel,
                                                   float4* __restrict__ posOut, uint32_t n,
                                                   int last, PbfConsts c, int remap, LearnArgs la, uint32_t qdepth,
                     PassOrder po)
{
  RTP_TILED_PROLOGUE((MODE & kClampIn) != 0)
  if (active)
    correct_body<MODE>(i, pred, lambda, se, predOut, pos, vel, posOut, last, c, la, T, po);
}

__global__ __launch_bounds__(256) void k_vorticity_t(const float4* __restrict__ pred,
                                                     const uint2* __restrict__ se,
                                                     const float4* __restrict__ vel,
                                                     float4* __restrict__ vort, uint32_t n,
                                                     PbfConsts c, int remap, uint32_t qdepth,
                     PassOrder po)
{
  RTP_TILED_PROLOGUE(false)
  if (active)
    vorticity_body(i, pred, se, vel, vort, c, T);
}

__global__ __launch_bounds__(256) void k_confinement_t(const float4* __restrict__ pred,
                                                       const uint2* __restrict__ se,
                                                       const float4* __restrict__ vort,
                                                       const float4* __restrict__ vel,
                                                       float4* __restrict__ velOut, uint32_t n,
                                                       PbfConsts c, int remap, uint32_t qdepth,
                     PassOrder po)
{
  RTP_TILED_PROLOGUE(false)
  if (active)
    confinement_body(i, pred, se, vort, vel, velOut, c, T);
}

__global__ __launch_bounds__(256) void k_xsph_t(const float4* __restrict__ pred,
                                                const uint2* __restrict__ se,
                                                const float4* __restrict__ velIn,
                                                float4* __restrict__ velOut,
                                                float4* __restrict__ posOut, uint32_t n,
                                                PbfConsts c, int remap, uint32_t qdepth,
                     PassOrder po)
{
  RTP_TILED_PROLOGUE(false)
  if (active)
    xsph_body(i, pred, se, velIn, velOut, posOut, c, T);
}

#undef RTP_TILED_PROLOGUE

// ------------------------------------------------------------------- batched wrappers
// One launch covers every scene of a batch: block b works on scene blocks[b].x, particles
// blocks[b].y + threadIdx.x of it.  All particle pointers are shifted to the scene's first
// particle and the cell table to its key base, so each body runs exactly as in a model of
// its own (same indices, same summation order).

// A batched pass's survivor lists: slot = the thread's launch-wide id.
#define RTP_BATCH_NB_CTX \
  W.nb = bv.nb;          \
  W.slot = blockIdx.x * blockDim.x + threadIdx.x;

struct BatchSite {
  uint32_t i;       // scene-local particle
  uint32_t scene;
  SceneSlot slot;
};

__device__ __forceinline__ bool batch_site(const BatchView& bv, BatchSite& s)
{
  const uint2 blk = bv.blocks[blockIdx.x];
  s.scene = blk.x;
  s.slot = bv.scenes[blk.x];
  const uint32_t t = blk.y + threadIdx.x;
  if (t >= s.slot.n)
    return false;
  s.i = bv.order ? bv.order[s.slot.pbase + t] - s.slot.pbase : t;
  return true;
}

// The correction pass's key output of a batched scene: batch-wide keys.
__device__ __forceinline__ PassOrder batch_order(const BatchView& bv, const BatchSite& b)
{
  PassOrder po;
  if (bv.keyOut) {
    po.keyOut = bv.keyOut + b.slot.pbase;
    po.keyBase = b.slot.kbase;
  }
  return po;
}

__global__ __launch_bounds__(256) void k_fill_start_end(uint2* __restrict__ se, uint32_t count)
{
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count)
    se[t] = make_uint2(1u, 0u);  // grid.cl:91-96
}

__global__ __launch_bounds__(256) void k_predict_batch(BatchView bv, const float4* __restrict__ pos,
                                                       const float4* __restrict__ vel,
                                                       float4* __restrict__ pred,
                                                       uint32_t* __restrict__ key)
{
  BatchSite b;
  if (!batch_site(bv, b))
    return;
  const PbfConsts c = bv.consts[b.scene];
  const uint32_t o = b.slot.pbase;
  predict_body(b.i, pos + o, vel + o, pred + o, key + o, b.slot.kbase, c);
}

template <int NB>
__global__ __launch_bounds__(256) void k_density_lambda_batch(BatchView bv,
                                                              const float4* __restrict__ pred,
                                                              const uint2* __restrict__ se,
                                                              float* __restrict__ density,
                                                              float* __restrict__ lambda)
{
  BatchSite b;
  if (!batch_site(bv, b))
    return;
  const PbfConsts c = bv.consts[b.scene];
  const uint32_t o = b.slot.pbase;
  RTP_PLAIN_CTX(bv.qdepth)
  RTP_BATCH_NB_CTX
  density_lambda_body<NB>(b.i, pred + o, se + b.slot.kbase, density + o, lambda + o, c, LearnArgs(), W);
}

template <int NB>
__global__ __launch_bounds__(256) void k_correct_batch(BatchView bv, const float4* __restrict__ pred,
                                                       const float* __restrict__ lambda,
                                                       const uint2* __restrict__ se,
                                                       float4* __restrict__ predOut,
                                                       const float4* __restrict__ pos,
                                                       float4* __restrict__ vel,
                                                       float4* __restrict__ posOut, int last)
{
  BatchSite b;
  if (!batch_site(bv, b))
    return;
  const PbfConsts c = bv.consts[b.scene];
  const uint32_t o = b.slot.pbase;
  RTP_PLAIN_CTX(bv.qdepth)
  RTP_BATCH_NB_CTX
  correct_body<NB>(b.i, pred + o, lambda + o, se + b.slot.kbase, predOut + o, pos + o, vel + o,
                  posOut ? posOut + o : nullptr, last, c, LearnArgs(), W, batch_order(bv, b));
}

__global__ __launch_bounds__(256) void k_update_vel_batch(BatchView bv, const float4* __restrict__ pred,
                                                          const float4* __restrict__ pos,
                                                          float4* __restrict__ vel,
                                                          float4* __restrict__ posOut)
{
  BatchSite b;
  if (!batch_site(bv, b))
    return;
  const PbfConsts c = bv.consts[b.scene];
  const uint32_t o = b.slot.pbase;
  update_vel_body(b.i, pred + o, pos + o, vel + o, posOut ? posOut + o : nullptr, c);
}

template <int NB>
__global__ __launch_bounds__(256) void k_vorticity_batch(BatchView bv, const float4* __restrict__ pred,
                                                         const uint2* __restrict__ se,
                                                         const float4* __restrict__ vel,
                                                         float4* __restrict__ vort)
{
  BatchSite b;
  if (!batch_site(bv, b))
    return;
  const PbfConsts c = bv.consts[b.scene];
  const uint32_t o = b.slot.pbase;
  RTP_PLAIN_CTX(bv.qdepth)
  RTP_BATCH_NB_CTX
  vorticity_body<NB>(b.i, pred + o, se + b.slot.kbase, vel + o, vort + o, c, W);
}

template <int NB>
__global__ __launch_bounds__(256) void k_confinement_batch(BatchView bv,
                                                           const float4* __restrict__ pred,
                                                           const uint2* __restrict__ se,
                                                           const float4* __restrict__ vort,
                                                           const float4* __restrict__ vel,
                                                           float4* __restrict__ velOut)
{
  BatchSite b;
  if (!batch_site(bv, b))
    return;
  const PbfConsts c = bv.consts[b.scene];
  const uint32_t o = b.slot.pbase;
  RTP_PLAIN_CTX(bv.qdepth)
  RTP_BATCH_NB_CTX
  confinement_body<NB>(b.i, pred + o, se + b.slot.kbase, vort + o, vel + o, velOut + o, c, W);
}

template <int NB>
__global__ __launch_bounds__(256) void k_xsph_batch(BatchView bv, const float4* __restrict__ pred,
                                                    const uint2* __restrict__ se,
                                                    const float4* __restrict__ velIn,
                                                    float4* __restrict__ velOut,
                                                    float4* __restrict__ posOut)
{
  BatchSite b;
  if (!batch_site(bv, b))
    return;
  const PbfConsts c = bv.consts[b.scene];
  const uint32_t o = b.slot.pbase;
  RTP_PLAIN_CTX(bv.qdepth)
  RTP_BATCH_NB_CTX
  xsph_body<NB>(b.i, pred + o, se + b.slot.kbase, velIn + o, velOut + o, posOut + o, c, W);
}

__global__ __launch_bounds__(64) void k_zero3(float* __restrict__ out)
{
  if (threadIdx.x < 3)
    out[threadIdx.x] = 0.0f;
}

// Block-wide (max, sum, max) of three per-thread values, in a fixed order: the same
// inputs give the same bits on every run.  Thread 0 holds the result.
__device__ inline void block_reduce3(float& mx, float& sm, float& cm)
{
  __shared__ float smax[4], ssum[4], scmp[4];
  for (int o = 32; o > 0; o >>= 1) {
    mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    sm += __shfl_xor(sm, o, 64);
    cm = fmaxf(cm, __shfl_xor(cm, o, 64));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    smax[w] = mx;
    ssum[w] = sm;
    scmp[w] = cm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; ++k) {
      mx = fmaxf(mx, smax[k]);
      sm += ssum[k];
      cm = fmaxf(cm, scmp[k]);
    }
  }
}

// Density error |rho/rho0 - 1| (max, sum) and the largest compression, in two passes with
// a fixed grid of at most kErrBlocks blocks so the float sum does not depend on timing.
constexpr uint32_t kErrBlocks = 1024;
__global__ __launch_bounds__(256) void k_density_error(const float* __restrict__ density, uint32_t n,
                                                       float rho0, float* __restrict__ partial)
{
  float mx = 0.0f, sm = 0.0f, cm = 0.0f;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float e = density[i] / rho0 - 1.0f;
    const float a = fabsf(e);
    mx = fmaxf(mx, a);
    sm += a;
    cm = fmaxf(cm, e);
  }
  block_reduce3(mx, sm, cm);
  if (threadIdx.x == 0) {
    partial[3u * blockIdx.x] = mx;
    partial[3u * blockIdx.x + 1u] = sm;
    partial[3u * blockIdx.x + 2u] = cm;
  }
}

__global__ __launch_bounds__(256) void k_density_error_final(const float* __restrict__ partial,
                                                             uint32_t blocks, float* __restrict__ out)
{
  float mx = 0.0f, sm = 0.0f, cm = 0.0f;
  for (uint32_t b = threadIdx.x; b < blocks; b += blockDim.x) {
    mx = fmaxf(mx, partial[3u * b]);
    sm += partial[3u * b + 1u];
    cm = fmaxf(cm, partial[3u * b + 2u]);
  }
  block_reduce3(mx, sm, cm);
  if (threadIdx.x == 0) {
    out[0] = mx;
    out[1] = sm;
    out[2] = cm;
  }
}

inline uint32_t blocks_for(uint32_t n, uint32_t b) { return (n + b - 1u) / b; }

// Dynamic LDS of a neighbour pass: the survivor queue, depth x block size words.
inline uint32_t queue_bytes(const Launch& L) { return (uint32_t)(L.queue * L.block) * 4u; }

}  // namespace

void launch_pbf_predict(const float4* pos, const float4* vel, float4* pred, uint32_t* key,
                        uint2* startEnd, uint32_t n, const PbfConsts& c, hipStream_t s)
{
  const uint32_t m = n > c.g.numCells ? n : c.g.numCells;
  if (m == 0)
    return;
  hipLaunchKernelGGL(k_predict, dim3(blocks_for(m, 256)), dim3(256), 0, s, pos, vel, pred, key,
                     startEnd, n, c);
}

// A quad pass's grid: sixteen particles per 64-lane workgroup.
static dim3 quad_grid(uint32_t n)
{
  return dim3(blocks_for(n, kQuadGroups));
}

void launch_pbf_density_lambda(const float4* pred, const uint2* startEnd, float* density,
                               float* lambda, uint32_t n, const PbfConsts& c, const Launch& L,
                               hipStream_t s, int mode, const LearnArgs& la, const PassOrder& po)
{
  if (n == 0)
    return;
  if (L.tile)  // survivor lists are kept by the untiled kernels only
    mode &= ~(kNbRecord | kNbReplay);
  if (L.quad && !L.tile && (mode & ~kNbRecord) == 0) {
    hipLaunchKernelGGL(mode ? k_density_lambda_q<kNbRecord> : k_density_lambda_q<0>,
                       quad_grid(n), dim3(kQuadBlock), 0, s, pred, startEnd, density, lambda, n,
                       c, po);
    return;
  }
  const dim3 g(blocks_for(n, L.block)), b(L.block);
  const uint32_t q = (uint32_t)L.queue, lds = queue_bytes(L);
#define RTP_DL(M)                                                                              \
  case M:                                                                                    \
    if (L.tile)                                                                              \
      hipLaunchKernelGGL(k_density_lambda_t<M>, g, b, lds, s, pred, startEnd, density, lambda, \
                         n, c, L.xcdRemap, la, q, po);                                       \
    else                                                                                     \
      hipLaunchKernelGGL(k_density_lambda<M>, g, b, lds, s, pred, startEnd, density, lambda, \
                         n, c, L.xcdRemap, la, q, po);                                       \
    break;
  switch (mode) {
    RTP_DL(0)
    RTP_DL(kFeatures)
    RTP_DL(kAccum)
    RTP_DL(kFeatures | kAccum)
    RTP_DL(kDensOnly)
    RTP_DL(kNbRecord)
    RTP_DL(kFeatures | kNbRecord)
  default: break;
  }
#undef RTP_DL
}

void launch_pbf_correct(const float4* pred, const float* lambda, const uint2* startEnd,
                        float4* predOut, const float4* pos, float4* vel, float4* posOut,
                        uint32_t n, int last, const PbfConsts& c, const Launch& L, hipStream_t s,
                        int mode, const LearnArgs& la, const PassOrder& po)
{
  if (n == 0)
    return;
  if (L.tile)
    mode &= ~(kNbRecord | kNbReplay);
  if (L.quad && !L.tile && (mode & ~kNbReplay) == 0) {
    hipLaunchKernelGGL(mode ? k_correct_q<kNbReplay> : k_correct_q<0>, quad_grid(n),
                       dim3(kQuadBlock), 0, s, pred, lambda, startEnd, predOut, pos, vel, posOut,
                       n, last, c, po);
    return;
  }
  const dim3 g(blocks_for(n, L.block)), b(L.block);
  const uint32_t q = (uint32_t)L.queue, lds = queue_bytes(L);
#define RTP_CO(M)                                                                              \
  case M:                                                                                    \
    if (L.tile)                                                                              \
      hipLaunchKernelGGL(k_correct_t<M>, g, b, lds, s, pred, lambda, startEnd, predOut, pos,   \
                         vel, posOut, n, last, c, L.xcdRemap, la, q, po);                    \
    else                                                                                     \
      hipLaunchKernelGGL(k_correct<M>, g, b, lds, s, pred, lambda, startEnd, predOut, pos,     \
                         vel, posOut, n, last, c, L.xcdRemap, la, q, po);                    \
    break;
  switch (mode) {
    RTP_CO(0)
    RTP_CO(kNbReplay)
  default: break;
  }
#undef RTP_CO
}

// Threads of a pass over a compacted list: the list holds at most n entries, and a
// grid-stride loop covers any count with at most 2^16 threads (sets are a few percent of
// the particles; a larger grid only adds workgroups that find no entry).
static uint32_t list_blocks(uint32_t n, uint32_t block)
{
  const uint32_t cap = (1u << 16) / block;
  const uint32_t b = blocks_for(n, block);
  return b < cap ? b : cap;
}

void launch_fallback_set(const float4* pred, const uint2* startEnd, const float* resDensity,
                         float tau, uint32_t n, const PbfConsts& c, const FallbackSet& fs,
                         hipStream_t s)
{
  launch_fill_u32(fs.count, 3, 0u, s);
  if (n == 0)
    return;
  hipLaunchKernelGGL(k_fb_mark, dim3(blocks_for(n, 256)), dim3(256), 0, s, resDensity, n,
                     c.p.restDensity, tau, fs);
  // one item per (particle, neighbour cell); at most 2^18 threads loop over the items
  const uint32_t cap = (1u << 18) / 256u;
  const uint32_t b = blocks_for(n, 256) * 27u + 1u;
  const dim3 g(b < cap ? b : cap), cg(blocks_for(n, 1024));
  hipLaunchKernelGGL(k_fb_grow, g, dim3(256), 0, s, pred, startEnd, fs.seeds, fs.count, 1u,
                     resDensity, tau, n, c, fs.act);
  hipLaunchKernelGGL(k_fb_compact, cg, dim3(1024), 0, s, n, 1u, fs.act, fs.list, fs.count + 1);
  hipLaunchKernelGGL(k_fb_grow, g, dim3(256), 0, s, pred, startEnd, fs.list, fs.count + 1, 2u,
                     resDensity, tau, n, c, fs.act);
  hipLaunchKernelGGL(k_fb_compact, cg, dim3(1024), 0, s, n, 2u, fs.act, fs.halo, fs.count + 2);
}

void launch_fallback_iteration(float4* pred, float4* partner, float* density, float* lambda,
                               const uint2* startEnd, uint32_t n, uint32_t f, bool last,
                               const PbfConsts& c, const Launch& L, const FallbackSet& fs,
                               const NbList* nb, hipStream_t s)
{
  if (n == 0)
    return;
  if (f == 0)
    hipLaunchKernelGGL(k_fb_begin, dim3(blocks_for(n, 256)), dim3(256), 0, s, pred, partner,
                       lambda, n, c, fs);
  (void)L;
  // iteration f reads one buffer and writes the set's corrections into the other: pred for
  // odd f, the partner for even f; both hold the same positions outside the set
  const float4* src = (f & 1u) ? partner : pred;
  float4* dst = (f & 1u) ? pred : partner;
  // a quad per particle, at most 2^16 particles per round of the grid-stride loop
  const uint32_t qb = blocks_for(n, kQuadGroups), qcap = (1u << 16) / kQuadGroups;
  const dim3 g(qb < qcap ? qb : qcap), b(kQuadBlock);
  const NbList lists = nb ? *nb : NbList();
  const int lastMode = last ? kLastOnly : kNextIter;
  // the correction replays the density pass's lists: the same positions
  if (nb) {
    hipLaunchKernelGGL(k_density_lambda_quad<true>, g, b, 0, s, src, startEnd, density, lambda,
                       c, fs, lists);
    hipLaunchKernelGGL(k_correct_quad<true>, g, b, 0, s, src, lambda, startEnd, dst, lastMode, c,
                       fs, lists);
  } else {
    hipLaunchKernelGGL(k_density_lambda_quad<false>, g, b, 0, s, src, startEnd, density, lambda,
                       c, fs, lists);
    hipLaunchKernelGGL(k_correct_quad<false>, g, b, 0, s, src, lambda, startEnd, dst, lastMode,
                       c, fs, lists);
  }
  if (last && dst == partner)  // an odd number of iterations: the result is in the partner
    hipLaunchKernelGGL(k_fb_scatter, dim3(list_blocks(n, 256)), dim3(256), 0, s, pred, partner,
                       fs);
}

void launch_pbf_update_vel(const float4* pred, const float4* pos, float4* vel, float4* posOut,
                           uint32_t n, const PbfConsts& c, hipStream_t s)
{
  if (n == 0)
    return;
  hipLaunchKernelGGL(k_update_vel, dim3(blocks_for(n, 256)), dim3(256), 0, s, pred, pos, vel,
                     posOut, n, c);
}

// The untiled kernel for survivor-list mode nb (0, kNbRecord, kNbReplay).
#define RTP_BY_NB(KERNEL)                                                                      \
  [](int nb) {                                                                               \
    return nb == kNbRecord ? KERNEL<kNbRecord> : (nb == kNbReplay ? KERNEL<kNbReplay> : KERNEL<0>); \
  }

void launch_pbf_vorticity(const float4* pred, const uint2* startEnd, const float4* vel,
                          float4* vort, uint32_t n, const PbfConsts& c, const Launch& L,
                          hipStream_t s, const PassOrder& po, int nb)
{
  if (n == 0)
    return;
  if (L.quad && !L.tile) {
    hipLaunchKernelGGL(RTP_BY_NB(k_vorticity_q)(nb), quad_grid(n), dim3(kQuadBlock), 0, s, pred,
                       startEnd, vel, vort, n, c, po);
    return;
  }
  hipLaunchKernelGGL(L.tile ? k_vorticity_t : RTP_BY_NB(k_vorticity)(nb),
                     dim3(blocks_for(n, L.block)), dim3(L.block), queue_bytes(L), s, pred,
                     startEnd, vel, vort, n, c, L.xcdRemap, (uint32_t)L.queue, po);
}

void launch_pbf_confinement(const float4* pred, const uint2* startEnd, const float4* vort,
                            const float4* vel, float4* velOut, uint32_t n, const PbfConsts& c,
                            const Launch& L, hipStream_t s, const PassOrder& po, int nb)
{
  if (n == 0)
    return;
  if (L.quad && !L.tile) {
    hipLaunchKernelGGL(RTP_BY_NB(k_confinement_q)(nb), quad_grid(n), dim3(kQuadBlock), 0, s,
                       pred, startEnd, vort, vel, velOut, n, c, po);
    return;
  }
  hipLaunchKernelGGL(L.tile ? k_confinement_t : RTP_BY_NB(k_confinement)(nb),
                     dim3(blocks_for(n, L.block)), dim3(L.block), queue_bytes(L), s, pred,
                     startEnd, vort, vel, velOut, n, c, L.xcdRemap, (uint32_t)L.queue, po);
}

void launch_pbf_xsph(const float4* pred, const uint2* startEnd, const float4* velIn,
                     float4* velOut, float4* posOut, uint32_t n, const PbfConsts& c,
                     const Launch& L, hipStream_t s, const PassOrder& po, int nb)
{
  if (n == 0)
    return;
  if (L.quad && !L.tile) {
    hipLaunchKernelGGL(RTP_BY_NB(k_xsph_q)(nb), quad_grid(n), dim3(kQuadBlock), 0, s, pred,
                       startEnd, velIn, velOut, posOut, n, c, po);
    return;
  }
  hipLaunchKernelGGL(L.tile ? k_xsph_t : RTP_BY_NB(k_xsph)(nb), dim3(blocks_for(n, L.block)),
                     dim3(L.block), queue_bytes(L), s, pred, startEnd, velIn, velOut, posOut, n,
                     c, L.xcdRemap, (uint32_t)L.queue, po);
}
#undef RTP_BY_NB

void launch_density_error(const float* density, uint32_t n, float rho0, float* metric,
                          hipStream_t s)
{
  if (n == 0) {
    hipLaunchKernelGGL(k_zero3, dim3(1), dim3(64), 0, s, metric);
    return;
  }
  uint32_t blocks = blocks_for(n, 256);
  if (blocks > kErrBlocks)
    blocks = kErrBlocks;
  float* partial = metric + 4;
  hipLaunchKernelGGL(k_density_error, dim3(blocks), dim3(256), 0, s, density, n, rho0, partial);
  hipLaunchKernelGGL(k_density_error_final, dim3(1), dim3(256), 0, s, partial, blocks, metric);
}


// ------------------------------------------------------------------- batched launchers

void launch_fill_start_end(uint2* startEnd, uint32_t count, hipStream_t s)
{
  if (count)
    hipLaunchKernelGGL(k_fill_start_end, dim3(blocks_for(count, 256)), dim3(256), 0, s, startEnd, count);
}

#define RTP_BATCH_STREAM(kernel, ...)                                                          \
  do {                                                                                       \
    if (bv.numBlocks)                                                                        \
      hipLaunchKernelGGL(kernel, dim3(bv.numBlocks), dim3(bv.block), 0, s, bv, __VA_ARGS__); \
  } while (0)

#define RTP_BATCH_LAUNCH(kernel, ...)                                                          \
  do {                                                                                       \
    if (bv.numBlocks)                                                                        \
      hipLaunchKernelGGL(kernel, dim3(bv.numBlocks), dim3(bv.block),                         \
                         bv.qdepth * bv.block * 4u, s, bv, __VA_ARGS__);                     \
  } while (0)

void launch_pbf_predict_batch(const BatchView& bv, const float4* pos, const float4* vel,
                              float4* pred, uint32_t* key, hipStream_t s)
{
  RTP_BATCH_STREAM(k_predict_batch, pos, vel, pred, key);
}

// The batched kernel instantiation for survivor-list mode nb.
#define RTP_BATCH_NB(KERNEL) \
  (nb == kNbRecord ? KERNEL<kNbRecord> : (nb == kNbReplay ? KERNEL<kNbReplay> : KERNEL<0>))

void launch_pbf_density_lambda_batch(const BatchView& bv, const float4* pred, const uint2* startEnd,
                                     float* density, float* lambda, hipStream_t s, int nb)
{
  RTP_BATCH_LAUNCH(RTP_BATCH_NB(k_density_lambda_batch), pred, startEnd, density, lambda);
}

void launch_pbf_correct_batch(const BatchView& bv, const float4* pred, const float* lambda,
                              const uint2* startEnd, float4* predOut, const float4* pos,
                              float4* vel, float4* posOut, int last, hipStream_t s, int nb)
{
  RTP_BATCH_LAUNCH(RTP_BATCH_NB(k_correct_batch), pred, lambda, startEnd, predOut, pos, vel, posOut,
                   last);
}

void launch_pbf_update_vel_batch(const BatchView& bv, const float4* pred, const float4* pos,
                                 float4* vel, float4* posOut, hipStream_t s)
{
  RTP_BATCH_STREAM(k_update_vel_batch, pred, pos, vel, posOut);
}

void launch_pbf_vorticity_batch(const BatchView& bv, const float4* pred, const uint2* startEnd,
                                const float4* vel, float4* vort, hipStream_t s, int nb)
{
  RTP_BATCH_LAUNCH(RTP_BATCH_NB(k_vorticity_batch), pred, startEnd, vel, vort);
}

void launch_pbf_confinement_batch(const BatchView& bv, const float4* pred, const uint2* startEnd,
                                  const float4* vort, const float4* vel, float4* velOut,
                                  hipStream_t s, int nb)
{
  RTP_BATCH_LAUNCH(RTP_BATCH_NB(k_confinement_batch), pred, startEnd, vort, vel, velOut);
}

void launch_pbf_xsph_batch(const BatchView& bv, const float4* pred, const uint2* startEnd,
                           const float4* velIn, float4* velOut, float4* posOut, hipStream_t s, int nb)
{
  RTP_BATCH_LAUNCH(RTP_BATCH_NB(k_xsph_batch), pred, startEnd, velIn, velOut, posOut);
}

#undef RTP_BATCH_NB
#undef RTP_BATCH_LAUNCH
#undef RTP_BATCH_STREAM

}  // namespace rtp
