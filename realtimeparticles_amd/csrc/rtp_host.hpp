// Host-side helpers shared by the model handles of the C ABI: constant folding as the
// reference's host does it, grid set-up, device buffers and per-launch HIP-event timing.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "rtp_internal.hpp"

#define RTP_REQUIRE(cond, msg)                    \
  do {                                            \
    if (!(cond))                                  \
      return ::rtp::fail(RTP_ERR_INVALID, msg);   \
  } while (0)

namespace rtp {
namespace {

// A roctx range around host work that enqueues device work (visible with
// rocprofv3 --marker-trace): update calls, graph launches and direct substeps.
struct MarkRange {
  explicit MarkRange(const char* name) { roctxRangePush(name); }
  ~MarkRange() { roctxRangePop(); }
  MarkRange(const MarkRange&) = delete;
  MarkRange& operator=(const MarkRange&) = delete;
};

// ------------------------------------------------------------------------- host constants

// utils/Utils.cpp:24-29 FloatToStr(v, 10) followed by the OpenCL compiler's parse.
float float_roundtrip(float v)
{
  char buf[128];
  std::snprintf(buf, sizeof(buf), "%.10f", (double)v);
  return std::strtof(buf, nullptr);
}

constexpr float kPiF = 3.1415927f;  // utils/Math.hpp:69

int make_grid(const rtp_model_desc* d, uint32_t defaultCap, GridConsts* g)
{
  for (int a = 0; a < 3; ++a) {
    if (!(d->boxSize[a] > 0.0f) || d->gridRes[a] <= 0)
      return fail(RTP_ERR_INVALID, "boxSize and gridRes must be positive");
    g->absWall[a] = float_roundtrip(d->boxSize[a] / 2.0f);  // Fluids.cpp:108-110
    g->res[a] = d->gridRes[a];
  }
  const uint64_t nc = (uint64_t)d->gridRes[0] * d->gridRes[1] * d->gridRes[2];
  if (nc == 0 || nc > (1u << 30))
    return fail(RTP_ERR_INVALID, "grid has too many cells");
  g->cellSize = float_roundtrip(d->boxSize[0] / (float)d->gridRes[0]);  // Fluids.cpp:114
  g->numCells = (uint32_t)nc;
  g->cap = d->maxPartsInCell ? d->maxPartsInCell : defaultCap;
  return RTP_OK;
}

// Largest cell key fillCellIDs can produce (each 3-D index is at most res when a position
// sits exactly on the +wall) — the sort only needs that many bits.
uint32_t max_active_key(const GridConsts& g)
{
  const uint64_t ry = (uint64_t)g.res[1], rz = (uint64_t)g.res[2];
  const uint64_t m = (uint64_t)g.res[0] * rz * ry + ry * rz + rz;
  return (uint32_t)std::min<uint64_t>(m, 0xffffffffu);
}

// Entries of the per-cell count table of the thread order: every key the hash can produce,
// rounded up to the scan's 16-word groups.
uint32_t order_cells(const GridConsts& g)
{
  return (max_active_key(g) + 1u + 15u) & ~15u;
}

// Count-table entries for either key kind.
uint32_t order_count_words(const GridConsts& g);

// The same for the brick keys of RTP_LAUNCH_BRICKS (brick_key: cell coordinates up to res
// in each dimension, (y, z) tiles of kBrick x kBrick).
uint32_t brick_order_cells(const GridConsts& g)
{
  const uint64_t nTy = ((uint64_t)g.res[1] + kBrick) / kBrick;
  const uint64_t nTz = ((uint64_t)g.res[2] + kBrick) / kBrick;
  const uint64_t m = nTy * nTz * ((uint64_t)g.res[0] + 1u) * kBrick * kBrick;
  return (uint32_t)std::min<uint64_t>((m + 15u) & ~(uint64_t)15u, 0xfffffff0u);
}

uint32_t order_count_words(const GridConsts& g)
{
  return std::max(order_cells(g), brick_order_cells(g));
}

// Smallest float >= v, largest float <= v.
float float_above(double v)
{
  float t = (float)v;
  if ((double)t < v)
    t = std::nextafter(t, INFINITY);
  return t;
}

float float_below(double v)
{
  float t = (float)v;
  if ((double)t > v)
    t = std::nextafter(t, -INFINITY);
  return t;
}

// Smallest float that is >= h*h as a real number.
float cull_bound(float h)
{
  return float_above((double)h * (double)h);
}

// --------------------------------------------------------------------------------- buffers

template <class T>
int dev_alloc(T** p, size_t count)
{
  *p = nullptr;
  if (count == 0)
    count = 1;
  RTP_HIP_TRY(hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * count));
  return RTP_OK;
}

template <class T>
void dev_free(T*& p)
{
  if (p)
    (void)hipFree(p);
  p = nullptr;
}

// ------------------------------------------------------------------------------- profiling

struct Profiler {
  bool enabled = false;
  struct Pending {
    int id;
    hipEvent_t a, b;
  };
  std::vector<std::string> names;
  std::vector<uint32_t> launches;
  std::vector<double> totals;
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;

  int id_of(const char* name)
  {
    for (size_t i = 0; i < names.size(); ++i)
      if (names[i] == name)
        return (int)i;
    names.emplace_back(name);
    launches.push_back(0);
    totals.push_back(0.0);
    return (int)names.size() - 1;
  }
  hipEvent_t take()
  {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  // Wrap one launch.  Call begin() before and end() after enqueueing it.
  int beginId = -1;
  hipEvent_t beginEv = nullptr;
  void begin(const char* name, hipStream_t s)
  {
    if (!enabled)
      return;
    beginId = id_of(name);
    beginEv = take();
    (void)hipEventRecord(beginEv, s);
  }
  void end(hipStream_t s)
  {
    if (!enabled || beginId < 0)
      return;
    hipEvent_t e = take();
    (void)hipEventRecord(e, s);
    pending.push_back(Pending{ beginId, beginEv, e });
    beginId = -1;
  }
  int collect()
  {
    for (auto& p : pending) {
      RTP_HIP_TRY(hipEventSynchronize(p.b));
      float ms = 0.0f;
      RTP_HIP_TRY(hipEventElapsedTime(&ms, p.a, p.b));
      launches[p.id] += 1;
      totals[p.id] += ms;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
    return RTP_OK;
  }
  void reset()
  {
    (void)collect();
    for (auto& l : launches)
      l = 0;
    for (auto& t : totals)
      t = 0.0;
  }
  int report(rtp_kernel_time* out, int capacity, int* count)
  {
    const int rc = collect();
    if (rc)
      return rc;
    const int n = (int)names.size();
    if (count)
      *count = n;
    for (int i = 0; i < n && i < capacity && out; ++i) {
      std::memset(out[i].name, 0, sizeof(out[i].name));
      std::strncpy(out[i].name, names[i].c_str(), sizeof(out[i].name) - 1);
      out[i].launches = launches[i];
      out[i].totalMs = totals[i];
    }
    return RTP_OK;
  }
  void destroy()
  {
    (void)collect();
    for (auto e : pool)
      (void)hipEventDestroy(e);
    pool.clear();
  }
};

#define RTP_PROF(prof, name, stream, launch) \
  do {                                       \
    (prof).begin(name, stream);              \
    launch;                                  \
    (prof).end(stream);                      \
  } while (0)

int check_device(int device)
{
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    return fail(RTP_ERR_NO_DEVICE, "no HIP device visible");
  if (device < 0 || device >= count)
    return fail(RTP_ERR_INVALID, "device ordinal out of range");
  RTP_HIP_TRY(hipSetDevice(device));
  return RTP_OK;
}

int upload_f4(float4* dst, const float* src, uint32_t count, uint32_t maxN, float4 fill,
              hipStream_t s)
{
  std::vector<float4> h(maxN, fill);
  for (uint32_t i = 0; i < count; ++i)
    h[i] = make_float4(src[4 * i + 0], src[4 * i + 1], src[4 * i + 2], 0.0f);
  RTP_HIP_TRY(hipMemcpyAsync(dst, h.data(), sizeof(float4) * maxN, hipMemcpyHostToDevice, s));
  RTP_HIP_TRY(hipStreamSynchronize(s));
  return RTP_OK;
}

// ----------------------------------------------------------------------------- tuning

// Launch flags a single model / a batch takes.
constexpr int kFluidsLaunchFlags = RTP_LAUNCH_XCD_MAP | RTP_LAUNCH_LDS_TILES | RTP_LAUNCH_RESORT |
                                   RTP_LAUNCH_NB_LIST | RTP_LAUNCH_BRICKS | RTP_LAUNCH_QUAD;
constexpr int kBatchLaunchFlags = RTP_LAUNCH_RESORT | RTP_LAUNCH_NB_LIST;

// Validates a launch tuning; allowedFlags masks the RTP_LAUNCH_* bits the caller supports.
int check_tuning(const rtp_tuning& t, int allowedFlags)
{
  RTP_REQUIRE(t.blockSize >= 64 && t.blockSize <= 256 && t.blockSize % 64 == 0,
              "blockSize must be 64, 128, 192 or 256");
  RTP_REQUIRE((t.flags & ~allowedFlags) == 0, "launch flags not supported here");
  RTP_REQUIRE(!(t.flags & RTP_LAUNCH_BRICKS) || (t.flags & RTP_LAUNCH_RESORT),
              "RTP_LAUNCH_BRICKS orders the threads of RTP_LAUNCH_RESORT: set both");
  RTP_REQUIRE(!(t.flags & RTP_LAUNCH_QUAD) || !(t.flags & RTP_LAUNCH_LDS_TILES),
              "RTP_LAUNCH_QUAD and RTP_LAUNCH_LDS_TILES exclude each other");
  RTP_REQUIRE(t.sortBits == 0 || (t.sortBits >= 4 && t.sortBits <= 11), "sortBits must be 0 or 4..11");
  const int q = t.queueDepth ? t.queueDepth : kDefaultQueue;
  RTP_REQUIRE(q >= 8 && q <= 64, "queueDepth must be 0 or 8..64");
  const int limit = (t.flags & RTP_LAUNCH_LDS_TILES) ? 24 * 1024 : 64 * 1024;
  RTP_REQUIRE(q * t.blockSize * 4 <= limit, "queueDepth x blockSize exceeds the LDS budget");
  return RTP_OK;
}

// Times fixed-input substeps for candidate tunings and returns the fastest in *best.
// set(t) applies a tuning, steps(k) runs k fixed-input substeps.  The sort's digit width
// only affects the sort, so it is chosen first; then every valid combination of block size,
// launch flags (single models) and queue depth is timed with it.  The quad passes of a
// single model ignore the block size and queue depth, so their flag combinations are timed
// last, once each, with the best lane-per-particle block size and depth.
template <class Set, class Steps>
int tune_search(hipStream_t s, uint32_t trialSteps, const rtp_tuning& start, bool singleModel,
                Set set, Steps steps, rtp_tuning* best)
{
  hipEvent_t e0 = nullptr, e1 = nullptr;
  RTP_HIP_TRY(hipEventCreate(&e0));
  RTP_HIP_TRY(hipEventCreate(&e1));
  int rc = RTP_OK;
  auto time_one = [&](const rtp_tuning& t, float* ms) -> int {
    int r = set(t);
    if (r || (r = steps(1)))
      return r;
    (void)hipEventRecord(e0, s);
    if ((r = steps(trialSteps)))
      return r;
    (void)hipEventRecord(e1, s);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(ms, e0, e1) != hipSuccess)
      return fail(RTP_ERR_HIP, "autotune timing failed");
    *ms /= (float)trialSteps;
    return RTP_OK;
  };
  rtp_tuning top = start;
  top.msPerStep = 1e30f;
  for (int bits : { 7, 8, 9, 10, 11 }) {
    rtp_tuning t = start;
    t.sortBits = bits;
    float ms = 0.0f;
    if (check_tuning(t, singleModel ? kFluidsLaunchFlags : kBatchLaunchFlags))
      continue;
    if ((rc = time_one(t, &ms)))
      break;
    if (ms < top.msPerStep) {
      top = t;
      top.msPerStep = ms;
    }
  }
  auto consider = [&](const rtp_tuning& t) -> int {
    if (check_tuning(t, singleModel ? kFluidsLaunchFlags : kBatchLaunchFlags))
      return RTP_OK;
    float ms = 0.0f;
    const int r = time_one(t, &ms);
    if (!r && ms < top.msPerStep) {
      top = t;
      top.msPerStep = ms;
    }
    return r;
  };
  // every combination of the flags the caller takes
  const int laneFlags = singleModel ? kFluidsLaunchFlags & ~RTP_LAUNCH_QUAD : 3;
  for (int b = 64; b <= 256 && !rc; b += 64)
    for (int fi = 0; fi <= laneFlags && !rc; ++fi)
      for (int q : { 8, 12, 16, 24, 32 }) {
        const int f = singleModel ? fi : ((fi & 1) ? RTP_LAUNCH_RESORT : 0) | ((fi & 2) ? RTP_LAUNCH_NB_LIST : 0);
        rtp_tuning t = top;
        t.blockSize = b;
        t.flags = f;
        t.queueDepth = q;
        if ((rc = consider(t)))
          break;
      }
  if (singleModel) {
    const rtp_tuning lane = top;
    for (int fi = 0; fi <= laneFlags && !rc; ++fi) {
      if (fi & (RTP_LAUNCH_XCD_MAP | RTP_LAUNCH_LDS_TILES))  // no effect on / not with quads
        continue;
      rtp_tuning t = lane;
      t.flags = fi | RTP_LAUNCH_QUAD;
      rc = consider(t);
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc)
    return rc;
  fail(RTP_OK, "");  // candidates rejected by check_tuning leave no error behind
  *best = top;
  return RTP_OK;
}

// Fluids::createProgram constants (Fluids.cpp:97-126) of a scene with box edge boxX split
// into resX cells; the rest of PbfConsts follows from the parameters (refresh_pbf_consts).
void init_pbf_consts(PbfConsts& c, const GridConsts& g, float boxX, int32_t resX, float gravity)
{
  c.g = g;
  const float effectRadius = boxX / (float)resX;                                      // :104
  c.h = float_roundtrip(effectRadius);
  c.poly6 = float_roundtrip(315.0f / (64.0f * kPiF * std::pow(effectRadius, 9.f)));  // :117
  c.spiky = float_roundtrip(15.0f / (kPiF * std::pow(effectRadius, 6.f)));           // :118
  c.maxVel = float_roundtrip(30.0f);                                                  // :119
  c.gravity = gravity > 0.0f ? gravity : 9.81f;                                       // define.cl:8
  c.p = rtp_fluid_params{ 450.0f, 600.0f, 0.010f, 3u, 1u, 0.006f, 0.001f, 4u, 1u, 0.0004f, 0.0001f };
}

// The host-side constant folding of Fluids::createProgram (Fluids.cpp:97-126) and of the
// kernels' constant subexpressions.
void refresh_pbf_consts(PbfConsts& c)
{
  const float h = c.h;
  c.hh = h * h;
  c.cullSq = cull_bound(h);
  const rtp_fluid_params& p = c.p;
  const float dq = p.artPressureRadius * h;
  const float t = dq < h ? 1.0f : 0.0f;  // 1 - step(h, dq)
  const float d = h * h - dq * dq;
  c.wdq = (t * c.poly6) * ((d * d) * d);
  c.rho0sq = p.restDensity * p.restDensity;
  c.dtEps = p.timeStep + 0.00000001f;
  c.gdt = -c.gravity * p.timeStep;
  for (int a = 0; a < 3; ++a) {
    c.lo[a] = -c.g.absWall[a] + 0.01f;
    c.hi[a] = c.g.absWall[a] - 0.1f;
  }
  c.cullFused = float_above((double)c.cullSq * (1.0 + 0x1p-20));
  c.farFused = float_below((double)kFarD2 * (1.0 - 0x1p-20));
  // The short forms' operand ranges hold for every neighbour with r < h when
  // 2^-7 <= h <= 2^4: r in (1e-8, h), gradient numerators |n| <= 45/(pi h^3) <= 2^25
  // (tiny nonzero ones are checked per neighbour), so quotients in [2^-94, 2^52]; poly6
  // values in [1.5 h^-3 2^-72, 1.6 h^-3], within [2^-84, 2^22], against a divisor wdq in
  // [2^-30, 2^30], so quotients in [2^-114, 2^52].
  c.shortMath = 0;
  if (h >= 0x1p-7f && h <= 0x1p4f) {
    c.shortMath |= kShortSqrtGrad;
    if (c.wdq >= 0x1p-30f && c.wdq <= 0x1p30f)
      c.shortMath |= kShortArtDiv;
  }
}

// ------------------------------------------------------------------------- shared grid state

// Buffers of the uniform grid that both models own: the reference's p_cellID,
// RadixSortIndices and c_startEndPartID (plus the sort's ping-pong partners and counts).
struct GridState {
  GridConsts g{};
  uint32_t maxN = 0;
  uint32_t* key[2] = { nullptr, nullptr };
  uint32_t* idx[2] = { nullptr, nullptr };
  uint2* startEnd = nullptr;
  uint32_t* counts = nullptr;
  int keyCur = 0;   // key[keyCur] is p_cellID
  int idxCur = 0;   // idx[idxCur] holds the last permutation over [0, sortedN)
  uint32_t sortedN = 0;
  uint32_t dirtyEnd = 0;  // keys in [n, dirtyEnd) are not the reset tail 2*NC+i
  uint32_t permN = 0;     // range of the permutation in idx[idxCur] (cell or camera sort)
  int sortBits = 0;

  int alloc(uint32_t maxN_)
  {
    maxN = maxN_;
    int rc = 0;
    if ((rc = dev_alloc(&key[0], maxN)) || (rc = dev_alloc(&key[1], maxN)) ||
        (rc = dev_alloc(&idx[0], maxN)) || (rc = dev_alloc(&idx[1], maxN)) ||
        (rc = dev_alloc(&startEnd, g.numCells)) ||
        (rc = dev_alloc(&counts, count_capacity(maxN))))
      return rc;
    return RTP_OK;
  }
  void release()
  {
    dev_free(key[0]);
    dev_free(key[1]);
    dev_free(idx[0]);
    dev_free(idx[1]);
    dev_free(startEnd);
    dev_free(counts);
  }
  // resetCellIDs (grid.cl:65-71) into both key buffers so their tails always agree.
  void reset_keys(hipStream_t s)
  {
    launch_reset_cell_ids(key[0], maxN, g.numCells, s);
    launch_reset_cell_ids(key[1], maxN, g.numCells, s);
    keyCur = 0;
    idxCur = 0;
    sortedN = 0;
    dirtyEnd = 0;
    permN = 0;
  }
  uint32_t sort_range(uint32_t n) const { return std::max(n, dirtyEnd); }
  uint32_t max_key(uint32_t n) const
  {
    uint32_t m = max_active_key(g);
    const uint32_t sn = sort_range(n);
    if (sn > n) {  // stale keys may be anything fillCellIDs wrote or a reset tail value
      const uint64_t tail = (uint64_t)g.numCells * 2 + sn - 1;
      m = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(m, tail), 0xffffffffu);
    }
    return m;
  }
  // RadixSort::sort on p_cellID over [0, sort_range(n)); the rest is already in order.
  int sort(uint32_t n, hipStream_t s, Profiler& prof)
  {
    const uint32_t sn = sort_range(n);
    const SortPlan plan = make_sort_plan(sn, max_key(n), sortBits);
    bool inB = false;
    int rc = 0;
    RTP_PROF(prof, "radix_sort", s,
             rc = sort_pairs(plan, key[keyCur], key[1 - keyCur], idx[0], idx[1], counts, &inB, s));
    if (rc)
      return rc;
    if (inB)
      keyCur = 1 - keyCur;
    idxCur = inB ? 1 : 0;
    sortedN = sn;
    permN = sn;
    return RTP_OK;
  }
  // Copy p_cellID, the permutation (identity past the sorted range) and the cell table.
  int read(uint32_t* keys, uint32_t* perm, uint32_t* se, hipStream_t s)
  {
    if (keys)
      RTP_HIP_TRY(hipMemcpyAsync(keys, key[keyCur], sizeof(uint32_t) * maxN, hipMemcpyDeviceToHost, s));
    if (perm && permN)
      RTP_HIP_TRY(hipMemcpyAsync(perm, idx[idxCur], sizeof(uint32_t) * permN, hipMemcpyDeviceToHost, s));
    if (se)
      RTP_HIP_TRY(hipMemcpyAsync(se, startEnd, sizeof(uint2) * g.numCells, hipMemcpyDeviceToHost, s));
    RTP_HIP_TRY(hipStreamSynchronize(s));
    if (perm)
      for (uint32_t i = permN; i < maxN; ++i)
        perm[i] = i;
    return RTP_OK;
  }
};

}  // namespace
}  // namespace rtp
