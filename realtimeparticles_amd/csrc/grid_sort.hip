// Uniform-grid kernels for gfx950: cell hashing, a stable LSD radix sort of
// (cellID, index) pairs and the cell start/end table.
//
// Reference behaviour reproduced here (bit-exact):
//   fillCellIDs / resetCellIDs      physics/ocl/kernels/grid.cl:65-86
//   RadixSort::sort                 physics/utils/RadixSort.cpp:122-189, radixSort.cl:19-188
//   reset/fillStart/fillEnd/adjust  physics/ocl/kernels/grid.cl:91-152
//
// The sort is a different algorithm from the reference's (512 work-items scanning
// contiguous chunks serially): every pass is a reduce-then-scan over wave-owned subtiles.
//   upsweep   : each wave histograms the digits of its subtile in LDS
//   scan      : one exclusive scan over the digit-major count table
//   downsweep : each wave ranks its keys with ballot-based match (one ballot per digit bit,
//               mbcnt for the rank among lower lanes), adds its subtile's digit base and
//               scatters; a wave-private LDS counter per digit carries the running rank
//               from one 64-key round to the next, so no workgroup barrier sits in the loop.
// A stable sort has exactly one result, so it equals the reference's for any digit width.
#include <algorithm>

#include "rtp_device.hpp"
#include "rtp_grid_device.hpp"

namespace rtp {

namespace {

__device__ __forceinline__ uint32_t mbcnt64(uint64_t m)
{
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// fillCellIDs over n particles plus resetStartEndCell over the grid.
__global__ __launch_bounds__(256) void k_fill_keys_reset(const float4* __restrict__ pos, uint32_t n,
                                                         uint32_t* __restrict__ key,
                                                         uint2* __restrict__ startEnd, GridConsts g)
{
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n)
    key[t] = cell1d(pos[t], g);
  if (t < g.numCells)
    startEnd[t] = make_uint2(1u, 0u);
}

__global__ __launch_bounds__(256) void k_reset_cell_ids(uint32_t* __restrict__ key, uint32_t maxN,
                                                        uint32_t numCells)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < maxN)
    key[i] = numCells * 2u + i;
}

__global__ __launch_bounds__(256) void k_fill_identity(uint32_t* __restrict__ idx, uint32_t begin,
                                                       uint32_t end)
{
  const uint32_t i = begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < end)
    idx[i] = i;
}

__global__ __launch_bounds__(256) void k_copy_f4(float4* __restrict__ dst, const float4* __restrict__ src,
                                                 uint32_t begin, uint32_t end)
{
  const uint32_t i = begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < end)
    dst[i] = src[i];
}

template <int BITS, int ITEMS>
__global__ __launch_bounds__(256) void k_sort_upsweep(const uint32_t* __restrict__ keys, uint32_t n,
                                                      int shift, uint32_t S,
                                                      uint32_t* __restrict__ counts)
{
  constexpr int RADIX = 1 << BITS;
  __shared__ uint32_t hist[4][RADIX];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t s = blockIdx.x * 4u + (uint32_t)w;
  for (int d = lane; d < RADIX; d += 64)
    hist[w][d] = 0;
  __syncthreads();
  if (s < S) {
    const uint32_t base = s * (uint32_t)(ITEMS * 64);
#pragma unroll 8
    for (int k = 0; k < ITEMS; ++k) {
      const uint32_t i = base + (uint32_t)(k * 64 + lane);
      if (i < n)
        atomicAdd(&hist[w][(keys[i] >> shift) & (RADIX - 1)], 1u);
    }
  }
  __syncthreads();
  if (s < S)
    for (int d = lane; d < RADIX; d += 64)
      counts[(size_t)d * S + s] = hist[w][d];
}

// Exclusive scan of the digit-major count table (m words, m % 4 == 0) in two launches:
// per-segment sums, then a rescan of every segment from its base (the sum of the segment
// sums before it).
// A thread owns 16 consecutive words (four 16-byte loads), a block owns kScanSeg words.
constexpr uint32_t kScanPer = 16;
static_assert(kScanSeg == 256 * kScanPer, "a scan block owns 256 threads x kScanPer words");

__device__ __forceinline__ void load16(const uint32_t* __restrict__ data, uint32_t b, uint32_t m,
                                       uint32_t v[kScanPer])
{
#pragma unroll
  for (uint32_t k = 0; k < kScanPer; k += 4) {
    uint4 q = make_uint4(0u, 0u, 0u, 0u);
    if (b + k < m)
      q = *reinterpret_cast<const uint4*>(data + b + k);
    v[k] = q.x;
    v[k + 1] = q.y;
    v[k + 2] = q.z;
    v[k + 3] = q.w;
  }
}

// Block-wide exclusive scan of one value per thread (256 threads); returns the prefix and
// writes the block total to *total.
__device__ __forceinline__ uint32_t block_exclusive_256(uint32_t v, uint32_t* total)
{
  __shared__ uint32_t waveTot[4];
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(inc, d, 64);
    if ((int)lane >= d)
      inc += o;
  }
  if (lane == 63)
    waveTot[w] = inc;
  __syncthreads();
  uint32_t base = 0;
  for (uint32_t k = 0; k < w; ++k)
    base += waveTot[k];
  *total = waveTot[0] + waveTot[1] + waveTot[2] + waveTot[3];
  return base + inc - v;
}

__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* __restrict__ data, uint32_t m,
                                                     uint32_t* __restrict__ partial)
{
  const uint32_t b = blockIdx.x * kScanSeg + threadIdx.x * kScanPer;
  uint32_t v[kScanPer];
  load16(data, b, m, v);
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanPer; ++k)
    sum += v[k];
  uint32_t total = 0;
  (void)block_exclusive_256(sum, &total);
  if (threadIdx.x == 0)
    partial[blockIdx.x] = total;
}

// Each block's base is the sum of the segment sums before it, summed here: the segments
// are few (m / kScanSeg), so this costs less than a launch of its own.  Block 0 reads none,
// so a table of one segment needs no k_scan_reduce.
__global__ __launch_bounds__(256) void k_scan_apply(uint32_t* __restrict__ data, uint32_t m,
                                                    const uint32_t* __restrict__ partial)
{
  uint32_t before = 0;
  for (uint32_t j = threadIdx.x; j < blockIdx.x; j += 256u)
    before += partial[j];
  uint32_t segBase = 0;
  (void)block_exclusive_256(before, &segBase);
  __syncthreads();  // block_exclusive_256's LDS is used again below
  const uint32_t b = blockIdx.x * kScanSeg + threadIdx.x * kScanPer;
  uint32_t v[kScanPer];
  load16(data, b, m, v);
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanPer; ++k)
    sum += v[k];
  uint32_t total = 0;
  uint32_t run = segBase + block_exclusive_256(sum, &total);
#pragma unroll
  for (uint32_t k = 0; k < kScanPer; k += 4) {
    uint4 q;
    q.x = run;
    run += v[k];
    q.y = run;
    run += v[k + 1];
    q.z = run;
    run += v[k + 2];
    q.w = run;
    run += v[k + 3];
    if (b + k < m)
      *reinterpret_cast<uint4*>(data + b + k) = q;
  }
}

template <int BITS, int ITEMS>
__global__ __launch_bounds__(256) void k_sort_downsweep(
  const uint32_t* __restrict__ keysIn, const uint32_t* __restrict__ idxIn, uint32_t n, int shift,
  uint32_t S, const uint32_t* __restrict__ offsets, uint32_t* __restrict__ keysOut,
  uint32_t* __restrict__ idxOut)
{
  constexpr int RADIX = 1 << BITS;
  __shared__ uint32_t run[4][RADIX];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t s = blockIdx.x * 4u + (uint32_t)w;
  if (s < S)
    for (int d = lane; d < RADIX; d += 64)
      run[w][d] = offsets[(size_t)d * S + s];
  __syncthreads();
  if (s >= S)
    return;
  const uint32_t base = s * (uint32_t)(ITEMS * 64);
  uint32_t kr[ITEMS], vr[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const uint32_t i = base + (uint32_t)(k * 64 + lane);
    const bool valid = i < n;
    kr[k] = valid ? keysIn[i] : 0u;
    vr[k] = valid ? (idxIn ? idxIn[i] : i) : 0u;
  }
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const uint32_t i = base + (uint32_t)(k * 64 + lane);
    const bool valid = i < n;
    const uint32_t d = (kr[k] >> shift) & (uint32_t)(RADIX - 1);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < BITS; ++bit) {
      const bool set = (d >> bit) & 1u;
      const uint64_t m = __ballot(set);
      peers &= set ? m : ~m;
    }
    const uint32_t rank = mbcnt64(peers);
    const uint32_t start = run[w][d];
    if (valid && rank == 0)
      run[w][d] = start + (uint32_t)__popcll(peers);
    if (valid) {
      keysOut[start + rank] = kr[k];
      idxOut[start + rank] = vr[k];
    }
  }
}

__global__ __launch_bounds__(256) void k_reorder(ReorderArgs a)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.sortN)
    reorder_body(i, a, 0u, 0u);
}

// One launch over a batch: every scene's slice is a complete, fully active model
// (sortN = n = maxN = the scene's count) with its own grid, bounds and clamp.
__global__ __launch_bounds__(256) void k_reorder_batch(ReorderArgs a, BatchView bv)
{
  const uint2 blk = bv.blocks[blockIdx.x];
  const SceneSlot sc = bv.scenes[blk.x];
  const uint32_t i = blk.y + threadIdx.x;
  if (i >= sc.n)
    return;
  const PbfConsts& c = bv.consts[blk.x];
  const uint32_t o = sc.pbase;
  ReorderArgs l = a;
  l.key = a.key + o;
  l.idx = a.idx + o;
  l.sortN = l.n = l.maxN = sc.n;
  l.srcA = a.srcA ? a.srcA + o : nullptr;  // gather sources are indexed scene-locally
  l.dstA = a.dstA ? a.dstA + o : nullptr;
  l.srcB = a.srcB ? a.srcB + o : nullptr;
  l.dstB = a.dstB ? a.dstB + o : nullptr;
  l.srcC = a.srcC ? a.srcC + o : nullptr;
  l.dstC = a.dstC ? a.dstC + o : nullptr;
  for (int d = 0; d < 3; ++d) {
    l.lo[d] = c.lo[d];
    l.hi[d] = c.hi[d];
  }
  l.startEnd = a.startEnd + sc.kbase;
  l.g = c.g;
  reorder_body(i, l, sc.kbase, o);
}

template <int CH>
__global__ __launch_bounds__(512) void k_sort_small(uint32_t* __restrict__ keyA,
                                                    uint32_t* __restrict__ keyB,
                                                    uint32_t* __restrict__ idxA,
                                                    uint32_t* __restrict__ idxB, uint32_t n,
                                                    int passes)
{
  __shared__ SmallSortLds lds;
  small_sort_block<CH>(keyA, keyB, idxA, idxB, n, passes, lds);
}

template <int BITS, int ITEMS>
int run_pass(const SortPlan& p, int shift, const uint32_t* kin, const uint32_t* iin,
             uint32_t* kout, uint32_t* iout, uint32_t* counts, hipStream_t s)
{
  const uint32_t blocks = (p.subtiles + 3u) / 4u;
  hipLaunchKernelGGL((k_sort_upsweep<BITS, ITEMS>), dim3(blocks), dim3(256), 0, s, kin, p.n,
                     shift, p.subtiles, counts);
  const uint32_t m = (uint32_t)((size_t)(1u << BITS) * p.subtiles);
  const uint32_t nseg = (m + kScanSeg - 1u) / kScanSeg;
  uint32_t* partial = counts + (size_t)2048 * p.subtiles;  // past the largest table for n
  if (nseg > 1)  // one segment: its apply block reads no segment sums
    hipLaunchKernelGGL(k_scan_reduce, dim3(nseg), dim3(256), 0, s, counts, m, partial);
  hipLaunchKernelGGL(k_scan_apply, dim3(nseg), dim3(256), 0, s, counts, m, partial);
  hipLaunchKernelGGL((k_sort_downsweep<BITS, ITEMS>), dim3(blocks), dim3(256), 0, s, kin, iin,
                     p.n, shift, p.subtiles, counts, kout, iout);
  return RTP_OK;
}

template <int ITEMS>
int dispatch_bits(const SortPlan& p, int shift, const uint32_t* kin, const uint32_t* iin,
                  uint32_t* kout, uint32_t* iout, uint32_t* counts, hipStream_t s)
{
  switch (p.bits) {
  case 4: return run_pass<4, ITEMS>(p, shift, kin, iin, kout, iout, counts, s);
  case 5: return run_pass<5, ITEMS>(p, shift, kin, iin, kout, iout, counts, s);
  case 6: return run_pass<6, ITEMS>(p, shift, kin, iin, kout, iout, counts, s);
  case 7: return run_pass<7, ITEMS>(p, shift, kin, iin, kout, iout, counts, s);
  case 8: return run_pass<8, ITEMS>(p, shift, kin, iin, kout, iout, counts, s);
  case 9: return run_pass<9, ITEMS>(p, shift, kin, iin, kout, iout, counts, s);
  case 10: return run_pass<10, ITEMS>(p, shift, kin, iin, kout, iout, counts, s);
  case 11: return run_pass<11, ITEMS>(p, shift, kin, iin, kout, iout, counts, s);
  default: return fail(RTP_ERR_INVALID, "sort digit width out of range");
  }
}

int dispatch_pass(const SortPlan& p, int shift, const uint32_t* kin, const uint32_t* iin,
                  uint32_t* kout, uint32_t* iout, uint32_t* counts, hipStream_t s)
{
  switch (p.items) {
  case 4: return dispatch_bits<4>(p, shift, kin, iin, kout, iout, counts, s);
  case 8: return dispatch_bits<8>(p, shift, kin, iin, kout, iout, counts, s);
  case 16: return dispatch_bits<16>(p, shift, kin, iin, kout, iout, counts, s);
  case kSortItems: return dispatch_bits<kSortItems>(p, shift, kin, iin, kout, iout, counts, s);
  default: return fail(RTP_ERR_INVALID, "sort subtile size out of range");
  }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t v)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = v;
}

// Counting-sort ranks: each particle's place among the particles of its cell.  Thread t
// counts particle prev[t] (t without a previous order) and keeps its rank at rank[t]: in
// the previous thread order a block's particles sat in few cells and mostly still do.
// They are counted per cell in an LDS hash table first (open addressing over twice as many
// slots as the block has threads, so a probe always finds the cell or a free slot), and
// one device atomic per cell of the block reserves their places: far fewer device atomics
// than one per particle.
constexpr uint32_t kRankThreads = 1024;
constexpr uint32_t kRankSlots = 2 * kRankThreads;
constexpr uint32_t kNoKey = 0xffffffffu;  // keys are cell indices below 2^32 - 16
__global__ __launch_bounds__(kRankThreads) void k_count_ranks(const uint32_t* __restrict__ key,
                                                              const uint32_t* __restrict__ prev,
                                                              uint32_t n, uint32_t* __restrict__ count,
                                                              uint32_t* __restrict__ rank)
{
  __shared__ uint32_t tkey[kRankSlots];
  __shared__ uint32_t tcnt[kRankSlots];
  for (uint32_t s = threadIdx.x; s < kRankSlots; s += blockDim.x) {
    tkey[s] = kNoKey;
    tcnt[s] = 0u;
  }
  __syncthreads();
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t k = kNoKey, slot = 0u, local = 0u;
  if (i < n) {
    k = key[prev ? prev[i] : i];
    slot = (k * 2654435761u) >> 21;  // 11 bits: kRankSlots
    for (;;) {
      const uint32_t old = atomicCAS(&tkey[slot], kNoKey, k);
      if (old == kNoKey || old == k)
        break;
      slot = (slot + 1u) & (kRankSlots - 1u);
    }
    local = atomicAdd(&tcnt[slot], 1u);
  }
  __syncthreads();
  if (i < n && local == 0u)
    tcnt[slot] = atomicAdd(count + k, tcnt[slot]);  // the cell's count in, its first place out
  __syncthreads();
  if (i < n)
    rank[i] = tcnt[slot] + local;
}

// The particle i of thread t (i = prev[t]) goes to slot offset[key[i]] + rank[t] of the
// new thread order.
// ... and the next call's count table cleared (clearWords words of clearNext).
__global__ __launch_bounds__(256) void k_order_scatter(const uint32_t* __restrict__ key,
                                                       const uint32_t* __restrict__ prev,
                                                       const uint32_t* __restrict__ rank,
                                                       const uint32_t* __restrict__ offset, uint32_t n,
                                                       uint32_t* __restrict__ order,
                                                       uint32_t* __restrict__ clearNext,
                                                       uint32_t clearWords)
{
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    const uint32_t i = prev ? prev[t] : t;
    order[offset[key[i]] + rank[t]] = i;
  }
  if (t < clearWords)
    clearNext[t] = 0u;
}

inline uint32_t blocks_for(uint32_t n, uint32_t b) { return (n + b - 1u) / b; }

}  // namespace

void launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s)
{
  if (n)
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n, v);
}

bool launch_counting_order(const CountTables& ct, uint32_t m16, const uint32_t* key,
                           const uint32_t* prev, uint32_t* rank, uint32_t n, uint32_t* partial,
                           uint32_t* order, hipStream_t s)
{
  if (n == 0 || m16 == 0)
    return false;
  uint32_t* count = ct.table[ct.cur];
  const uint32_t nseg = (m16 + kScanSeg - 1u) / kScanSeg;
  hipLaunchKernelGGL(k_count_ranks, dim3(blocks_for(n, kRankThreads)), dim3(kRankThreads), 0, s,
                     key, prev, n, count, rank);
  if (nseg > 1)  // one segment: its apply block reads no segment sums
    hipLaunchKernelGGL(k_scan_reduce, dim3(nseg), dim3(256), 0, s, count, m16, partial);
  hipLaunchKernelGGL(k_scan_apply, dim3(nseg), dim3(256), 0, s, count, m16, partial);
  hipLaunchKernelGGL(k_order_scatter, dim3(blocks_for(std::max(n, ct.words), 256)), dim3(256), 0, s,
                     key, prev, rank, count, n, order, ct.table[1 - ct.cur], ct.words);
  return true;
}

uint32_t scan_partials_needed(uint32_t m16) { return (m16 + kScanSeg - 1u) / kScanSeg; }

SortPlan make_sort_plan(uint32_t n, uint32_t maxKey, int bitsOverride, int fixedItems)
{
  SortPlan p;
  p.n = n;
  int keyBits = 0;
  while (keyBits < 32 && (maxKey >> keyBits) != 0u)
    ++keyBits;
  if (keyBits == 0)
    keyBits = 1;
  const int maxBits = (bitsOverride >= 4 && bitsOverride <= 11) ? bitsOverride : 9;
  p.passes = (keyBits + maxBits - 1) / maxBits;
  p.bits = (keyBits + p.passes - 1) / p.passes;
  if (p.bits < 4)  // the count table scan works on whole 16-byte groups
    p.bits = 4;
  p.items = fixedItems ? fixedItems : sort_items(n);  // 4, 8, 16 or 32
  p.subtiles = (n + (uint32_t)(p.items * 64) - 1u) / (uint32_t)(p.items * 64);
  p.countWords = ((size_t)1u << p.bits) * p.subtiles;
  p.keyBits = keyBits;
  return p;
}

// Whether the current device gives one workgroup the LDS k_sort_small declares (checked
// once per device; otherwise small sorts take the multi-launch path).
bool small_sort_fits()
{
  static int ok[64] = {};  // 0 unknown, 1 yes, 2 no
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
    return false;
  if (ok[dev] == 0) {
    hipFuncAttributes a{};
    int limit = 0;
    const bool q = hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&k_sort_small<16>)) == hipSuccess &&
                   hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess;
    ok[dev] = (q && a.sharedSizeBytes <= (size_t)limit) ? 1 : 2;
  }
  return ok[dev] == 1;
}

// Passes of the one-workgroup sort: 4-bit digits over the significant bits.
static int small_sort_passes(const SortPlan& p) { return (p.keyBits + 3) / 4; }

// sort_pairs takes the one-workgroup sort up to kSmallSortFaster keys, where it beats the
// multi-launch sort (21.0 against 23.8 us at 4,096 keys, 32.5 against 25.1 at 6,144, on the
// boids flock); the persistent boids update, whose step runs in one launch, takes it up to
// kSmallSortMax.
constexpr uint32_t kSmallSortFaster = 4096;

static bool use_small_sort(const SortPlan& p) { return p.n <= kSmallSortFaster && small_sort_fits(); }

int small_sort_pass_count(const SortPlan& p)
{
  return (p.n <= kSmallSortMax && small_sort_fits()) ? small_sort_passes(p) : 0;
}

bool sort_result_in_b(const SortPlan& p)
{
  if (p.n == 0)
    return false;
  return ((use_small_sort(p) ? small_sort_passes(p) : p.passes) & 1) != 0;
}

int sort_pairs(const SortPlan& p, uint32_t* keyA, uint32_t* keyB, uint32_t* idxA, uint32_t* idxB,
               uint32_t* counts, bool* resultInB, hipStream_t s)
{
  *resultInB = false;
  if (p.n == 0)
    return RTP_OK;
  if (use_small_sort(p)) {  // one workgroup, all passes in LDS
    const int passes = small_sort_passes(p);
    if (p.n <= 4096u)
      hipLaunchKernelGGL(k_sort_small<8>, dim3(1), dim3(kSmallThreads), 0, s, keyA, keyB, idxA, idxB,
                         p.n, passes);
    else
      hipLaunchKernelGGL(k_sort_small<16>, dim3(1), dim3(kSmallThreads), 0, s, keyA, keyB, idxA,
                         idxB, p.n, passes);
    *resultInB = (passes & 1) != 0;
    return RTP_OK;
  }
  uint32_t *kin = keyA, *kout = keyB, *iin = idxA, *iout = idxB;
  bool inB = false;
  for (int pass = 0; pass < p.passes; ++pass) {
    const int rc = dispatch_pass(p, pass * p.bits, kin, pass == 0 ? nullptr : iin, kout, iout,
                                 counts, s);
    if (rc)
      return rc;
    std::swap(kin, kout);
    std::swap(iin, iout);
    inB = !inB;
  }
  *resultInB = inB;
  return RTP_OK;
}

void launch_fill_keys_and_reset_cells(const float4* pos, uint32_t n, uint32_t* key,
                                      uint2* startEnd, const GridConsts& g, hipStream_t s)
{
  const uint32_t m = n > g.numCells ? n : g.numCells;
  if (m == 0)
    return;
  hipLaunchKernelGGL(k_fill_keys_reset, dim3(blocks_for(m, 256)), dim3(256), 0, s, pos, n, key,
                     startEnd, g);
}

void launch_reset_cell_ids(uint32_t* key, uint32_t maxN, uint32_t numCells, hipStream_t s)
{
  if (maxN == 0)
    return;
  hipLaunchKernelGGL(k_reset_cell_ids, dim3(blocks_for(maxN, 256)), dim3(256), 0, s, key, maxN,
                     numCells);
}

void launch_fill_identity(uint32_t* idx, uint32_t begin, uint32_t end, hipStream_t s)
{
  if (end <= begin)
    return;
  hipLaunchKernelGGL(k_fill_identity, dim3(blocks_for(end - begin, 256)), dim3(256), 0, s, idx,
                     begin, end);
}

void launch_copy_f4(float4* dst, const float4* src, uint32_t begin, uint32_t end, hipStream_t s)
{
  if (end <= begin)
    return;
  hipLaunchKernelGGL(k_copy_f4, dim3(blocks_for(end - begin, 256)), dim3(256), 0, s, dst, src,
                     begin, end);
}

void launch_reorder_batch(const BatchView& bv, const ReorderArgs& a, hipStream_t s)
{
  if (bv.numBlocks)
    hipLaunchKernelGGL(k_reorder_batch, dim3(bv.numBlocks), dim3(bv.block), 0, s, a, bv);
}

void launch_reorder(const ReorderArgs& a, hipStream_t s)
{
  if (a.sortN == 0)
    return;
  hipLaunchKernelGGL(k_reorder, dim3(blocks_for(a.sortN, 256)), dim3(256), 0, s, a);
}

}  // namespace rtp
