// Device code of the grid build shared by its own kernels and by the persistent boids
// update: the sorted-slot gather with the cell table (reorder_body) and the one-workgroup
// LDS radix sort (small_sort_block).
#pragma once

#include "rtp_device.hpp"

namespace rtp {

constexpr uint32_t kSmallThreads = 512;

// The LDS of small_sort_block (144 KB): both ping-pong buffers of keys and indices, the
// digit-count table and the per-wave sums of its scan.
struct SmallSortLds {
  uint32_t sk[2][kSmallSortMax];
  uint32_t si[2][kSmallSortMax];
  uint16_t cnt[16 * kSmallThreads];
  uint32_t wsum[kSmallThreads / 64];
};

// Stable LSD sort of up to kSmallSortMax (key, index) pairs by one workgroup of
// kSmallThreads, 4-bit digits, all passes in one workgroup and in LDS: the pairs are read
// once (coalesced), ping-pong between two LDS buffers, and the result is written once to
// the buffer the multi-launch sort would leave it in (B after an odd number of passes).
// Thread t owns the contiguous chunk [t*CH, (t+1)*CH) of a pass's input.  Its digit counts
// sit in LDS column t of a digit-major table whose exclusive scan gives, for each
// (digit, thread), the number of keys of smaller digits plus those of the same digit in
// earlier chunks; the thread then scatters its chunk in order, so equal keys keep their
// input order.
// Every thread of the workgroup (kSmallThreads) must call it.
template <int CH>
__device__ __forceinline__ void small_sort_block(uint32_t* __restrict__ keyA,
                                                 uint32_t* __restrict__ keyB,
                                                 uint32_t* __restrict__ idxA,
                                                 uint32_t* __restrict__ idxB, uint32_t n, int passes,
                                                 SmallSortLds& lds)
{
  auto& sk = lds.sk;
  auto& si = lds.si;
  auto& cnt = lds.cnt;
  auto& wsum = lds.wsum;
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < n; i += kSmallThreads) {
    sk[0][i] = keyA[i];
    si[0][i] = i;
  }
  __syncthreads();
  const uint32_t lo = t * CH;
  for (int pass = 0; pass < passes; ++pass) {
    const int in = pass & 1;
    const int shift = 4 * pass;
    uint32_t k[CH], ix[CH];
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const uint32_t i = lo + (uint32_t)j;
      k[j] = i < n ? sk[in][i] : 0u;
      ix[j] = i < n ? si[in][i] : 0u;
    }
#pragma unroll
    for (int d = 0; d < 16; ++d)
      cnt[d * kSmallThreads + t] = 0;
#pragma unroll
    for (int j = 0; j < CH; ++j)
      if (lo + (uint32_t)j < n)
        cnt[((k[j] >> shift) & 15u) * kSmallThreads + t] += 1;
    __syncthreads();
    // exclusive scan of the table in linear order: thread t takes entries [16t, 16t + 16)
    uint32_t loc[16];
    uint32_t run = 0;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      loc[e] = run;
      run += cnt[16 * t + e];
    }
    uint32_t incl = run;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t v = __shfl_up(incl, o, 64);
      if ((t & 63u) >= (uint32_t)o)
        incl += v;
    }
    if ((t & 63u) == 63u)
      wsum[t >> 6] = incl;
    __syncthreads();
    uint32_t base = incl - run;
    for (uint32_t w = 0; w < (t >> 6); ++w)
      base += wsum[w];
#pragma unroll
    for (int e = 0; e < 16; ++e)
      cnt[16 * t + e] = (uint16_t)(base + loc[e]);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      if (lo + (uint32_t)j < n) {
        const uint32_t slot = ((k[j] >> shift) & 15u) * kSmallThreads + t;
        const uint32_t pos = cnt[slot];
        cnt[slot] = (uint16_t)(pos + 1u);
        sk[1 - in][pos] = k[j];
        si[1 - in][pos] = ix[j];
      }
    }
    __syncthreads();
  }
  const int res = passes & 1;
  uint32_t* kout = res ? keyB : keyA;
  uint32_t* iout = res ? idxB : idxA;
  for (uint32_t i = t; i < n; i += kSmallThreads) {
    kout[i] = sk[res][i];
    iout[i] = si[res][i];
  }
}


// Gather by idx + cell start/end (with the cap folded in: the thread that closes a run
// knows the run's start from a probe CAP+1 entries back, and a binary search when the run
// is longer than that).
// Sorted slot i: gather the payloads of the element that landed there and record the cell
// boundaries.  kbase and ibase are the key and index bases of a batched scene (0 for a
// model of its own): keys are compared as stored, cells and gather sources are local.
__device__ __forceinline__ void reorder_body(uint32_t i, const ReorderArgs& a, uint32_t kbase,
                                             uint32_t ibase)
{
  const uint32_t j = a.idx[i] - ibase;
  if (a.dstA)
    a.dstA[i] = a.srcA[j];
  if (a.dstB)
    a.dstB[i] = a.srcB[j];
  if (a.dstC) {
    float4 p = a.srcC[j];
    if (a.clampC && i < a.n) {  // fld_applyBoundaryCondition, fluids.cl:435-439
      p.x = clampf(p.x, a.lo[0], a.hi[0]);
      p.y = clampf(p.y, a.lo[1], a.hi[1]);
      p.z = clampf(p.z, a.lo[2], a.hi[2]);
      p.w = 0.0f;
    }
    a.dstC[i] = p;
  }
  if (i >= a.n)
    return;
  const uint32_t kr = a.key[i];
  const uint32_t k = kr - kbase;
  if (k >= a.g.numCells)  // grid.cl:107,128: only keys of real cells are recorded
    return;
  uint32_t* se = reinterpret_cast<uint32_t*>(a.startEnd);
  if (i > 0 && kr != a.key[i - 1])  // fillStartCell, grid.cl:101-117 (never for i == 0)
    se[2u * k] = i;
  // fillEndCell, grid.cl:122-138: key[i+1] is read even for the last active particle.
  const bool isEnd = (i + 1u >= a.maxN) || (kr != a.key[i + 1u]);
  if (!isEnd)
    return;
  // adjustEndCell, grid.cl:143-152, with start = run start, or 1 (the reset value) when
  // the run starts at 0.
  const uint32_t cap = a.g.cap;
  uint32_t endv = i;
  if (i >= cap + 1u && a.key[i - cap - 1u] == kr) {
    uint32_t lo = 0, hi = i - cap - 1u;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (a.key[mid] < kr)
        lo = mid + 1u;
      else
        hi = mid;
    }
    const uint32_t startEff = lo == 0 ? 1u : lo;
    if (i > startEff)
      endv = startEff + min(i - startEff, cap);
  }
  se[2u * k + 1u] = endv;
}


}  // namespace rtp
