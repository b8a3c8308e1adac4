// The rendering side of a frame: camera-distance keys (utils.cl:35-52), the fluid colour
// (fluids.cl:458-478), the boids colour (boids.cl:38-41), the grid detector (grid.cl:43-60)
// and the payload gathers of the camera-distance sort (RadixSort.cpp:163-175).
//
// All of these are streaming kernels: one particle (or cell) per thread, no neighbour
// reads, bounded by HBM bandwidth.  The sort itself is the grid's radix sort (grid_sort.hip)
// on 20-bit keys.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rtp_device.hpp"
#include "rtp_internal.hpp"

namespace rtp {
namespace {

constexpr float kFarDist = 1000000.0f;  // FAR_DIST, define.cl:10

inline uint32_t blocks_for(uint32_t n, uint32_t b) { return (n + b - 1u) / b; }

// fillCameraDist, utils.cl:43-52: (uint)max(FAR_DIST - length(p.xyz - cam) * 100, 0).
// cam holds u_cameraPos in xyz; length is the correctly rounded square root of the
// left-to-right sum of squares.
__global__ __launch_bounds__(256) void k_camera_key(const float4* __restrict__ pos,
                                                    const float4* __restrict__ cam,
                                                    uint32_t* __restrict__ key, uint32_t n)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const float4 p = pos[i];
  const float4 c = *cam;
  const float dx = p.x - c.x, dy = p.y - c.y, dz = p.z - c.z;
  const float len = sqrtf((dx * dx + dy * dy) + dz * dz);
  key[i] = (uint32_t)fmaxf(kFarDist - len * 100.0f, 0.0f);
}

// fld_fillFluidColor, fluids.cl:458-478: blue at rest density, towards light blue where the
// fluid is sparse (constraint > 0) and dark blue where it is compressed.  Each component is
// blue + (constraint * (other - blue)) / 0.35 as the reference's float4 expression computes it.
__global__ __launch_bounds__(256) void k_fluid_color(const float* __restrict__ density,
                                                     float restDensity, float4* __restrict__ col,
                                                     uint32_t n)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const float cst = 1.0f - density[i] / restDensity;
  float4 c = make_float4(0.0f, 0.1f, 1.0f, 0.5f);
  if (cst > 0.0f) {  // lightBlue - blue
    c.x = c.x + (cst * (0.7f - 0.0f)) / 0.35f;
    c.y = c.y + (cst * (0.7f - 0.1f)) / 0.35f;
    c.z = c.z + (cst * (1.0f - 1.0f)) / 0.35f;
    c.w = c.w + (cst * (0.5f - 0.5f)) / 0.35f;
  } else if (cst < 0.0f) {  // blue - darkBlue
    c.x = c.x + (cst * (0.0f - 0.0f)) / 0.35f;
    c.y = c.y + (cst * (0.1f - 0.0f)) / 0.35f;
    c.z = c.z + (cst * (1.0f - 0.8f)) / 0.35f;
    c.w = c.w + (cst * (0.5f - 0.5f)) / 0.35f;
  }
  col[i] = c;
}

__global__ __launch_bounds__(256) void k_fill_f4v(float4* __restrict__ dst, float4 v,
                                                  uint32_t begin, uint32_t end)
{
  const uint32_t i = begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < end)
    dst[i] = v;
}

// fillGridDetector, grid.cl:51-60: the float8 of every cell holding an active particle
// becomes all ones (the scalar 1.0f widens to every component).  det has two float4 per
// cell.
__global__ __launch_bounds__(256) void k_detector_fill(const float4* __restrict__ pos,
                                                       float4* __restrict__ det, uint32_t n,
                                                       GridConsts g)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const uint32_t cell = cell1d(pos[i], g);
  if (cell < g.numCells) {
    const float4 one = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    det[2u * cell] = one;
    det[2u * cell + 1u] = one;
  }
}

// permutateFloat4 (radixSort.cl:179-188) for up to four payloads at once over [begin, end):
// dst[i] = src[idx[i]].  Every idx[i] lies in the range the permutation was sorted over.
__global__ __launch_bounds__(256) void k_gather_f4(const uint32_t* __restrict__ idx,
                                                   uint32_t begin, uint32_t end, GatherF4 g)
{
  const uint32_t i = begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= end)
    return;
  const uint32_t j = idx[i];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (g.src[k])
      g.dst[k][i] = g.src[k][j];
}

}  // namespace

void launch_camera_key(const float4* pos, const float4* cam, uint32_t* key, uint32_t n,
                       hipStream_t s)
{
  if (n)
    hipLaunchKernelGGL(k_camera_key, dim3(blocks_for(n, 256)), dim3(256), 0, s, pos, cam, key, n);
}

void launch_fluid_color(const float* density, float restDensity, float4* col, uint32_t n,
                        hipStream_t s)
{
  if (n)
    hipLaunchKernelGGL(k_fluid_color, dim3(blocks_for(n, 256)), dim3(256), 0, s, density,
                       restDensity, col, n);
}

void launch_fill_f4v(float4* dst, float4 v, uint32_t begin, uint32_t end, hipStream_t s)
{
  if (end > begin)
    hipLaunchKernelGGL(k_fill_f4v, dim3(blocks_for(end - begin, 256)), dim3(256), 0, s, dst, v,
                       begin, end);
}

void launch_detector(const float4* pos, float4* det, uint32_t n, const GridConsts& g,
                     hipStream_t s)
{
  // resetGridDetector (grid.cl:43-46) by a fill kernel, so the step can sit in a hipGraph
  launch_fill_f4v(det, make_float4(0.0f, 0.0f, 0.0f, 0.0f), 0, 2u * g.numCells, s);
  if (n)
    hipLaunchKernelGGL(k_detector_fill, dim3(blocks_for(n, 256)), dim3(256), 0, s, pos, det, n, g);
}

void launch_gather_f4(const uint32_t* idx, uint32_t begin, uint32_t end, const GatherF4& g,
                      hipStream_t s)
{
  if (end > begin)
    hipLaunchKernelGGL(k_gather_f4, dim3(blocks_for(end - begin, 256)), dim3(256), 0, s, idx,
                       begin, end, g);
}

}  // namespace rtp
