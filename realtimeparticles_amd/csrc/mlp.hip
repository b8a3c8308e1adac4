// The learned warm-start network: an 8 -> 64 -> 64 -> 1 MLP with GELU, BF16 weights and
// activations on the matrix cores (v_mfma_f32_16x16x16_bf16) with FP32 accumulation.
// No reference counterpart: it predicts a particle's initial lambda for the PBF solver.
//
// One wave owns 64 rows (particles or training samples).  Layers 1 and 2 are 4 x 4 tiles
// of 16 x 16 outputs per wave; the 64 -> 1 head and the elementwise work run on the VALU.
// Activations are staged through LDS to turn the accumulator layout into the next layer's
// operand layout.
//
// Operand layout of v_mfma_f32_16x16x16_bf16 (lane l, j = 0..3):
//   A (16 x 16, M x K): A[l % 16][4 * (l / 16) + j]
//   B (16 x 16, K x N): B[4 * (l / 16) + j][l % 16]
//   D (16 x 16, M x N): D[4 * (l / 16) + j][l % 16]
//
// Parameter vector (fp32 master copy, 4801 values):
//   W1[8][64] (input k, hidden n) | b1[64] | W2[64][64] (k, n) | b2[64] | w3[64] | b3
#include "rtp_mlp.hpp"

namespace rtp {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short f2bf(float f)
{
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (short)(u >> 16);
}

__device__ __forceinline__ float bf2f(short b) { return __uint_as_float(((uint32_t)(uint16_t)b) << 16); }

constexpr float kGeluA = 0.7978845608028654f;  // sqrt(2 / pi)
constexpr float kGeluB = 0.044715f;

// tanh-form GELU, 0.5 z (1 + tanh u) with u = sqrt(2/pi) (z + 0.044715 z^3), evaluated as
// z * sigmoid(2u) = z / (1 + exp(-2u)) with the hardware exponential and reciprocal (about
// 1e-6 relative; the network is trained with these same functions).
__device__ __forceinline__ float sigmoid2u(float z)
{
  const float u = kGeluA * (z + kGeluB * (z * z) * z);
  return __builtin_amdgcn_rcpf(1.0f + __expf(-2.0f * u));
}

__device__ __forceinline__ float gelu(float z) { return z * sigmoid2u(z); }

// d/dz [z sigmoid(2u)] = s + z s (1 - s) 2 u'(z),  u' = sqrt(2/pi) (1 + 3 * 0.044715 z^2)
__device__ __forceinline__ float gelu_grad(float z)
{
  const float sg = sigmoid2u(z);
  return sg + z * sg * (1.0f - sg) * (2.0f * kGeluA) * (1.0f + 3.0f * kGeluB * z * z);
}

// Four consecutive bf16 (8-byte aligned) as one operand register.
__device__ __forceinline__ bf16x4 ld4(const short* p) { return *reinterpret_cast<const bf16x4*>(p); }

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

struct WaveLds {
  short X[64][16];   // normalised inputs (bf16, k padded to 16 with zeros)
  short H1[64][64];  // layer-1 activations (bf16)
};

// Inputs of the wave's 64 rows into LDS; returns whether this lane's row exists.
__device__ __forceinline__ bool load_inputs(const float* __restrict__ feat, uint32_t row0, uint32_t n,
                                            const MlpDev& m, short (*X)[16], const uint32_t* perm)
{
  const uint32_t l = threadIdx.x;
  const uint32_t r = row0 + l;
  const bool valid = r < n;
  const uint32_t src = valid ? (perm ? perm[r] : r) : 0u;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const float x = valid ? feat[(size_t)src * 8 + k] : 0.0f;
    X[l][k] = f2bf((x - m.mu[k]) * m.invSigma[k]);
    X[l][8 + k] = 0;
  }
  return valid;
}

// Layer 1 for the wave: z1 = x W1 + b1, writing GELU(z1) (bf16) to H1 and optionally z1.
__device__ __forceinline__ void layer1(const MlpDev& m, short (*X)[16], short (*H1)[64], float (*Z1)[64])
{
  const uint32_t l = threadIdx.x, lr = l % 16u, lq = l / 16u;
  float b1[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
    b1[nt] = m.p[kMlpB1 + nt * 16 + lr];
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      f32x4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
      acc = mfma(ld4(&X[mt * 16 + lr][4 * lq]), ld4(m.w1t + (nt * 16 + lr) * 16 + 4 * lq), acc);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = mt * 16 + 4 * lq + j, col = nt * 16 + lr;
        const float z = acc[j] + b1[nt];
        if (Z1)
          Z1[row][col] = z;
        H1[row][col] = f2bf(gelu(z));
      }
    }
}

// Layer 2 tile (mt, nt): z2 = h1 W2 + b2 for rows mt*16.., columns nt*16..
__device__ __forceinline__ f32x4 layer2_tile(const MlpDev& m, short (*H1)[64], int mt, int nt)
{
  const uint32_t l = threadIdx.x, lr = l % 16u, lq = l / 16u;
  f32x4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
  for (int kt = 0; kt < 4; ++kt)
    acc = mfma(ld4(&H1[mt * 16 + lr][kt * 16 + 4 * lq]), ld4(m.w2t + (nt * 16 + lr) * 64 + kt * 16 + 4 * lq), acc);
  return acc;
}

// Sum over the 16 lanes that share l / 16.
__device__ __forceinline__ float sum16(float v)
{
  v += __shfl_xor(v, 8, 64);
  v += __shfl_xor(v, 4, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 1, 64);
  return v;
}

// Inference: each wave keeps its lanes' weight operands (W1, W2 fragments, biases, head)
// in registers and runs 64-row tiles in a grid-stride loop.
__global__ __launch_bounds__(64) void k_mlp_forward(const float* __restrict__ feat, uint32_t n, MlpDev m,
                                                    float* __restrict__ out)
{
  __shared__ WaveLds s;
  const uint32_t l = threadIdx.x, lr = l % 16u, lq = l / 16u;
  bf16x4 w1f[4], w2f[4][4];
  float b1[4], b2[4], w3[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    w1f[nt] = ld4(m.w1t + (nt * 16 + lr) * 16 + 4 * lq);
#pragma unroll
    for (int kt = 0; kt < 4; ++kt)
      w2f[nt][kt] = ld4(m.w2t + (nt * 16 + lr) * 64 + kt * 16 + 4 * lq);
    b1[nt] = m.p[kMlpB1 + nt * 16 + lr];
    b2[nt] = m.p[kMlpB2 + nt * 16 + lr];
    w3[nt] = m.p[kMlpW3 + nt * 16 + lr];
  }
  const float b3 = m.p[kMlpB3];
  for (uint32_t row0 = blockIdx.x * 64u; row0 < n; row0 += gridDim.x * 64u) {
    (void)load_inputs(feat, row0, n, m, s.X, nullptr);
    __syncthreads();
    for (int mt = 0; mt < 4; ++mt) {
      const bf16x4 xa = ld4(&s.X[mt * 16 + lr][4 * lq]);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
        acc = mfma(xa, w1f[nt], acc);
#pragma unroll
        for (int j = 0; j < 4; ++j)
          s.H1[mt * 16 + 4 * lq + j][nt * 16 + lr] = f2bf(gelu(acc[j] + b1[nt]));
      }
    }
    __syncthreads();
    for (int mt = 0; mt < 4; ++mt) {
      bf16x4 ha[4];
#pragma unroll
      for (int kt = 0; kt < 4; ++kt)
        ha[kt] = ld4(&s.H1[mt * 16 + lr][kt * 16 + 4 * lq]);
      float y[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
        for (int kt = 0; kt < 4; ++kt)
          acc = mfma(ha[kt], w2f[nt][kt], acc);
#pragma unroll
        for (int j = 0; j < 4; ++j)
          y[j] += gelu(acc[j] + b2[nt]) * w3[nt];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float v = sum16(y[j]);
        const uint32_t r = row0 + mt * 16 + 4 * lq + j;
        if (lr == 0 && r < n)
          out[r] = (v + b3) * m.outScale;
      }
    }
    __syncthreads();  // the next tile overwrites X and H1
  }
}

// Forward + backward of the mean squared error over a batch; gradients are added to grad
// (the parameter layout above) and the squared-error sum to loss.
struct TrainLds {
  short X[64][16];
  short H1[64][64];
  float Z1[64][64];
  float Z2[64][64];
  float DZ[64][64];  // dz2, then dz1
  float Y[64];
  float DY[64];
};

__global__ __launch_bounds__(64) void k_mlp_train(const float* __restrict__ feat,
                                                  const float* __restrict__ target,
                                                  const uint32_t* __restrict__ perm, uint32_t n,
                                                  float invBatch, MlpDev m, float* __restrict__ grad,
                                                  float* __restrict__ loss)
{
  __shared__ TrainLds s;
  const uint32_t l = threadIdx.x, lr = l % 16u, lq = l / 16u;
  const uint32_t row0 = blockIdx.x * 64u;
  const bool valid = load_inputs(feat, row0, n, m, s.X, perm);
  __syncthreads();
  layer1(m, s.X, s.H1, s.Z1);
  __syncthreads();
  // layer 2 and the head
  for (int mt = 0; mt < 4; ++mt) {
    float y[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
    for (int nt = 0; nt < 4; ++nt) {
      const f32x4 acc = layer2_tile(m, s.H1, mt, nt);
      const int col = nt * 16 + lr;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float z = acc[j] + m.p[kMlpB2 + col];
        s.Z2[mt * 16 + 4 * lq + j][col] = z;
        y[j] += gelu(z) * m.p[kMlpW3 + col];
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float v = sum16(y[j]);
      if (lr == 0)
        s.Y[mt * 16 + 4 * lq + j] = v + m.p[kMlpB3];
    }
  }
  __syncthreads();
  // d loss / d y for this lane's row
  {
    const uint32_t r = row0 + l;
    const float t = valid ? target[perm ? perm[r] : r] / m.outScale : 0.0f;
    const float e = valid ? s.Y[l] - t : 0.0f;
    s.DY[l] = 2.0f * e * invBatch;
    float se = e * e, db3 = 2.0f * e * invBatch;
    for (int o = 32; o > 0; o >>= 1) {
      se += __shfl_xor(se, o, 64);
      db3 += __shfl_xor(db3, o, 64);
    }
    if (l == 0) {
      atomicAdd(loss, se);
      atomicAdd(grad + kMlpB3, db3);
    }
  }
  __syncthreads();
  // head and layer-2 activation: lane l owns hidden unit l
  {
    float dw3 = 0.0f, db2 = 0.0f;
    const float w3 = m.p[kMlpW3 + l];
    for (int r = 0; r < 64; ++r) {
      const float z = s.Z2[r][l];
      const float dy = s.DY[r];
      dw3 += gelu(z) * dy;
      const float dz = dy * w3 * gelu_grad(z);
      s.DZ[r][l] = dz;
      db2 += dz;
    }
    atomicAdd(grad + kMlpW3 + l, dw3);
    atomicAdd(grad + kMlpB2 + l, db2);
  }
  __syncthreads();
  // dW2[k][n] = sum_rows h1[row][k] dz2[row][n]   (M = k, N = n, K = rows)
  for (int mt = 0; mt < 4; ++mt)
    for (int nt = 0; nt < 4; ++nt) {
      f32x4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
      for (int kt = 0; kt < 4; ++kt) {
        bf16x4 a, b;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int row = kt * 16 + 4 * lq + j;
          a[j] = s.H1[row][mt * 16 + lr];
          b[j] = f2bf(s.DZ[row][nt * 16 + lr]);
        }
        acc = mfma(a, b, acc);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
        atomicAdd(grad + kMlpW2 + (mt * 16 + 4 * lq + j) * 64 + nt * 16 + lr, acc[j]);
    }
  // dh1[row][k] = sum_n dz2[row][n] W2[k][n]   (M = rows, N = k, K = n) -> dz1
  float dz1[4][4][4];
  for (int mt = 0; mt < 4; ++mt)
    for (int nt = 0; nt < 4; ++nt) {
      f32x4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
      for (int kt = 0; kt < 4; ++kt) {
        bf16x4 a;
#pragma unroll
        for (int j = 0; j < 4; ++j)
          a[j] = f2bf(s.DZ[mt * 16 + lr][kt * 16 + 4 * lq + j]);
        acc = mfma(a, ld4(m.w2 + (nt * 16 + lr) * 64 + kt * 16 + 4 * lq), acc);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
        dz1[mt][nt][j] = acc[j] * gelu_grad(s.Z1[mt * 16 + 4 * lq + j][nt * 16 + lr]);
    }
  __syncthreads();  // every lane has read dz2
  for (int mt = 0; mt < 4; ++mt)
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        s.DZ[mt * 16 + 4 * lq + j][nt * 16 + lr] = dz1[mt][nt][j];
  __syncthreads();
  {
    float db1 = 0.0f;
    for (int r = 0; r < 64; ++r)
      db1 += s.DZ[r][l];
    atomicAdd(grad + kMlpB1 + l, db1);
  }
  // dW1[k][n] = sum_rows x[row][k] dz1[row][n]   (M = k padded to 16, N = n, K = rows)
  for (int nt = 0; nt < 4; ++nt) {
    f32x4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
    for (int kt = 0; kt < 4; ++kt) {
      bf16x4 a, b;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = kt * 16 + 4 * lq + j;
        a[j] = s.X[row][lr];
        b[j] = f2bf(s.DZ[row][nt * 16 + lr]);
      }
      acc = mfma(a, b, acc);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = 4 * lq + j;
      if (k < 8)
        atomicAdd(grad + kMlpW1 + k * 64 + nt * 16 + lr, acc[j]);
    }
  }
}

// Adam on the fp32 master parameters, then the bf16 operand copies; clears the gradient.
__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ p, float* __restrict__ grad,
                                                  float* __restrict__ m1, float* __restrict__ m2,
                                                  float lr, float b1, float b2, float eps, float c1,
                                                  float c2, MlpDev m)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint32_t)kMlpParams)
    return;
  const float g = grad[i];
  grad[i] = 0.0f;
  const float a = b1 * m1[i] + (1.0f - b1) * g;
  const float v = b2 * m2[i] + (1.0f - b2) * g * g;
  m1[i] = a;
  m2[i] = v;
  p[i] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

// Rebuild the operand copies from the master parameters.
__global__ __launch_bounds__(256) void k_mlp_pack(const float* __restrict__ p, MlpDev m)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 64u * 16u) {  // w1t[n][k], k padded
    const uint32_t nn = i / 16u, k = i % 16u;
    m.w1tMut[i] = k < 8u ? f2bf(p[kMlpW1 + k * 64u + nn]) : (short)0;
  }
  if (i < 64u * 64u) {
    const uint32_t a = i / 64u, b = i % 64u;
    m.w2tMut[i] = f2bf(p[kMlpW2 + b * 64u + a]);  // w2t[n = a][k = b]
    m.w2Mut[i] = f2bf(p[kMlpW2 + a * 64u + b]);   // w2[k = a][n = b]
  }
}

__global__ void k_zero(float* __restrict__ p, uint32_t n)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = 0.0f;
}

__global__ void k_affine_perm(uint32_t* __restrict__ perm, uint32_t n, uint32_t a, uint32_t c)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    perm[i] = (uint32_t)(((uint64_t)a * i + c) % n);
}

inline uint32_t blocks_for(uint32_t n, uint32_t b) { return (n + b - 1u) / b; }

}  // namespace

void launch_mlp_forward(const float* feat, uint32_t n, const MlpDev& m, float* out, hipStream_t s)
{
  if (n == 0)
    return;
  const uint32_t tiles = blocks_for(n, 64);
  hipLaunchKernelGGL(k_mlp_forward, dim3(tiles < 4096u ? tiles : 4096u), dim3(64), 0, s, feat, n, m, out);
}

void launch_mlp_train(const float* feat, const float* target, const uint32_t* perm, uint32_t n,
                      const MlpDev& m, float* grad, float* loss, hipStream_t s)
{
  if (n == 0)
    return;
  hipLaunchKernelGGL(k_mlp_train, dim3(blocks_for(n, 64)), dim3(64), 0, s, feat, target, perm, n,
                     1.0f / (float)n, m, grad, loss);
}

void launch_mlp_adam(float* p, float* grad, float* m1, float* m2, float lr, int step,
                     const MlpDev& m, hipStream_t s)
{
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_mlp_adam, dim3(blocks_for(kMlpParams, 256)), dim3(256), 0, s, p, grad, m1,
                     m2, lr, b1, b2, 1e-8f, c1, c2, m);
}

void launch_mlp_pack(const float* p, const MlpDev& m, hipStream_t s)
{
  hipLaunchKernelGGL(k_mlp_pack, dim3(blocks_for(64 * 64, 256)), dim3(256), 0, s, p, m);
}

void launch_zero(float* p, uint32_t n, hipStream_t s)
{
  if (n)
    hipLaunchKernelGGL(k_zero, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n);
}

void launch_affine_perm(uint32_t* perm, uint32_t n, uint32_t a, uint32_t c, hipStream_t s)
{
  if (n)
    hipLaunchKernelGGL(k_affine_perm, dim3(blocks_for(n, 256)), dim3(256), 0, s, perm, n, a, c);
}

}  // namespace rtp
