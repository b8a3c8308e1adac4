// Learned warm-start network: device parameter views and launchers (mlp.hip).
#pragma once

#include "rtp_internal.hpp"

namespace rtp {

constexpr int kMlpIn = 8;
constexpr int kMlpHidden = 64;
// Offsets in the fp32 parameter vector.
constexpr int kMlpW1 = 0;                      // [8][64]
constexpr int kMlpB1 = kMlpW1 + 8 * 64;        // [64]
constexpr int kMlpW2 = kMlpB1 + 64;            // [64][64]
constexpr int kMlpB2 = kMlpW2 + 64 * 64;       // [64]
constexpr int kMlpW3 = kMlpB2 + 64;            // [64]
constexpr int kMlpB3 = kMlpW3 + 64;            // [1]
constexpr int kMlpParams = kMlpB3 + 1;         // 4801

// What the kernels read: the fp32 master vector (biases, head), the bf16 matrix operands
// and the feature normalisation.
struct MlpDev {
  const float* p;
  const short* w1t;  // [64][16]  W1 transposed, k padded with zeros
  const short* w2t;  // [64][64]  W2 transposed (forward operand)
  const short* w2;   // [64][64]  W2 (backward operand)
  short* w1tMut;
  short* w2tMut;
  short* w2Mut;
  float mu[8];
  float invSigma[8];
  float outScale;
};

void launch_mlp_forward(const float* feat, uint32_t n, const MlpDev& m, float* out, hipStream_t s);
void launch_mlp_train(const float* feat, const float* target, const uint32_t* perm, uint32_t n,
                      const MlpDev& m, float* grad, float* loss, hipStream_t s);
void launch_mlp_adam(float* p, float* grad, float* m1, float* m2, float lr, int step,
                     const MlpDev& m, hipStream_t s);
void launch_mlp_pack(const float* p, const MlpDev& m, hipStream_t s);
void launch_zero(float* p, uint32_t n, hipStream_t s);
void launch_affine_perm(uint32_t* perm, uint32_t n, uint32_t a, uint32_t c, hipStream_t s);

}  // namespace rtp
