// Learned SA closure: a 5 -> 64 -> 64 -> 1 tanh MLP (see sg_mlp.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace sg {

constexpr int MLP_IN = 5;
constexpr int MLP_H = 64;
// parameter offsets in the packed FP32 vector
constexpr int MLP_W1 = 0;
constexpr int MLP_B1 = MLP_W1 + MLP_H * MLP_IN;  // 320
constexpr int MLP_W2 = MLP_B1 + MLP_H;           // 384
constexpr int MLP_B2 = MLP_W2 + MLP_H * MLP_H;   // 4480
constexpr int MLP_W3 = MLP_B2 + MLP_H;           // 4544
constexpr int MLP_B3 = MLP_W3 + MLP_H;           // 4608
constexpr int MLP_NPARAM = MLP_B3 + 1;           // 4609

// Device work space of one network (all pointers device memory).
struct MlpWork {
  float* P = nullptr;      // FP32 master parameters
  float* m = nullptr;      // Adam first moment
  float* v = nullptr;      // Adam second moment
  float* grad = nullptr;   // last gradient
  __bf16* W2b = nullptr;   // BF16 copy of W2 used by the matrix cores
  int cap = 0;             // rows the activation buffers hold
  int nsplit = 16;         // split-K factor of the dW2 product
  float *H1 = nullptr, *H2 = nullptr, *dH1 = nullptr, *dH2 = nullptr, *dY = nullptr;
  __bf16 *H1b = nullptr, *H2b = nullptr, *dH2b = nullptr;
  float *loss_part = nullptr, *colpart = nullptr, *w2part = nullptr, *loss = nullptr;
};

hipError_t gemm_bf16(int M, int N, int K, const __bf16* A, long sam, long sak, const __bf16* B, long sbk, long sbn,
                     float* C, long scm, long scn, int splitk, long splitc, int epi, const float* bias, __bf16* Cb,
                     hipStream_t s);
hipError_t mlp_w2_to_bf16(const float* P, __bf16* W2b, hipStream_t s);
hipError_t mlp_forward_backward(const MlpWork& w, int n, const float* X, const int* idx, const float* T, int ntot,
                                float* Y, float* grad, float* loss, hipStream_t s);
hipError_t mlp_adam(const MlpWork& w, const float* grad, float lr, int step, hipStream_t s);

}  // namespace sg
