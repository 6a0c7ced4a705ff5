// Learned closure: a 5 -> 64 -> 64 -> 1 tanh MLP mapping local flow features
// to the SA production multiplier beta = 1 + MLP(features).
//
// The 64 x 64 layer runs on the matrix cores: BF16 operands, FP32
// accumulation (v_mfma_f32_16x16x32_bf16; fragment layout, lane l of the
// wave: A[l%16][8(l/16)+t], B[8(l/16)+t][l%16], C[4(l/16)+i][l%16]).  The
// 5 -> 64 and 64 -> 1 layers, activations, loss, Adam and every reduction
// are FP32; reductions use fixed orders, so training is bitwise
// reproducible.  Parameters (FP32 master copy, 4609 floats):
//   W1[64][5] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3
#include <hip/hip_runtime.h>

#include "sg_closure.hpp"
#include "sg_mlp.hpp"

namespace sg {

// C[m][n] = sum_k A[m][k] B[k][n] (+ epilogue).  A, B bf16 with element
// strides; C f32.  Block = 4 waves computing a 64 x 64 tile (each wave 32 x
// 32 = 2 x 2 MFMA tiles); blockIdx.z splits K (partials at C + z * splitc).
// epi 0: C = acc; 1: C = tanh(acc + bias[n]) and Cb (bf16) = C.
__global__ __launch_bounds__(256) void gemm_bf16_kernel(int M, int N, int K, int kchunk, const __bf16* A, long sam,
                                                        long sak, const __bf16* B, long sbk, long sbn, float* C,
                                                        long scm, long scn, long splitc, int epi, const float* bias,
                                                        __bf16* Cb) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m_base = blockIdx.y * 64 + (w >> 1) * 32, n_base = blockIdx.x * 64 + (w & 1) * 32;
  const int k_begin = blockIdx.z * kchunk, k_end = min(K, k_begin + kchunk);
  f32x4 acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int r16 = lane & 15, kq = 8 * (lane >> 4);
  for (int k0 = k_begin; k0 < k_end; k0 += 32) {
    bf16x8 af[2], bf[2];
    for (int a = 0; a < 2; ++a) {
      const int m = m_base + 16 * a + r16;
      for (int t = 0; t < 8; ++t) {
        const int k = k0 + kq + t;
        af[a][t] = (m < M && k < k_end) ? A[m * sam + k * sak] : (__bf16)0.0f;
      }
    }
    for (int b = 0; b < 2; ++b) {
      const int n = n_base + 16 * b + r16;
      for (int t = 0; t < 8; ++t) {
        const int k = k0 + kq + t;
        bf[b][t] = (n < N && k < k_end) ? B[k * sbk + n * sbn] : (__bf16)0.0f;
      }
    }
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bf[b], acc[a][b], 0, 0, 0);
  }
  float* Cz = C + (long)blockIdx.z * splitc;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int i = 0; i < 4; ++i) {
        const int m = m_base + 16 * a + 4 * (lane >> 4) + i, n = n_base + 16 * b + r16;
        if (m >= M || n >= N) continue;
        float v = acc[a][b][i];
        if (epi == 1) {
          v = tanh_fast(v + bias[n]);
          Cb[m * scm + n * scn] = (__bf16)v;
        }
        Cz[m * scm + n * scn] = v;
      }
}

hipError_t gemm_bf16(int M, int N, int K, const __bf16* A, long sam, long sak, const __bf16* B, long sbk, long sbn,
                     float* C, long scm, long scn, int splitk, long splitc, int epi, const float* bias, __bf16* Cb,
                     hipStream_t s) {
  const int kchunk = ((K + splitk - 1) / splitk + 31) / 32 * 32;
  dim3 grid((N + 63) / 64, (M + 63) / 64, splitk);
  hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, s, M, N, K, kchunk, A, sam, sak, B, sbk, sbn, C, scm, scn,
                     splitc, epi, bias, Cb);
  return hipGetLastError();
}

// ---- elementwise / reduction kernels -------------------------------------
// layer 1: a1 = W1 x + b1, h1 = tanh(a1) (f32 and bf16), rows gathered by idx
__global__ void mlp_layer1_kernel(int n, const float* X, const int* idx, const float* P, float* H1, __bf16* H1b) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int src = idx ? idx[r] : r;
  float x[MLP_IN];
  for (int c = 0; c < MLP_IN; ++c) x[c] = X[(long)src * MLP_IN + c];
  const float* W1 = P + MLP_W1;
  const float* b1 = P + MLP_B1;
  for (int k = 0; k < MLP_H; ++k) {
    float a = b1[k];
    for (int c = 0; c < MLP_IN; ++c) a = fmaf(W1[k * MLP_IN + c], x[c], a);
    const float h = tanh_fast(a);
    H1[(long)r * MLP_H + k] = h;
    H1b[(long)r * MLP_H + k] = (__bf16)h;
  }
}

// layer 3, loss and the output-side backward: y = W3 h2 + b3, e = y - t,
// dy = 2 e / ntot, dH2 = dy W3 (1 - h2^2) (f32 and bf16).  Per-block partial
// sums of e^2 (one float per block).
__global__ __launch_bounds__(256) void mlp_layer3_kernel(int n, const float* H2, const float* P, const float* T,
                                                         const int* idx, float inv_ntot, float* Y, float* dY,
                                                         float* dH2, __bf16* dH2b, float* loss_partial) {
  __shared__ float red[256];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  float e2 = 0.f;
  if (r < n) {
    const float* W3 = P + MLP_W3;
    float y = P[MLP_B3];
    for (int k = 0; k < MLP_H; ++k) y = fmaf(W3[k], H2[(long)r * MLP_H + k], y);
    if (Y) Y[r] = y;
    if (T) {
      const float e = y - T[idx ? idx[r] : r];
      e2 = e * e;
      const float dy = 2.f * e * inv_ntot;
      dY[r] = dy;
      for (int k = 0; k < MLP_H; ++k) {
        const float h = H2[(long)r * MLP_H + k];
        const float d = dy * W3[k] * (1.f - h * h);
        dH2[(long)r * MLP_H + k] = d;
        dH2b[(long)r * MLP_H + k] = (__bf16)d;
      }
    }
  }
  if (!loss_partial) return;
  red[threadIdx.x] = e2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_partial[blockIdx.x] = red[0];
}

// dH1 = dH1pre * (1 - h1^2)
__global__ void mlp_dh1_kernel(long total, const float* dH1pre, const float* H1, float* dH1) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const float h = H1[e];
  dH1[e] = dH1pre[e] * (1.f - h * h);
}

// Per-block column partial sums for the FP32 parameter gradients:
//   dW1[k][c] = sum_r dH1[r][k] x[r][c], db1[k] = sum_r dH1[r][k],
//   db2[k] = sum_r dH2[r][k], dW3[k] = sum_r dY[r] H2[r][k], db3 = sum_r dY[r].
// One thread per hidden unit k (64 threads), rows [blockIdx.x*rows, +rows).
__global__ void mlp_colsum_kernel(int n, int rows, const float* X, const int* idx, const float* dH1, const float* dH2,
                                  const float* H2, const float* dY, float* part) {
  const int k = threadIdx.x;
  const int r0 = blockIdx.x * rows, r1 = min(n, r0 + rows);
  float w1[MLP_IN] = {0, 0, 0, 0, 0};
  float b1 = 0.f, b2 = 0.f, w3 = 0.f, b3 = 0.f;
  for (int r = r0; r < r1; ++r) {
    const int src = idx ? idx[r] : r;
    const float d1 = dH1[(long)r * MLP_H + k];
    for (int c = 0; c < MLP_IN; ++c) w1[c] = fmaf(d1, X[(long)src * MLP_IN + c], w1[c]);
    b1 += d1;
    b2 += dH2[(long)r * MLP_H + k];
    w3 = fmaf(dY[r], H2[(long)r * MLP_H + k], w3);
    if (k == 0) b3 += dY[r];
  }
  float* p = part + (long)blockIdx.x * MLP_NPARAM;
  for (int c = 0; c < MLP_IN; ++c) p[MLP_W1 + k * MLP_IN + c] = w1[c];
  p[MLP_B1 + k] = b1;
  p[MLP_B2 + k] = b2;
  p[MLP_W3 + k] = w3;
  if (k == 0) p[MLP_B3] = b3;
}

// grad[p] = sum over partial blocks (fixed order); W2 part from the split-K
// GEMM partials.
__global__ void mlp_grad_reduce_kernel(int nblk, const float* part, int nsplit, const float* w2part, float* grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_NPARAM) return;
  float s = 0.f;
  if (p >= MLP_W2 && p < MLP_B2) {
    for (int z = 0; z < nsplit; ++z) s += w2part[(long)z * MLP_H * MLP_H + (p - MLP_W2)];
  } else {
    for (int b = 0; b < nblk; ++b) s += part[(long)b * MLP_NPARAM + p];
  }
  grad[p] = s;
}

__global__ void mlp_sum_kernel(int n, const float* v, float scale, float* out) {
  __shared__ float red[256];
  float s = 0.f;
  for (int i = threadIdx.x; i < n; i += 256) s += v[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0] * scale;
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters, then the BF16 copy of
// W2 used by the matrix cores.
__global__ void mlp_adam_kernel(float* P, float* m, float* v, const float* g, float lr, float b1, float b2, float eps,
                                float c1, float c2, __bf16* W2b) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_NPARAM) return;
  const float gp = g[p];
  const float mp = b1 * m[p] + (1.f - b1) * gp;
  const float vp = b2 * v[p] + (1.f - b2) * gp * gp;
  m[p] = mp;
  v[p] = vp;
  const float w = P[p] - lr * (mp / c1) / (sqrtf(vp / c2) + eps);
  P[p] = w;
  if (p >= MLP_W2 && p < MLP_B2) W2b[p - MLP_W2] = (__bf16)w;
}

__global__ void mlp_w2_to_bf16_kernel(const float* P, __bf16* W2b) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < MLP_H * MLP_H) W2b[e] = (__bf16)P[MLP_W2 + e];
}

// ---- host-side drivers ------------------------------------------------------
hipError_t mlp_w2_to_bf16(const float* P, __bf16* W2b, hipStream_t s) {
  hipLaunchKernelGGL(mlp_w2_to_bf16_kernel, dim3((MLP_H * MLP_H + 255) / 256), dim3(256), 0, s, P, W2b);
  return hipGetLastError();
}

// Forward (and, with T, the loss and the full gradient) on n rows.
hipError_t mlp_forward_backward(const MlpWork& w, int n, const float* X, const int* idx, const float* T, int ntot,
                                float* Y, float* grad, float* loss, hipStream_t s) {
  const int nb = (n + 255) / 256;
  hipLaunchKernelGGL(mlp_layer1_kernel, dim3(nb), dim3(256), 0, s, n, X, idx, w.P, w.H1, w.H1b);
  // H2 = tanh(H1b W2b^T + b2): A = H1b [n][64], B[k][j] = W2b[j][k]
  gemm_bf16(n, MLP_H, MLP_H, w.H1b, MLP_H, 1, w.W2b, 1, MLP_H, w.H2, MLP_H, 1, 1, 0, 1, w.P + MLP_B2, w.H2b, s);
  hipLaunchKernelGGL(mlp_layer3_kernel, dim3(nb), dim3(256), 0, s, n, w.H2, w.P, T, idx, 1.0f / (float)ntot, Y, w.dY,
                     w.dH2, w.dH2b, T ? w.loss_part : nullptr);
  if (!T) return hipGetLastError();
  hipLaunchKernelGGL(mlp_sum_kernel, dim3(1), dim3(256), 0, s, nb, w.loss_part, 1.0f / (float)n, loss);
  // dW2[j][k] = sum_r dH2[r][j] H1[r][k]  (split-K over rows)
  const int nsplit = w.nsplit;
  gemm_bf16(MLP_H, MLP_H, n, w.dH2b, 1, MLP_H, w.H1b, MLP_H, 1, w.w2part, MLP_H, 1, nsplit, (long)MLP_H * MLP_H, 0,
            nullptr, nullptr, s);
  // dH1pre[r][k] = sum_j dH2[r][j] W2[j][k]
  gemm_bf16(n, MLP_H, MLP_H, w.dH2b, MLP_H, 1, w.W2b, MLP_H, 1, w.dH1, MLP_H, 1, 1, 0, 0, nullptr, nullptr, s);
  const long tot = (long)n * MLP_H;
  hipLaunchKernelGGL(mlp_dh1_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, tot, w.dH1, w.H1, w.dH1);
  const int rows = 512, ncb = (n + rows - 1) / rows;
  hipLaunchKernelGGL(mlp_colsum_kernel, dim3(ncb), dim3(MLP_H), 0, s, n, rows, X, idx, w.dH1, w.dH2, w.H2, w.dY,
                     w.colpart);
  hipLaunchKernelGGL(mlp_grad_reduce_kernel, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, ncb, w.colpart, nsplit,
                     w.w2part, grad);
  return hipGetLastError();
}

hipError_t mlp_adam(const MlpWork& w, const float* grad, float lr, int step, hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, w.P, w.m, w.v, grad, lr, b1,
                     b2, eps, c1, c2, w.W2b);
  return hipGetLastError();
}

}  // namespace sg
