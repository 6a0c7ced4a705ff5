// Closures for the SA production multiplier beta, evaluated inside the
// residual kernel: local flow features, the analytic FP64 closure, and the
// learned one (5 -> 64 -> 64 -> 1 tanh MLP with the 64 x 64 layer on the
// matrix cores).  All are templated on the scalar type, so the J.v
// instantiation carries the closure's derivative.
#pragma once
#include "sg_device.hpp"
#include "sg_scalar.hpp"

namespace sg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kNFeat = 5;
constexpr int kMlpH = 64;
constexpr int kHld = kMlpH + 8;  // LDS row stride (bf16) of the staged hidden layer

// tanh x = 1 - 2 / (1 + e^(2x)) from the hardware exp2 / reciprocal (absolute
// error ~1e-7; saturates to +-1 for large |x|).  Used by every MLP kernel.
__device__ __forceinline__ float tanh_fast(float x) {
  const float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + e);
}

__device__ __forceinline__ double slog1p(double x) { return log1p(x); }
__device__ __forceinline__ Dual slog1p(Dual x) { return Dual(log1p(x.v), x.d * frcp(1.0 + x.v)); }
__device__ __forceinline__ double sasinh(double x) { return asinh(x); }
__device__ __forceinline__ Dual sasinh(Dual x) { return Dual(asinh(x.v), x.d * frcp(sqrt(1.0 + x.v * x.v))); }
__device__ __forceinline__ double stanh(double x) { return tanh(x); }
__device__ __forceinline__ Dual stanh(Dual x) {
  const double t = tanh(x.v);
  return Dual(t, (1.0 - t * t) * x.d);
}

// Local features of one cell (all O(1), smooth where the flow is):
//   f0 = |W| / (|W| + |S|)                 vorticity / strain ratio
//   f1 = log(1 + d^2 |W| / nu) / 10        wall-distance Reynolds number
//   f2 = asinh(nu~ / nu) / 8               eddy-viscosity ratio
//   f3 = tanh(d (u . grad p) / (|u| q))    streamwise pressure-gradient indicator
//   f4 = P / (|P| + |D| + eps)             SA production / destruction balance
// |S| = sqrt(2 S_ij S_ij + eps_s^2); |u| and the dynamic pressure q carry
// small freestream-scaled floors so that the features stay finite at walls.
template <class S>
__device__ __forceinline__ void sa_features(S r, S u, S v, S n, S nu, const S* gu, const S* gv, const S* gp,
                                            double d, S P1, S D, const DevPhys& ph, S* f) {
  const double U2 = ph.u_inf * ph.u_inf + ph.v_inf * ph.v_inf;
  const double U = sqrt(U2);
  S om = sabs(gv[0] - gu[1]);
  S s12 = gu[1] + gv[0];
  S smag = ssqrt(2.0 * (gu[0] * gu[0] + gv[1] * gv[1]) + s12 * s12 + 1e-12 * U2);
  f[0] = sdiv(om, om + smag);
  f[1] = slog1p(sdiv(om * (d * d), nu)) * 0.1;
  f[2] = sasinh(sdiv(n, nu)) * 0.125;
  S uu = u * u + v * v;
  S umag = ssqrt(uu + 1e-6 * U2);
  S qd = 0.5 * r * uu + 0.025 * ph.rho_inf * U2;
  f[3] = stanh(sdiv(d * (u * gp[0] + v * gp[1]), umag * qd));
  f[4] = sdiv(P1, sabs(P1) + sabs(D) + 1e-8 * ph.rho_inf * U2 * U);
}

template <class S>
__device__ __forceinline__ S closure_analytic(const DevClosure& c, const S* f) {
  S a = S(c.coef[5]);
  for (int k = 0; k < kNFeat; ++k) a = a + c.coef[k] * f[k];
  return 1.0 + c.coef[6] * stanh(a);
}

// MLP inference for the 64 cells of one wavefront (one cell per lane; every
// lane of the block must call it, the block synchronises once inside).
// hs / ht: this wave's 64 x kHld BF16 LDS rows for the hidden layer and its
// tangent.  Returns y = MLP(x) and, with TAN, its directional derivative
// yt along xt.  Layer 1 and 3 are FP32 FMAs; layer 2 is 32 (64 with TAN)
// v_mfma_f32_16x16x32_bf16 per wave.
template <bool TAN>
__device__ __forceinline__ void mlp_wave(const float* __restrict__ P, const __bf16* __restrict__ W2, const float* x,
                                         const float* xt, __bf16* hs, __bf16* ht, float& y, float& yt) {
  const int lane = threadIdx.x & 63;
  const float* W1 = P;                       // [64][5]
  const float* b1 = P + kMlpH * kNFeat;      // [64]
  const float* b2 = b1 + kMlpH + kMlpH * kMlpH;
  const float* W3 = b2 + kMlpH;
  const float b3 = W3[kMlpH];
#pragma unroll
  for (int k0 = 0; k0 < kMlpH; k0 += 8) {
    bf16x8 hv, tv;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int k = k0 + t;
      float a = b1[k];
#pragma unroll
      for (int c = 0; c < kNFeat; ++c) a = fmaf(W1[k * kNFeat + c], x[c], a);
      const float h = tanh_fast(a);
      hv[t] = (__bf16)h;
      if (TAN) {
        float at = 0.f;
#pragma unroll
        for (int c = 0; c < kNFeat; ++c) at = fmaf(W1[k * kNFeat + c], xt[c], at);
        tv[t] = (__bf16)((1.f - h * h) * at);
      }
    }
    *reinterpret_cast<bf16x8*>(hs + lane * kHld + k0) = hv;
    if (TAN) *reinterpret_cast<bf16x8*>(ht + lane * kHld + k0) = tv;
  }
  __syncthreads();
  const int r16 = lane & 15, kq = 8 * (lane >> 4);
  float yp[4] = {0.f, 0.f, 0.f, 0.f}, tp[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const bf16x8 w0 = *reinterpret_cast<const bf16x8*>(W2 + (16 * a + r16) * kMlpH + kq);
    const bf16x8 w1 = *reinterpret_cast<const bf16x8*>(W2 + (16 * a + r16) * kMlpH + 32 + kq);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const __bf16* row = hs + (16 * b + r16) * kHld + kq;
      f32x4 c = {0.f, 0.f, 0.f, 0.f};
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w0, *reinterpret_cast<const bf16x8*>(row), c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, *reinterpret_cast<const bf16x8*>(row + 32), c, 0, 0, 0);
      f32x4 ct = {0.f, 0.f, 0.f, 0.f};
      if (TAN) {
        const __bf16* trow = ht + (16 * b + r16) * kHld + kq;
        ct = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w0, *reinterpret_cast<const bf16x8*>(trow), ct, 0, 0, 0);
        ct = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, *reinterpret_cast<const bf16x8*>(trow + 32), ct, 0, 0, 0);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int j = 16 * a + 4 * (lane >> 4) + i;
        const float h2 = tanh_fast(c[i] + b2[j]);
        yp[b] = fmaf(W3[j], h2, yp[b]);
        if (TAN) tp[b] = fmaf(W3[j] * (1.f - h2 * h2), ct[i], tp[b]);
      }
    }
  }
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    yp[b] += __shfl_xor(yp[b], 16);
    yp[b] += __shfl_xor(yp[b], 32);
    if (TAN) {
      tp[b] += __shfl_xor(tp[b], 16);
      tp[b] += __shfl_xor(tp[b], 32);
    }
  }
  const int mb = lane >> 4;
  const float ys = mb == 0 ? yp[0] : mb == 1 ? yp[1] : mb == 2 ? yp[2] : yp[3];
  y = b3 + ys;
  yt = TAN ? (mb == 0 ? tp[0] : mb == 1 ? tp[1] : mb == 2 ? tp[2] : tp[3]) : 0.f;
}

}  // namespace sg
