// C ABI of the closure network: creation (seeded Glorot-uniform weights),
// parameter I/O, batched inference, loss + gradient, and Adam training over
// device-resident samples.  Kernels are in sg_mlp.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "sg_context.hpp"
#include "sg_mlp.hpp"

using namespace sg;

struct sg_mlp {
  MlpWork w;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<void*> allocs;
  int step = 0;                 // Adam step counter (bias correction)
  float* Y = nullptr;           // forward output staging [cap]
  float* Xd = nullptr;          // staged inputs [cap][5] / targets [cap]
  float* Td = nullptr;
};

namespace {

int mlp_fail(sg_mlp* n, int code, const std::string& msg) {
  if (n) n->err = msg;
  return code;
}

#define MLP_TRY(expr, what)                                                            \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) return mlp_fail(net, SG_ERR_HIP, std::string(what) + ": " + \
                                                                  hipGetErrorString(e_)); \
  } while (0)

template <class T>
hipError_t dalloc(sg_mlp* net, T*& p, size_t count) {
  void* v = nullptr;
  hipError_t e = hipMalloc(&v, std::max<size_t>(count, 1) * sizeof(T));
  if (e == hipSuccess) {
    net->allocs.push_back(v);
    p = static_cast<T*>(v);
  }
  return e;
}

// splitmix64 (Steele, Lea & Flood 2014): seeded, platform-independent stream.
struct SplitMix {
  unsigned long long s;
  unsigned long long next() {
    unsigned long long z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

void glorot(std::vector<float>& p, unsigned long long seed) {
  SplitMix r{seed};
  std::fill(p.begin(), p.end(), 0.0f);
  auto fill = [&](int off, int fan_out, int fan_in) {
    const double lim = std::sqrt(6.0 / (fan_in + fan_out));
    for (int k = 0; k < fan_out * fan_in; ++k) p[off + k] = (float)((2.0 * r.uniform() - 1.0) * lim);
  };
  fill(MLP_W1, MLP_H, MLP_IN);
  fill(MLP_W2, MLP_H, MLP_H);
  fill(MLP_W3, 1, MLP_H);
}

}  // namespace

int sg_mlp_create(sg_mlp** out, int max_rows, unsigned long long seed) {
  if (!out || max_rows <= 0) return SG_ERR_ARG;
  *out = nullptr;
  sg_mlp* net = new sg_mlp();
  auto fail = [&](int code) {
    sg_mlp_destroy(net);
    return code;
  };
  if (hipStreamCreate(&net->stream) != hipSuccess) return fail(SG_ERR_HIP);
  MlpWork& w = net->w;
  const size_t cap = (size_t)max_rows, ch = cap * MLP_H;
  w.cap = max_rows;
  w.nsplit = (int)std::min<size_t>(64, std::max<size_t>(1, cap / 2048));
  const int ncb = (max_rows + 511) / 512;
  if (dalloc(net, w.P, MLP_NPARAM) || dalloc(net, w.m, MLP_NPARAM) || dalloc(net, w.v, MLP_NPARAM) ||
      dalloc(net, w.grad, MLP_NPARAM) || dalloc(net, w.W2b, MLP_H * MLP_H) || dalloc(net, w.H1, ch) ||
      dalloc(net, w.H2, ch) || dalloc(net, w.dH1, ch) || dalloc(net, w.dH2, ch) || dalloc(net, w.dY, cap) ||
      dalloc(net, w.H1b, ch) || dalloc(net, w.H2b, ch) || dalloc(net, w.dH2b, ch) ||
      dalloc(net, w.loss_part, (cap + 255) / 256) || dalloc(net, w.colpart, (size_t)ncb * MLP_NPARAM) ||
      dalloc(net, w.w2part, (size_t)w.nsplit * MLP_H * MLP_H) || dalloc(net, w.loss, 1) || dalloc(net, net->Y, cap) ||
      dalloc(net, net->Xd, cap * MLP_IN) || dalloc(net, net->Td, cap))
    return fail(SG_ERR_HIP);
  std::vector<float> p(MLP_NPARAM);
  glorot(p, seed);
  int rc = sg_mlp_set_params(net, p.data());
  if (rc) return fail(rc);
  *out = net;
  return SG_OK;
}

void sg_mlp_destroy(sg_mlp* net) {
  if (!net) return;
  if (net->stream) (void)hipStreamSynchronize(net->stream);
  for (void* p : net->allocs) (void)hipFree(p);
  if (net->stream) (void)hipStreamDestroy(net->stream);
  delete net;
}

const char* sg_mlp_last_error(const sg_mlp* net) { return net ? net->err.c_str() : "null network"; }

// New parameters also reset the Adam state.
int sg_mlp_set_params(sg_mlp* net, const float* params) {
  if (!net || !params) return SG_ERR_ARG;
  MlpWork& w = net->w;
  MLP_TRY(hipMemcpyAsync(w.P, params, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice, net->stream), "H2D");
  MLP_TRY(hipMemsetAsync(w.m, 0, MLP_NPARAM * sizeof(float), net->stream), "memset");
  MLP_TRY(hipMemsetAsync(w.v, 0, MLP_NPARAM * sizeof(float), net->stream), "memset");
  MLP_TRY(mlp_w2_to_bf16(w.P, w.W2b, net->stream), "w2 bf16");
  MLP_TRY(hipStreamSynchronize(net->stream), "set params");
  net->step = 0;
  return SG_OK;
}

int sg_mlp_get_params(sg_mlp* net, float* params) {
  if (!net || !params) return SG_ERR_ARG;
  MLP_TRY(hipMemcpyAsync(params, net->w.P, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, net->stream), "D2H");
  MLP_TRY(hipStreamSynchronize(net->stream), "get params");
  return SG_OK;
}

int sg_mlp_forward(sg_mlp* net, int n, const float* x, float* y) {
  if (!net || n < 0 || (n > 0 && (!x || !y))) return SG_ERR_ARG;
  const int cap = net->w.cap;
  for (int r0 = 0; r0 < n; r0 += cap) {
    const int m = std::min(cap, n - r0);
    MLP_TRY(hipMemcpyAsync(net->Xd, x + (size_t)r0 * MLP_IN, (size_t)m * MLP_IN * sizeof(float),
                           hipMemcpyHostToDevice, net->stream), "H2D");
    MLP_TRY(mlp_forward_backward(net->w, m, net->Xd, nullptr, nullptr, m, net->Y, nullptr, nullptr, net->stream),
            "forward");
    MLP_TRY(hipMemcpyAsync(y + r0, net->Y, (size_t)m * sizeof(float), hipMemcpyDeviceToHost, net->stream), "D2H");
  }
  MLP_TRY(hipStreamSynchronize(net->stream), "forward");
  return SG_OK;
}

int sg_mlp_loss_grad(sg_mlp* net, int n, const float* x, const float* t, float* loss, float* grad) {
  if (!net || n <= 0 || !x || !t) return SG_ERR_ARG;
  if (n > net->w.cap) return mlp_fail(net, SG_ERR_ARG, "more rows than the network was created for");
  MLP_TRY(hipMemcpyAsync(net->Xd, x, (size_t)n * MLP_IN * sizeof(float), hipMemcpyHostToDevice, net->stream), "H2D");
  MLP_TRY(hipMemcpyAsync(net->Td, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice, net->stream), "H2D");
  MLP_TRY(mlp_forward_backward(net->w, n, net->Xd, nullptr, net->Td, n, nullptr, net->w.grad, net->w.loss,
                               net->stream), "loss/grad");
  float l = 0.f;
  MLP_TRY(hipMemcpyAsync(&l, net->w.loss, sizeof(float), hipMemcpyDeviceToHost, net->stream), "D2H");
  if (grad)
    MLP_TRY(hipMemcpyAsync(grad, net->w.grad, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, net->stream), "D2H");
  MLP_TRY(hipStreamSynchronize(net->stream), "loss/grad");
  if (loss) *loss = l;
  return SG_OK;
}

int sg_mlp_train(sg_mlp* net, int n, const float* x, const float* t, int epochs, int batch, float lr,
                 unsigned long long seed, float* loss_hist) {
  if (!net || n <= 0 || !x || !t || epochs < 0 || batch <= 0 || !(lr > 0.f)) return SG_ERR_ARG;
  const int bs = std::min(batch, n);
  if (bs > net->w.cap) return mlp_fail(net, SG_ERR_ARG, "minibatch larger than the network was created for");
  const int nb = n / bs;
  float *X = nullptr, *T = nullptr, *L = nullptr;
  int* idx = nullptr;
  std::vector<void*> tmp;
  auto cleanup = [&]() {
    for (void* p : tmp) (void)hipFree(p);
  };
  auto alloc = [&](auto*& p, size_t count) {
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, std::max<size_t>(count, 1) * sizeof(*p));
    if (e == hipSuccess) {
      tmp.push_back(v);
      p = static_cast<std::remove_reference_t<decltype(p)>>(v);
    }
    return e;
  };
  if (alloc(X, (size_t)n * MLP_IN) || alloc(T, (size_t)n) || alloc(idx, (size_t)n) ||
      alloc(L, (size_t)std::max(1, epochs) * nb)) {
    cleanup();
    return mlp_fail(net, SG_ERR_HIP, "hipMalloc of the training set failed");
  }
  int rc = SG_OK;
  auto run = [&]() -> int {
    MLP_TRY(hipMemcpyAsync(X, x, (size_t)n * MLP_IN * sizeof(float), hipMemcpyHostToDevice, net->stream), "H2D");
    MLP_TRY(hipMemcpyAsync(T, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice, net->stream), "H2D");
    std::vector<int> perm(n);
    SplitMix r{seed};
    for (int e = 0; e < epochs; ++e) {
      for (int k = 0; k < n; ++k) perm[k] = k;
      for (int k = n - 1; k > 0; --k) std::swap(perm[k], perm[(int)(r.next() % (unsigned long long)(k + 1))]);
      // the previous epoch's kernels read idx: order the upload after them
      MLP_TRY(hipMemcpyAsync(idx, perm.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, net->stream), "H2D");
      MLP_TRY(hipStreamSynchronize(net->stream), "epoch");
      for (int b = 0; b < nb; ++b) {
        MLP_TRY(mlp_forward_backward(net->w, bs, X, idx + (size_t)b * bs, T, bs, nullptr, net->w.grad,
                                     L + (size_t)e * nb + b, net->stream), "forward/backward");
        MLP_TRY(mlp_adam(net->w, net->w.grad, lr, ++net->step, net->stream), "adam");
      }
    }
    std::vector<float> lh((size_t)std::max(1, epochs) * nb);
    MLP_TRY(hipMemcpyAsync(lh.data(), L, lh.size() * sizeof(float), hipMemcpyDeviceToHost, net->stream), "D2H");
    MLP_TRY(hipStreamSynchronize(net->stream), "train");
    if (loss_hist)
      for (int e = 0; e < epochs; ++e) {
        double s = 0.0;
        for (int b = 0; b < nb; ++b) s += lh[(size_t)e * nb + b];
        loss_hist[e] = (float)(s / nb);
      }
    return SG_OK;
  };
  rc = run();
  cleanup();
  return rc;
}

int sg_set_closure_mlp(sg_ctx* ctx, const sg_mlp* cnet) {
  if (!ctx || !cnet) return SG_ERR_ARG;
  sg_mlp* net = const_cast<sg_mlp*>(cnet);
  MLP_TRY(hipStreamSynchronize(net->stream), "network");
  std::vector<float> p(MLP_NPARAM);
  MLP_TRY(hipMemcpy(p.data(), net->w.P, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost), "D2H");
  return sg_set_closure(ctx, SG_CLOSURE_MLP, nullptr, p.data());
}
