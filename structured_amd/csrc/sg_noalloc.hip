// Residual entry point that allocates nothing, for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops).
//
// The fused residual's scratch -- the per-block partial sums of the L2 norms
// and the arrival counters of the last-block finalisation -- lives in a
// file-scope device array instead of context buffers.  Each GPU that the
// library runs on holds its own copy of the array until the process ends; the
// context looks its address up once per device at creation
// (hipGetSymbolAddress) and passes that address to the kernels.  Calls that
// overlap on one GPU share the array and race: serialise them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "sg_context.hpp"
#include "sg_kernels.hpp"

namespace {

// Layout: [0, 160) doubles hold the fused norm's arrival counters
// (sg::kNormStateWords 64-bit words; the rest keeps the partials 256-byte
// aligned); then the partials, 2 nvar per residual tile, component-major.
constexpr size_t kCounterDoubles = 160;
static_assert(kCounterDoubles >= (size_t)sg::kNormStateWords && kCounterDoubles % 32 == 0, "counter area");
constexpr size_t kScratchDoubles = kCounterDoubles + (size_t)SG_NOALLOC_MAX_TILES * 10;

static __device__ __align__(256) double scratch[kScratchDoubles];

// Zeroes the counter and the n partials the next launch uses: another call
// may have left them from a grid with another tile count.  Each call then
// starts at parity 0.
__global__ void clear_scratch_kernel(double* s, size_t n) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
    s[k] = 0.0;
}

std::mutex g_mu;
double* g_ptr[64] = {nullptr};  // scratch's device address, per device ordinal

}  // namespace

// Device address of this GPU's scratch, looked up once per device with that
// device current.
int sg_noalloc_setup(sg_ctx* ctx) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return sg_fail(ctx, SG_ERR_HIP, "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return sg_fail(ctx, SG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_ptr[dev] = static_cast<double*>(p);
  }
  ctx->noalloc_scratch = g_ptr[dev];
  return SG_OK;
}

extern "C" int sg_residual_dev_async_noalloc(sg_ctx* ctx, const double* q_dev, double* r_dev, double* l2_dev) {
  if (!ctx || !q_dev || !r_dev) return SG_ERR_ARG;
  if (!ctx->noalloc_scratch) return sg_fail(ctx, SG_ERR_STATE, "no static scratch on this device");
  const size_t tiles = sg::residual_partial_count(ctx->g);
  if (tiles > (size_t)SG_NOALLOC_MAX_TILES)
    return sg_fail(ctx, SG_ERR_ARG,
                   "grid too large for the static scratch: " + std::to_string(tiles) + " residual tiles, at most " +
                       std::to_string(SG_NOALLOC_MAX_TILES));
  double* partial = nullptr;
  unsigned long long* state = nullptr;
  if (l2_dev) {
    state = reinterpret_cast<unsigned long long*>(ctx->noalloc_scratch);
    partial = ctx->noalloc_scratch + kCounterDoubles;
    // counter and partials must start at zero (parity 0); a kernel, not a
    // memset, so that the call can be captured into a graph
    const size_t n = kCounterDoubles + tiles * 2 * (size_t)ctx->nvar;
    hipLaunchKernelGGL(clear_scratch_kernel, dim3((unsigned)std::min<size_t>((n + 255) / 256, 1024)), dim3(256), 0,
                       ctx->stream, ctx->noalloc_scratch, n);
    if (hipGetLastError() != hipSuccess) return sg_fail(ctx, SG_ERR_HIP, "scratch clear launch");
  }
  unsigned parity = 0;
  const hipError_t e = sg::launch_residual(ctx->g, ctx->ph, ctx->bc, ctx->sa, q_dev, r_dev, partial, state, l2_dev,
                                           ctx->stream, ctx->clo, ctx->nvar, &parity);
  if (e != hipSuccess) return sg_fail(ctx, SG_ERR_HIP, std::string("residual launch: ") + hipGetErrorString(e));
  return SG_OK;
}
