// Host-side context behind the C ABI: geometry (computed on the host in
// FP64 like Mesh::calc_metrics), the compiled boundary list, and the device
// buffers of one grid.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/structured_gpu.h"
#include "sg_device.hpp"

namespace sg {

// Device allocation owned by the context.
struct DevBuf {
  double* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  hipError_t alloc(size_t count) {
    release();
    n = count;
    if (count == 0) return hipSuccess;
    hipError_t e = hipMalloc(&p, count * sizeof(double));
    if (e != hipSuccess) { p = nullptr; n = 0; }
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// Page-locked host memory (hipHostMalloc), for the small per-step reads of
// the Krylov loop: a copy into it is a DMA the stream orders, where a copy
// into pageable memory goes through a staging buffer.
struct PinnedBuf {
  double* p = nullptr;
  size_t n = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() { release(); }
  hipError_t alloc(size_t count) {
    release();
    if (count == 0) return hipSuccess;
    hipError_t e = hipHostMalloc((void**)&p, count * sizeof(double), hipHostMallocDefault);
    if (e != hipSuccess) p = nullptr;
    else n = count;
    return e;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
};

// Mesh::calc_metrics, src/utils/mesh.cpp:171-205, reference layouts.
struct HostGeom {
  int ni = 0, nj = 0, nic = 0, njc = 0;
  std::vector<double> xv, yv;              // [ni][nj]
  std::vector<double> normal_chi;          // [ni][njc][2]
  std::vector<double> normal_eta;          // [nic][nj][2]
  std::vector<double> volume, xc, yc, ds_chi, ds_eta;  // [nic][njc]
  void build(int ni_, int nj_, const double* x, const double* y);
};

// Ordered boundary list -> evaluation nodes per ghost cell (see DevBc).
struct BcCompiled {
  std::vector<BcNode> nodes;
  std::vector<int> ghost_map;
  std::string err;
  // every ghost cell depends only on interior cells within one cell
  // tangentially and two normally (so every Jacobian coupling stays inside
  // the 13-point stencil); false for wake / wide periodic boundaries
  bool local = true;
  // the same inside the 3x3 box (so the first-order Jacobian, whose interior
  // couplings are the box, stays inside the box)
  bool local_box = true;
  bool compile(const HostGeom& g, const sg_bc* bcs, int nbc, bool viscous);
};

}  // namespace sg

struct SolverState;

struct sg_ctx {
  sg::HostGeom geo;
  sg::BcCompiled bcc;
  sg::DevGeom g{};
  sg::DevPhys ph{};
  sg::DevBc bc{};
  sg_freestream fs{};
  sg_opts opts{};
  int nvar = 4;
  bool sa = false;
  hipStream_t stream = nullptr;
  std::string err;
  // device buffers
  sg::DevBuf ncx, ncy, nex, ney, vol, dsmin, wd, beta, xv0, yv0;
  sg::DevBuf nodes_buf, map_buf;   // BcNode / int storage (raw doubles)
  sg::DevBuf aos_a, aos_b;         // AoS staging (nvar * ncell)
  sg::DevBuf q, r, v, w;           // SoA work vectors (nvar * plane)
  sg::DevBuf partial, l2, counter;  // norm partials; counter: the arrival count, 0 between launches
  unsigned norm_parity = 0;         // parity of the next fused norm on `partial`
  sg::DevBuf wall_cp, wall_cf, wall_f;
  sg::DevClosure clo;              // beta closure applied by every residual / J.v launch
  sg::DevBuf clo_params, clo_w2;   // MLP parameters (FP32) and W2 (BF16), raw storage
  sg::DevBuf aux;                  // features (5 planes) + beta, for sg_closure_eval
  sg::DevBuf ex_a, ex_b, ex_r;     // explicit stepping: stage states and residual
  SolverState* solver = nullptr;   // Newton-Krylov work space (sg_solver.cpp)
  double* noalloc_scratch = nullptr;  // this device's static scratch (sg_noalloc.hip)
  size_t ncell() const { return (size_t)geo.nic * geo.njc; }
  size_t nsoa() const { return (size_t)nvar * g.plane; }
};

int sg_fail(sg_ctx* ctx, int code, const std::string& msg);
int sg_noalloc_setup(sg_ctx* ctx);
void sg_solver_free(sg_ctx* ctx);
int sg_hip(sg_ctx* ctx, hipError_t e, const char* what);
