// C ABI (include/structured_gpu.h): setup, residual, tangent, time step and
// wall quantities on host or device buffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include "sg_context.hpp"
#include "sg_closure.hpp"
#include "sg_kernels.hpp"
#include "sg_mlp.hpp"

using namespace sg;

static thread_local std::string g_create_err;  // errors of sg_create (no context yet)

int sg_fail(sg_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  else g_create_err = msg;
  return code;
}

int sg_hip(sg_ctx* ctx, hipError_t e, const char* what) {
  if (e == hipSuccess) return SG_OK;
  return sg_fail(ctx, SG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define SG_TRY(expr, what)                         \
  do {                                             \
    int rc_ = sg_hip(ctx, (expr), what);           \
    if (rc_ != SG_OK) return rc_;                  \
  } while (0)

namespace sg {

void HostGeom::build(int ni_, int nj_, const double* x, const double* y) {
  ni = ni_; nj = nj_; nic = ni - 1; njc = nj - 1;
  xv.assign(x, x + (size_t)ni * nj);
  yv.assign(y, y + (size_t)ni * nj);
  auto X = [&](int i, int j) { return xv[(size_t)i * nj + j]; };
  auto Y = [&](int i, int j) { return yv[(size_t)i * nj + j]; };
  normal_chi.assign((size_t)ni * njc * 2, 0.0);
  normal_eta.assign((size_t)nic * nj * 2, 0.0);
  std::vector<double> reta((size_t)nic * nj * 2), rchi((size_t)ni * njc * 2);
  for (int i = 0; i < nic; ++i)
    for (int j = 0; j < nj; ++j) {
      size_t k = ((size_t)i * nj + j) * 2;
      reta[k] = X(i + 1, j) - X(i, j);
      reta[k + 1] = Y(i + 1, j) - Y(i, j);
      normal_eta[k] = -reta[k + 1];
      normal_eta[k + 1] = reta[k];
    }
  for (int i = 0; i < ni; ++i)
    for (int j = 0; j < njc; ++j) {
      size_t k = ((size_t)i * njc + j) * 2;
      rchi[k] = X(i, j + 1) - X(i, j);
      rchi[k + 1] = Y(i, j + 1) - Y(i, j);
      normal_chi[k] = rchi[k + 1];
      normal_chi[k + 1] = -rchi[k];
    }
  const size_t nc = (size_t)nic * njc;
  volume.assign(nc, 0.0); xc.assign(nc, 0.0); yc.assign(nc, 0.0); ds_chi.assign(nc, 0.0); ds_eta.assign(nc, 0.0);
  auto RE = [&](int i, int j, int k) { return reta[((size_t)i * nj + j) * 2 + k]; };
  auto RC = [&](int i, int j, int k) { return rchi[((size_t)i * njc + j) * 2 + k]; };
  for (int i = 0; i < nic; ++i)
    for (int j = 0; j < njc; ++j) {
      size_t c = (size_t)i * njc + j;
      volume[c] = 0.5 * (RE(i, j, 0) * RC(i, j, 1) - RC(i, j, 0) * RE(i, j, 1) + RE(i, j + 1, 0) * RC(i + 1, j, 1) -
                         RC(i + 1, j, 0) * RE(i, j + 1, 1));
      xc[c] = 0.25 * (X(i, j) + X(i + 1, j) + X(i, j + 1) + X(i + 1, j + 1));
      yc[c] = 0.25 * (Y(i, j) + Y(i + 1, j) + Y(i, j + 1) + Y(i + 1, j + 1));
      const double* ne = &normal_eta[((size_t)i * nj + j) * 2];
      const double* nh = &normal_chi[((size_t)i * njc + j) * 2];
      ds_eta[c] = std::sqrt(ne[0] * ne[0] + ne[1] * ne[1]);
      ds_chi[c] = std::sqrt(nh[0] * nh[0] + nh[1] * nh[1]);
    }
}

namespace {
struct Compiler {
  int nic, njc;
  std::vector<BcNode>& nodes;
  std::vector<int> depth;
  std::vector<int>& cur;
  std::string& err;
  int perim(int gi, int gj) const {
    if (gi < 0 || gi > nic + 1 || gj < 0 || gj > njc + 1) return -1;
    if (gj == 0) return gi;
    if (gj == njc + 1) return (nic + 2) + gi;
    if (gi == 0) return 2 * (nic + 2) + gj - 1;
    if (gi == nic + 1) return 2 * (nic + 2) + njc + gj - 1;
    return -1;
  }
  bool ref(int gi, int gj, int& out) {
    if (gi >= 1 && gi <= nic && gj >= 1 && gj <= njc) { out = (gj - 1) * nic + (gi - 1); return true; }
    int p = perim(gi, gj);
    if (p < 0) { err = "boundary reads outside the ghosted grid at (" + std::to_string(gi) + "," + std::to_string(gj) + ")"; return false; }
    if (cur[p] < 0) {
      err = "boundary list reads ghost cell (" + std::to_string(gi) + "," + std::to_string(gj) +
            ") before any boundary writes it; reorder the list";
      return false;
    }
    out = -(cur[p] + 1);
    return true;
  }
  int dep(int r) const { return r >= 0 ? 0 : depth[-(r + 1)]; }
  bool write(int gi, int gj, BcNode nd, bool uses_a, bool uses_b) {
    int p = perim(gi, gj);
    if (p < 0) { err = "boundary writes outside the ghost ring"; return false; }
    int d = 1 + std::max(uses_a ? dep(nd.ra) : 0, uses_b ? dep(nd.rb) : 0);
    if (d > 2) { err = "boundary chain deeper than two ghost cells"; return false; }
    cur[p] = (int)nodes.size();
    nodes.push_back(nd);
    depth.push_back(d);
    return true;
  }
};
}  // namespace

bool BcCompiled::compile(const HostGeom& g, const sg_bc* bcs, int nbc, bool viscous) {
  const int nic = g.nic, njc = g.njc;
  nodes.clear();
  ghost_map.assign(2 * (nic + 2) + 2 * njc, -1);
  Compiler C{nic, njc, nodes, {}, ghost_map, err};
  auto NE = [&](int i, int j, int k) { return g.normal_eta[((size_t)i * g.nj + j) * 2 + k]; };
  for (int b = 0; b < nbc; ++b) {
    const sg_bc& B = bcs[b];
    const bool horiz = B.face == SG_FACE_BOTTOM || B.face == SG_FACE_TOP;
    if (B.face < 0 || B.face > 3) { err = "bad face"; return false; }
    const long long nlen = horiz ? nic : njc;
    long long start = B.start, end = B.end < 0 ? nlen + 2 + B.end : B.end;  // src/model/bc.cpp:435-457
    if (start < 0 || end > nlen + 1) { err = "boundary range outside the face"; return false; }
    const int jend = B.face == SG_FACE_BOTTOM ? 0 : njc + 1;
    const int iend = B.face == SG_FACE_LEFT ? 0 : nic + 1;
    for (long long s = start; s <= end; ++s) {
      const int t = (int)s;
      BcNode nd{};
      switch (B.type) {
        case SG_BC_FREESTREAM:
          nd.op = OP_FREESTREAM;
          if (!(horiz ? C.write(t, jend, nd, false, false) : C.write(iend, t, nd, false, false))) return false;
          break;
        case SG_BC_SLIPWALL: {
          if (!horiz) break;  // left/right: not implemented in the reference (bc.cpp:121-126)
          if (t < 1 || t > nic) { err = "slipwall range must lie in 1..nic (the wall normal of cell i-1 is used)"; return false; }
          const bool bot = B.face == SG_FACE_BOTTOM;
          nd.op = OP_SLIP;
          if (!C.ref(t, bot ? 1 : njc, nd.ra) || !C.ref(t, bot ? 2 : njc - 1, nd.rb)) return false;
          nd.c0 = NE(t - 1, bot ? 0 : njc, 0);
          nd.c1 = NE(t - 1, bot ? 0 : njc, 1);
          if (!C.write(t, jend, nd, true, true)) return false;
        } break;
        case SG_BC_WALL:
        case SG_BC_ISOTHERMAL: {
          if (B.type == SG_BC_ISOTHERMAL && !horiz) break;  // not implemented in the reference
          nd.op = B.type == SG_BC_WALL ? OP_WALL : OP_ISO;
          nd.c0 = B.u; nd.c1 = B.v; nd.c2 = B.T;
          bool ok;
          if (horiz) {
            const bool bot = B.face == SG_FACE_BOTTOM;
            ok = C.ref(t, bot ? 1 : njc, nd.ra) && C.ref(t, bot ? 2 : njc - 1, nd.rb) && C.write(t, jend, nd, true, true);
          } else {
            const bool left = B.face == SG_FACE_LEFT;
            ok = C.ref(left ? 1 : nic, t, nd.ra) && C.ref(left ? 2 : nic - 1, t, nd.rb) && C.write(iend, t, nd, true, true);
          }
          if (!ok) return false;
        } break;
        case SG_BC_OUTFLOW: {
          if (horiz) break;  // not implemented in the reference (bc.cpp:288-292)
          if (B.face == SG_FACE_LEFT) { err = "outflow on the left face reads index -1 in the reference"; return false; }
          nd.op = OP_OUTFLOW;
          if (!C.ref(nic, t, nd.ra) || !C.write(iend, t, nd, true, false)) return false;
        } break;
        case SG_BC_PERIODIC: {
          nd.op = OP_COPY;
          if (horiz) {
            if (!C.ref(t, njc, nd.ra) || !C.write(t, 0, nd, true, false)) return false;
            if (!C.ref(t, 1, nd.ra) || !C.write(t, njc + 1, nd, true, false)) return false;
          } else {
            if (!C.ref(nic, t, nd.ra) || !C.write(0, t, nd, true, false)) return false;
            if (!C.ref(1, t, nd.ra) || !C.write(nic + 1, t, nd, true, false)) return false;
          }
        } break;
        case SG_BC_WAKE:
          break;  // handled after this loop (two sequential sweeps)
        default:
          err = "unknown boundary type";
          return false;
      }
    }
    if (B.type == SG_BC_WAKE && horiz) {  // src/model/bc.cpp:223-263: both sweeps read row 1
      for (long long s = start; s <= end; ++s) {
        BcNode nd{}; nd.op = OP_COPY;
        const int t = (int)s;
        if (!C.ref(nic + 1 - t, 1, nd.ra) || !C.write(t, jend, nd, true, false)) return false;
      }
      for (long long s = start; s <= end; ++s) {
        BcNode nd{}; nd.op = OP_COPY;
        const int t = (int)s;
        if (!C.ref(t, 1, nd.ra) || !C.write(nic + 1 - t, jend, nd, true, false)) return false;
      }
    }
  }
  // Locality of the ghost dependencies (see BcCompiled::local): every
  // interior cell that reads a ghost (those within one cell of it) must see
  // the ghost's interior operands inside its 13-point stencil.
  local = true;
  local_box = true;
  auto in_box = [](int di, int dj) { return std::abs(di) <= 1 && std::abs(dj) <= 1; };
  auto in_stencil = [&](int di, int dj) {
    return in_box(di, dj) || (di == 0 && std::abs(dj) == 2) || (dj == 0 && std::abs(di) == 2);
  };
  for (int gi = 0; gi <= nic + 1 && local; ++gi)
    for (int gj = 0; gj <= njc + 1 && local; ++gj) {
      const int p = C.perim(gi, gj);
      if (p < 0 || ghost_map[p] < 0) continue;
      std::vector<int> stack{ghost_map[p]}, leaves;
      while (!stack.empty()) {
        const BcNode& nd = nodes[stack.back()];
        stack.pop_back();
        if (nd.op == OP_FREESTREAM) continue;
        const bool two = nd.op == OP_SLIP || nd.op == OP_WALL || nd.op == OP_ISO;
        for (int r : {nd.ra, two ? nd.rb : nd.ra}) {
          if (r < 0) stack.push_back(-(r + 1));
          else leaves.push_back(r);
        }
      }
      for (int ui = gi - 1; ui <= gi + 1 && local; ++ui)
        for (int uj = gj - 1; uj <= gj + 1 && local; ++uj) {
          if (ui < 1 || ui > nic || uj < 1 || uj > njc) continue;  // users are interior cells
          for (int r : leaves) {
            const int di = r % nic + 1 - ui, dj = r / nic + 1 - uj;
            if (!in_box(di, dj)) local_box = false;
            if (!in_stencil(di, dj)) { local = false; local_box = false; break; }
          }
        }
    }
  // Edge ghosts feed the boundary faces; corners feed the viscous stencil.
  for (int gi = 1; gi <= nic; ++gi)
    if (ghost_map[C.perim(gi, 0)] < 0 || ghost_map[C.perim(gi, njc + 1)] < 0) { err = "bottom/top ghost row not fully set by the boundary list"; return false; }
  for (int gj = 1; gj <= njc; ++gj)
    if (ghost_map[C.perim(0, gj)] < 0 || ghost_map[C.perim(nic + 1, gj)] < 0) { err = "left/right ghost column not fully set by the boundary list"; return false; }
  if (viscous)
    for (int gi : {0, nic + 1})
      for (int gj : {0, njc + 1})
        if (ghost_map[C.perim(gi, gj)] < 0) { err = "corner ghost cell not set by the boundary list (needed by the viscous stencil)"; return false; }
  return true;
}

}  // namespace sg

// ------------------------------------------------------------- wall distance
namespace {
__global__ void wall_distance_kernel(int nic, int njc, int ldi, const double* xc, const double* yc, int nseg,
                                     const double* seg, double* d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= nic) return;
  const size_t c = (size_t)j * ldi + i;
  double px = xc[c], py = yc[c], best = INFINITY;
  for (int s = 0; s < nseg; ++s) {
    double x0 = seg[4 * s], y0 = seg[4 * s + 1], ex = seg[4 * s + 2] - x0, ey = seg[4 * s + 3] - y0;
    double t = ((px - x0) * ex + (py - y0) * ey) / (ex * ex + ey * ey);
    t = fmin(1.0, fmax(0.0, t));
    double dx = px - (x0 + t * ex), dy = py - (y0 + t * ey);
    best = fmin(best, sqrt(dx * dx + dy * dy));
  }
  d[c] = best;
}

template <class F>
std::vector<double> to_soa_cells(const sg::HostGeom& G, int ldi, F f) {
  std::vector<double> out((size_t)G.njc * ldi, 0.0);
  for (int i = 0; i < G.nic; ++i)
    for (int j = 0; j < G.njc; ++j) out[(size_t)j * ldi + i] = f(i, j);
  return out;
}

int upload(sg_ctx* ctx, sg::DevBuf& b, const std::vector<double>& h) {
  SG_TRY(b.alloc(h.size()), "hipMalloc");
  SG_TRY(hipMemcpy(b.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy H2D");
  return SG_OK;
}

int compute_wall_distance(sg_ctx* ctx, const sg_bc* bcs, int nbc) {
  const sg::HostGeom& G = ctx->geo;
  std::vector<double> seg;
  auto X = [&](int i, int j) { return G.xv[(size_t)i * G.nj + j]; };
  auto Y = [&](int i, int j) { return G.yv[(size_t)i * G.nj + j]; };
  for (int b = 0; b < nbc; ++b) {
    const sg_bc& B = bcs[b];
    if (B.type != SG_BC_WALL && B.type != SG_BC_ISOTHERMAL) continue;
    if (B.face != SG_FACE_BOTTOM && B.face != SG_FACE_TOP) continue;
    const int j = B.face == SG_FACE_BOTTOM ? 0 : G.nj - 1;
    long long end = B.end < 0 ? G.nic + 2 + B.end : B.end;
    for (long long s = std::max<long long>(B.start, 1); s <= std::min<long long>(end, G.nic); ++s) {
      int i = (int)s - 1;
      seg.insert(seg.end(), {X(i, j), Y(i, j), X(i + 1, j), Y(i + 1, j)});
    }
  }
  const int ldi = ctx->g.ldi;
  const size_t n = (size_t)G.njc * ldi;
  if (seg.empty()) {
    std::vector<double> big(n, 1e30);
    return upload(ctx, ctx->wd, big);
  }
  sg::DevBuf xc, yc, sg_;
  int rc;
  if ((rc = upload(ctx, xc, to_soa_cells(G, ldi, [&](int i, int j) { return G.xc[(size_t)i * G.njc + j]; }))) != SG_OK) return rc;
  if ((rc = upload(ctx, yc, to_soa_cells(G, ldi, [&](int i, int j) { return G.yc[(size_t)i * G.njc + j]; }))) != SG_OK) return rc;
  if ((rc = upload(ctx, sg_, seg)) != SG_OK) return rc;
  SG_TRY(ctx->wd.alloc(n), "hipMalloc");
  SG_TRY(hipMemset(ctx->wd.p, 0, n * sizeof(double)), "hipMemset");
  hipLaunchKernelGGL(wall_distance_kernel, dim3((G.nic + 255) / 256, G.njc), dim3(256), 0, ctx->stream, G.nic, G.njc,
                     ldi, xc.p, yc.p, (int)(seg.size() / 4), sg_.p, ctx->wd.p);
  SG_TRY(hipGetLastError(), "wall distance launch");
  SG_TRY(hipStreamSynchronize(ctx->stream), "wall distance");
  return SG_OK;
}

int check_finite(sg_ctx* ctx, const double* l2) {
  for (int k = 0; k < ctx->nvar; ++k)
    if (!std::isfinite(l2[k])) return sg_fail(ctx, SG_ERR_NAN, "non-finite residual norm");
  return SG_OK;
}

int upload_aos(sg_ctx* ctx, const double* h, double* soa) {
  const size_t n = ctx->ncell() * ctx->nvar;
  SG_TRY(hipMemcpyAsync(ctx->aos_a.p, h, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream), "H2D");
  SG_TRY(launch_aos_to_soa(ctx->geo.nic, ctx->geo.njc, ctx->g.ldi, ctx->nvar, ctx->aos_a.p, soa, ctx->stream), "aos->soa");
  return SG_OK;
}

int download_aos(sg_ctx* ctx, const double* soa, double* h) {
  const size_t n = ctx->ncell() * ctx->nvar;
  SG_TRY(launch_soa_to_aos(ctx->geo.nic, ctx->geo.njc, ctx->g.ldi, ctx->nvar, soa, ctx->aos_b.p, ctx->stream), "soa->aos");
  SG_TRY(hipMemcpyAsync(h, ctx->aos_b.p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "D2H");
  return SG_OK;
}

}  // namespace

// ------------------------------------------------------------------ C ABI --
extern "C" {

int sg_create(int ni, int nj, const double* xv, const double* yv, const sg_freestream* fs, const sg_bc* bcs, int nbc,
              const sg_opts* opts, sg_ctx** out) {
  if (!out) return SG_ERR_ARG;
  *out = nullptr;
  g_create_err.clear();
  if (ni < 3 || nj < 3 || !xv || !yv || !fs || !opts || (nbc > 0 && !bcs) || nbc < 0)
    return sg_fail(nullptr, SG_ERR_ARG, "bad shape or null argument (ni, nj >= 3 required)");
  if (opts->order != 1 && opts->order != 2) return sg_fail(nullptr, SG_ERR_ARG, "order must be 1 or 2");
  if (opts->flux != SG_FLUX_ROE && opts->flux != SG_FLUX_AUSM)
    return sg_fail(nullptr, SG_ERR_ARG, "flux must be SG_FLUX_ROE or SG_FLUX_AUSM");
  sg_ctx* ctx = new sg_ctx();
  ctx->fs = *fs;
  ctx->opts = *opts;
  ctx->sa = opts->turbulence != 0;
  ctx->nvar = ctx->sa ? 5 : 4;
  ctx->geo.build(ni, nj, xv, yv);
  const HostGeom& G = ctx->geo;
  const bool viscous = fs->mu_inf > 1e-15;  // src/utils/config.cpp:41
  if (ctx->sa && !viscous) { delete ctx; return sg_fail(nullptr, SG_ERR_ARG, "the SA model needs mu_inf > 0"); }
  if (!ctx->bcc.compile(G, bcs, nbc, viscous)) {
    g_create_err = ctx->bcc.err;
    delete ctx;
    return SG_ERR_BC;
  }
  int rc;
  auto fail = [&](int code) {
    g_create_err = ctx->err.empty() ? std::string("device setup failed") : ctx->err;
    sg_destroy(ctx);
    return code;
  };
  if (hipStreamCreate(&ctx->stream) != hipSuccess) return fail(SG_ERR_HIP);
  if ((rc = sg_noalloc_setup(ctx))) return fail(rc);
  // device geometry
  const int ldi = ((G.nic + 63) / 64) * 64, ldf = ((G.ni + 63) / 64) * 64;
  DevGeom& g = ctx->g;
  g.ni = G.ni; g.nj = G.nj; g.nic = G.nic; g.njc = G.njc; g.ldi = ldi; g.ldf = ldf;
  g.plane = (size_t)G.njc * ldi;
  std::vector<double> ncx((size_t)G.njc * ldf, 0.0), ncy(ncx.size(), 0.0), nex((size_t)G.nj * ldi, 0.0), ney(nex.size(), 0.0);
  for (int i = 0; i < G.ni; ++i)
    for (int j = 0; j < G.njc; ++j) {
      ncx[(size_t)j * ldf + i] = G.normal_chi[((size_t)i * G.njc + j) * 2];
      ncy[(size_t)j * ldf + i] = G.normal_chi[((size_t)i * G.njc + j) * 2 + 1];
    }
  for (int i = 0; i < G.nic; ++i)
    for (int j = 0; j < G.nj; ++j) {
      nex[(size_t)j * ldi + i] = G.normal_eta[((size_t)i * G.nj + j) * 2];
      ney[(size_t)j * ldi + i] = G.normal_eta[((size_t)i * G.nj + j) * 2 + 1];
    }
  if ((rc = upload(ctx, ctx->ncx, ncx)) || (rc = upload(ctx, ctx->ncy, ncy)) || (rc = upload(ctx, ctx->nex, nex)) ||
      (rc = upload(ctx, ctx->ney, ney)))
    return fail(rc);
  if ((rc = upload(ctx, ctx->vol, to_soa_cells(G, ldi, [&](int i, int j) { return G.volume[(size_t)i * G.njc + j]; }))))
    return fail(rc);
  if ((rc = upload(ctx, ctx->dsmin, to_soa_cells(G, ldi, [&](int i, int j) {
         size_t c = (size_t)i * G.njc + j;
         return std::min(G.ds_eta[c], G.ds_chi[c]);
       }))))
    return fail(rc);
  if ((rc = upload(ctx, ctx->beta, std::vector<double>(g.plane, 1.0)))) return fail(rc);
  std::vector<double> x0(G.ni), y0(G.ni);
  for (int i = 0; i < G.ni; ++i) { x0[i] = G.xv[(size_t)i * G.nj]; y0[i] = G.yv[(size_t)i * G.nj]; }
  if ((rc = upload(ctx, ctx->xv0, x0)) || (rc = upload(ctx, ctx->yv0, y0))) return fail(rc);
  if (ctx->sa) { if ((rc = compute_wall_distance(ctx, bcs, nbc))) return fail(rc); }
  else if ((rc = upload(ctx, ctx->wd, std::vector<double>(g.plane, 1.0)))) return fail(rc);
  g.ncx = ctx->ncx.p; g.ncy = ctx->ncy.p; g.nex = ctx->nex.p; g.ney = ctx->ney.p; g.vol = ctx->vol.p;
  g.wd = ctx->wd.p; g.beta = ctx->beta.p;
  // physics
  DevPhys& ph = ctx->ph;
  ph.rho_inf = fs->rho_inf; ph.u_inf = fs->u_inf; ph.v_inf = fs->v_inf; ph.p_inf = fs->p_inf; ph.T_inf = fs->T_inf;
  ph.mu_inf = fs->mu_inf; ph.pr = fs->pr_inf; ph.nut_inf = fs->nut_inf;
  ph.R = fs->p_inf / fs->rho_inf / fs->T_inf;          // src/model/fluid.cpp:12
  ph.iR = 1.0 / ph.R; ph.ipr = 1.0 / fs->pr_inf; ph.iTinf = 1.0 / fs->T_inf;
  ph.cp = kGamma * ph.R / (kGamma - 1.0);
  ph.eps_chi = std::pow(10.0 / G.nic, 3);               // src/model/reconstruction.cpp:62-63
  ph.eps_eta = std::pow(10.0 / G.njc, 3);
  ph.dpdx = opts->dpdx; ph.dpdy = opts->dpdy; ph.order = opts->order; ph.viscous = viscous ? 1 : 0;
  ph.flux = opts->flux == SG_FLUX_AUSM ? FLUX_AUSM : FLUX_ROE; ph.pad_ = 0;
  // boundary nodes
  {
    const size_t nb = ctx->bcc.nodes.size() * sizeof(BcNode), nm = ctx->bcc.ghost_map.size() * sizeof(int);
    if (ctx->nodes_buf.alloc((nb + 7) / 8 + 1) != hipSuccess || ctx->map_buf.alloc((nm + 7) / 8 + 1) != hipSuccess)
      return fail(SG_ERR_HIP);
    if (nb && hipMemcpy(ctx->nodes_buf.p, ctx->bcc.nodes.data(), nb, hipMemcpyHostToDevice) != hipSuccess) return fail(SG_ERR_HIP);
    if (hipMemcpy(ctx->map_buf.p, ctx->bcc.ghost_map.data(), nm, hipMemcpyHostToDevice) != hipSuccess) return fail(SG_ERR_HIP);
    ctx->bc.nodes = reinterpret_cast<const BcNode*>(ctx->nodes_buf.p);
    ctx->bc.ghost_map = reinterpret_cast<const int*>(ctx->map_buf.p);
  }
  // work buffers
  const size_t naos = ctx->ncell() * ctx->nvar, nsoa = ctx->nsoa();
  const size_t npart = residual_partial_count(g) * 2 * ctx->nvar;
  if (ctx->aos_a.alloc(naos) || ctx->aos_b.alloc(naos) || ctx->q.alloc(nsoa) || ctx->r.alloc(nsoa) ||
      ctx->v.alloc(nsoa) || ctx->w.alloc(nsoa) || ctx->partial.alloc(npart) || ctx->l2.alloc(16) || ctx->counter.alloc(sg::kNormStateWords) ||
      ctx->wall_cp.alloc(G.ni) || ctx->wall_cf.alloc(G.ni) || ctx->wall_f.alloc(4))
    return fail(SG_ERR_HIP);
  // zero the SoA padding once so padded lanes never hold garbage; the fused
  // norm's two state words and its partials start at zero
  for (DevBuf* b : {&ctx->q, &ctx->r, &ctx->v, &ctx->w, &ctx->counter, &ctx->partial})
    if (hipMemset(b->p, 0, b->n * sizeof(double)) != hipSuccess) return fail(SG_ERR_HIP);
  *out = ctx;
  return SG_OK;
}

void sg_destroy(sg_ctx* ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  sg_solver_free(ctx);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* sg_last_error(const sg_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }
int sg_nvar(const sg_ctx* ctx) { return ctx ? ctx->nvar : SG_ERR_ARG; }
int sg_ldi(const sg_ctx* ctx) { return ctx ? ctx->g.ldi : SG_ERR_ARG; }

static int set_cell_field(sg_ctx* ctx, DevBuf& b, const double* h) {
  if (!ctx || !h) return SG_ERR_ARG;
  const HostGeom& G = ctx->geo;
  std::vector<double> s = to_soa_cells(G, ctx->g.ldi, [&](int i, int j) { return h[(size_t)i * G.njc + j]; });
  SG_TRY(hipMemcpy(b.p, s.data(), s.size() * sizeof(double), hipMemcpyHostToDevice), "H2D");
  return SG_OK;
}

int sg_set_beta(sg_ctx* ctx, const double* beta) { return set_cell_field(ctx, ctx->beta, beta); }

// Closure buffers: MLP parameters as FP32 plus the BF16 copy of W2.
int sg_closure_alloc_mlp(sg_ctx* ctx) {
  if (!ctx->clo_params.p) SG_TRY(ctx->clo_params.alloc((MLP_NPARAM + 1) / 2), "hipMalloc");
  if (!ctx->clo_w2.p) SG_TRY(ctx->clo_w2.alloc(MLP_H * MLP_H / 4), "hipMalloc");
  return SG_OK;
}

int sg_set_closure(sg_ctx* ctx, int kind, const double* coef, const float* mlp_params) {
  if (!ctx) return SG_ERR_ARG;
  if (!ctx->sa && kind != SG_CLOSURE_FIELD) return sg_fail(ctx, SG_ERR_ARG, "a beta closure needs the SA model");
  DevClosure c;
  c.kind = kind;
  if (kind == SG_CLOSURE_ANALYTIC) {
    if (!coef) return sg_fail(ctx, SG_ERR_ARG, "analytic closure needs 7 coefficients");
    std::copy(coef, coef + 7, c.coef);
  } else if (kind == SG_CLOSURE_MLP) {
    if (!mlp_params) return sg_fail(ctx, SG_ERR_ARG, "MLP closure needs its parameters");
    int rc;
    if ((rc = sg_closure_alloc_mlp(ctx))) return rc;
    SG_TRY(hipMemcpyAsync(ctx->clo_params.p, mlp_params, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice,
                          ctx->stream), "H2D");
    SG_TRY(mlp_w2_to_bf16(reinterpret_cast<const float*>(ctx->clo_params.p),
                          reinterpret_cast<__bf16*>(ctx->clo_w2.p), ctx->stream), "w2 bf16");
    SG_TRY(hipStreamSynchronize(ctx->stream), "closure");
    c.params = reinterpret_cast<const float*>(ctx->clo_params.p);
    c.w2 = ctx->clo_w2.p;
  } else if (kind != SG_CLOSURE_FIELD) {
    return sg_fail(ctx, SG_ERR_ARG, "unknown closure kind");
  }
  ctx->clo = c;
  return SG_OK;
}

int sg_set_mlp_weights(sg_ctx* ctx, const float* params) {
  if (!ctx || !params) return SG_ERR_ARG;
  return sg_set_closure(ctx, SG_CLOSURE_MLP, nullptr, params);
}

int sg_closure_eval(sg_ctx* ctx, const double* q_aos, double* features, double* beta) {
  if (!ctx || !q_aos) return SG_ERR_ARG;
  if (!ctx->sa) return sg_fail(ctx, SG_ERR_ARG, "closure features need the SA model");
  if (!ctx->ph.viscous) return sg_fail(ctx, SG_ERR_ARG, "closure features need a viscous flow");
  int rc;
  if (!ctx->aux.p) SG_TRY(ctx->aux.alloc((kNFeat + 1) * ctx->g.plane), "hipMalloc");
  if ((rc = upload_aos(ctx, q_aos, ctx->q.p))) return rc;
  DevClosure c = ctx->clo;
  c.aux = ctx->aux.p;
  SG_TRY(launch_residual(ctx->g, ctx->ph, ctx->bc, true, ctx->q.p, ctx->r.p, nullptr, nullptr, nullptr, ctx->stream, c),
         "residual launch");
  const size_t nc = ctx->ncell();
  std::vector<double> f(nc * kNFeat), b(nc);
  SG_TRY(launch_soa_to_aos(ctx->geo.nic, ctx->geo.njc, ctx->g.ldi, kNFeat, ctx->aux.p, ctx->aos_b.p, ctx->stream),
         "soa->aos");
  SG_TRY(hipMemcpyAsync(f.data(), ctx->aos_b.p, f.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "D2H");
  SG_TRY(launch_soa_to_aos(ctx->geo.nic, ctx->geo.njc, ctx->g.ldi, 1, ctx->aux.p + kNFeat * ctx->g.plane,
                           ctx->aos_a.p, ctx->stream), "soa->aos");
  SG_TRY(hipMemcpyAsync(b.data(), ctx->aos_a.p, b.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "D2H");
  SG_TRY(hipStreamSynchronize(ctx->stream), "closure eval");
  if (features) std::copy(f.begin(), f.end(), features);
  if (beta) std::copy(b.begin(), b.end(), beta);
  return SG_OK;
}
int sg_set_walldist(sg_ctx* ctx, const double* d) { return set_cell_field(ctx, ctx->wd, d); }

int sg_get_walldist(sg_ctx* ctx, double* d) {
  if (!ctx || !d) return SG_ERR_ARG;
  const HostGeom& G = ctx->geo;
  std::vector<double> s(ctx->g.plane);
  SG_TRY(hipMemcpy(s.data(), ctx->wd.p, s.size() * sizeof(double), hipMemcpyDeviceToHost), "D2H");
  for (int i = 0; i < G.nic; ++i)
    for (int j = 0; j < G.njc; ++j) d[(size_t)i * G.njc + j] = s[(size_t)j * ctx->g.ldi + i];
  return SG_OK;
}

int sg_get_metrics(sg_ctx* ctx, double* nchi, double* neta, double* volume, double* xc, double* yc) {
  if (!ctx) return SG_ERR_ARG;
  const HostGeom& G = ctx->geo;
  if (nchi) std::copy(G.normal_chi.begin(), G.normal_chi.end(), nchi);
  if (neta) std::copy(G.normal_eta.begin(), G.normal_eta.end(), neta);
  if (volume) std::copy(G.volume.begin(), G.volume.end(), volume);
  if (xc) std::copy(G.xc.begin(), G.xc.end(), xc);
  if (yc) std::copy(G.yc.begin(), G.yc.end(), yc);
  return SG_OK;
}

int sg_residual_dev_async(sg_ctx* ctx, const double* q_dev, double* r_dev, double* l2_dev) {
  if (!ctx || !q_dev || !r_dev) return SG_ERR_ARG;
  // the caller's l2_dev holds nvar norms (||R_k||_2): only those are written
  SG_TRY(launch_residual(ctx->g, ctx->ph, ctx->bc, ctx->sa, q_dev, r_dev, l2_dev ? ctx->partial.p : nullptr,
                         reinterpret_cast<unsigned long long*>(ctx->counter.p), l2_dev, ctx->stream, ctx->clo, ctx->nvar,
                         &ctx->norm_parity),
         "residual launch");
  return SG_OK;
}

void* sg_get_stream(sg_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int sg_residual_dev(sg_ctx* ctx, const double* q_dev, double* r_dev, double* l2_host) {
  if (!ctx || !q_dev || !r_dev) return SG_ERR_ARG;
  SG_TRY(launch_residual(ctx->g, ctx->ph, ctx->bc, ctx->sa, q_dev, r_dev, ctx->partial.p,
                         reinterpret_cast<unsigned long long*>(ctx->counter.p), ctx->l2.p, ctx->stream, ctx->clo, 0,
                         &ctx->norm_parity),
         "residual launch");
  double l2[8];
  SG_TRY(hipMemcpyAsync(l2, ctx->l2.p, ctx->nvar * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "D2H");
  SG_TRY(hipStreamSynchronize(ctx->stream), "residual");
  if (l2_host) std::copy(l2, l2 + ctx->nvar, l2_host);
  return check_finite(ctx, l2);
}

int sg_residual(sg_ctx* ctx, const double* q_aos, double* r_aos, double* l2) {
  if (!ctx || !q_aos || !r_aos) return SG_ERR_ARG;
  int rc;
  if ((rc = upload_aos(ctx, q_aos, ctx->q.p))) return rc;
  double l2h[8];
  SG_TRY(launch_residual(ctx->g, ctx->ph, ctx->bc, ctx->sa, ctx->q.p, ctx->r.p, ctx->partial.p,
                         reinterpret_cast<unsigned long long*>(ctx->counter.p), ctx->l2.p, ctx->stream, ctx->clo, 0,
                         &ctx->norm_parity),
         "residual launch");
  SG_TRY(hipMemcpyAsync(l2h, ctx->l2.p, ctx->nvar * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "D2H");
  if ((rc = download_aos(ctx, ctx->r.p, r_aos))) return rc;
  SG_TRY(hipStreamSynchronize(ctx->stream), "residual");
  if (l2) std::copy(l2h, l2h + ctx->nvar, l2);
  return check_finite(ctx, l2h);
}

int sg_jv_dev_async(sg_ctx* ctx, const double* q_dev, const double* v_dev, double* jv_dev) {
  if (!ctx || !q_dev || !v_dev || !jv_dev) return SG_ERR_ARG;
  SG_TRY(launch_jv(ctx->g, ctx->ph, ctx->bc, ctx->sa, q_dev, v_dev, nullptr, jv_dev, ctx->stream, nullptr, ctx->clo),
         "jv launch");
  return SG_OK;
}

// EulerEquation::initialize, src/model/eulerequation.cpp:262-277: the uniform
// freestream state (rho, rho u, rho v, p/(gamma-1) + rho |u|^2 / 2), with
// rho nu~_inf as the fifth unknown of the SA model (the reference has no
// transport equation; its ntrans hook is src/solver/solution.cpp:9).
int sg_initialize(sg_ctx* ctx, double* q_aos) {
  if (!ctx || !q_aos) return SG_ERR_ARG;
  const sg_freestream& f = ctx->fs;
  const double e = f.p_inf / (kGamma - 1.0) + 0.5 * f.rho_inf * (f.u_inf * f.u_inf + f.v_inf * f.v_inf);
  const int nv = ctx->nvar;
  const size_t nc = ctx->ncell();
  for (size_t c = 0; c < nc; ++c) {
    double* q = q_aos + c * nv;
    q[0] = f.rho_inf;
    q[1] = f.rho_inf * f.u_inf;
    q[2] = f.rho_inf * f.v_inf;
    q[3] = e;
    if (nv > 4) q[4] = f.rho_inf * f.nut_inf;
  }
  return SG_OK;
}

int sg_jv_dev(sg_ctx* ctx, const double* q_dev, const double* v_dev, double* jv_dev) {
  if (!ctx || !q_dev || !v_dev || !jv_dev) return SG_ERR_ARG;
  SG_TRY(launch_jv(ctx->g, ctx->ph, ctx->bc, ctx->sa, q_dev, v_dev, nullptr, jv_dev, ctx->stream, nullptr, ctx->clo),
         "jv launch");
  SG_TRY(hipStreamSynchronize(ctx->stream), "jv");
  return SG_OK;
}

int sg_jv(sg_ctx* ctx, const double* q_aos, const double* v_aos, double* jv_aos) {
  if (!ctx || !q_aos || !v_aos || !jv_aos) return SG_ERR_ARG;
  int rc;
  if ((rc = upload_aos(ctx, q_aos, ctx->q.p))) return rc;
  if ((rc = upload_aos(ctx, v_aos, ctx->v.p))) return rc;
  SG_TRY(launch_jv(ctx->g, ctx->ph, ctx->bc, ctx->sa, ctx->q.p, ctx->v.p, nullptr, ctx->w.p, ctx->stream, nullptr,
                   ctx->clo), "jv launch");
  if ((rc = download_aos(ctx, ctx->w.p, jv_aos))) return rc;
  SG_TRY(hipStreamSynchronize(ctx->stream), "jv");
  return SG_OK;
}

int sg_dt(sg_ctx* ctx, const double* q_aos, double cfl, double* dt_aos) {
  if (!ctx || !q_aos || !dt_aos) return SG_ERR_ARG;
  int rc;
  if ((rc = upload_aos(ctx, q_aos, ctx->q.p))) return rc;
  SG_TRY(launch_dt(ctx->g, ctx->ph.mu_inf, ctx->nvar, ctx->q.p, ctx->dsmin.p, cfl, ctx->w.p, ctx->stream), "dt launch");
  if ((rc = download_aos(ctx, ctx->w.p, dt_aos))) return rc;
  SG_TRY(hipStreamSynchronize(ctx->stream), "dt");
  return SG_OK;
}

int sg_wall(sg_ctx* ctx, const double* q_aos, double* x, double* cp, double* cf, double* forces) {
  if (!ctx || !q_aos) return SG_ERR_ARG;
  if (!ctx->ph.viscous) return sg_fail(ctx, SG_ERR_ARG, "wall quantities need a viscous case");
  const HostGeom& G = ctx->geo;
  const long tail = ctx->opts.tail;
  const long first = tail - 1, nb = (long)G.ni - 2 * tail + 1;   // src/utils/mesh.cpp:349-350, io.cpp:232
  if (first < 0 || nb <= 0 || first + nb > G.nic) return sg_fail(ctx, SG_ERR_ARG, "tail outside the grid");
  int rc;
  if ((rc = upload_aos(ctx, q_aos, ctx->q.p))) return rc;
  const double qinf = 0.5 * ctx->fs.rho_inf * (ctx->fs.u_inf * ctx->fs.u_inf + ctx->fs.v_inf * ctx->fs.v_inf);
  SG_TRY(launch_wall(ctx->g, ctx->ph, ctx->bc, ctx->sa, ctx->q.p, (int)first, (int)nb, ctx->xv0.p, ctx->yv0.p, qinf,
                     ctx->wall_cp.p, ctx->wall_cf.p, ctx->wall_f.p, ctx->stream), "wall launch");
  std::vector<double> hcp(nb), hcf(nb);
  double f[4];
  SG_TRY(hipMemcpyAsync(hcp.data(), ctx->wall_cp.p, nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "D2H");
  SG_TRY(hipMemcpyAsync(hcf.data(), ctx->wall_cf.p, nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "D2H");
  SG_TRY(hipMemcpyAsync(f, ctx->wall_f.p, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "D2H");
  SG_TRY(hipStreamSynchronize(ctx->stream), "wall");
  for (long n = 0; n < nb; ++n) {
    if (x) x[n] = G.xc[(size_t)(first + n) * G.njc];
    if (cp) cp[n] = hcp[n];
    if (cf) cf[n] = hcf[n];
  }
  if (forces) {  // src/utils/io.cpp:239-247
    const double aoa = ctx->fs.aoa_deg * M_PI / 180.0, ca = std::cos(aoa), sa = std::sin(aoa);
    const double Fnp = f[0], Fcp = f[1], Fnv = f[2], Fcv = f[3];
    forces[0] = -Fcp * sa + Fnp * ca; forces[1] = Fcp * ca + Fnp * sa;
    forces[2] = -Fcv * sa + Fnv * ca; forces[3] = Fcv * ca + Fnv * sa;
    forces[4] = forces[0] + forces[2]; forces[5] = forces[1] + forces[3];
  }
  return SG_OK;
}

int sg_synchronize(sg_ctx* ctx) {
  if (!ctx) return SG_ERR_ARG;
  SG_TRY(hipStreamSynchronize(ctx->stream), "synchronize");
  return SG_OK;
}

int sg_aos_to_soa_dev(sg_ctx* ctx, const double* aos_dev, double* soa_dev, int nvar) {
  if (!ctx || !aos_dev || !soa_dev || nvar <= 0) return SG_ERR_ARG;
  SG_TRY(launch_aos_to_soa(ctx->geo.nic, ctx->geo.njc, ctx->g.ldi, nvar, aos_dev, soa_dev, ctx->stream), "aos->soa");
  return SG_OK;
}

int sg_soa_to_aos_dev(sg_ctx* ctx, const double* soa_dev, double* aos_dev, int nvar) {
  if (!ctx || !aos_dev || !soa_dev || nvar <= 0) return SG_ERR_ARG;
  SG_TRY(launch_soa_to_aos(ctx->geo.nic, ctx->geo.njc, ctx->g.ldi, nvar, soa_dev, aos_dev, ctx->stream), "soa->aos");
  return SG_OK;
}

// The adjoint entry points (sg_jtv, sg_adjoint_*, sg_objective_cf) are in sg_solver.cpp.

}  // extern "C"
