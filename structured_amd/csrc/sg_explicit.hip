// Explicit pseudo-time stepping on the device: the build of Solver::step
// without ADOL-C (src/solver/solver.cpp:103-121) and the loop of
// Solver::solve (:197-236).
//
// Each stage is one residual launch (norms fused into it) and one
// bandwidth-bound update kernel.  The update recomputes the cell's local time
// step from the step's base state instead of storing a dt array: the
// reference computes dt once per step from q (calc_dt, before the stages), and
// every stage reads that q anyway, so the dt field never reaches HBM.  Per
// cell and stage the update moves q (nv) + R (nv) + the shorter face length
// in and nv values out.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>

#include "sg_context.hpp"
#include "sg_kernels.hpp"

using namespace sg;

#define SG_TRY(expr, what)                         \
  do {                                             \
    int rc_ = sg_hip(ctx, (expr), what);           \
    if (rc_ != SG_OK) return rc_;                  \
  } while (0)

namespace {

// out = q0 + (R dt) / denom with dt = calc_dt(q0) of the cell: update_forward_euler
// (denom 1) and update_rk4 (denom 4 - stage), src/solver/solver.cpp:12-26.
// out may alias q0 (each thread reads its cell before writing it).
__global__ __launch_bounds__(256) void explicit_stage_kernel(DevGeom g, int nv, const double* q0, const double* R,
                                                             const double* dsmin, double mu_inf, double cfl,
                                                             double denom, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= g.nic) return;
  const size_t c = (size_t)j * g.ldi + i;
  double qv[5], rv[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    if (k < nv) { qv[k] = q0[k * g.plane + c]; rv[k] = R[k * g.plane + c]; }
  }
  const double dt = cell_dt(qv[0], qv[1], qv[2], qv[3], dsmin[c], mu_inf, cfl);
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    if (k < nv) {
#pragma clang fp contract(off)
      out[k * g.plane + c] = qv[k] + (rv[k] * dt) / denom;
    }
  }
}

hipError_t launch_stage(sg_ctx* ctx, const double* q0, const double* R, double cfl, double denom, double* out) {
  const DevGeom& g = ctx->g;
  hipLaunchKernelGGL(explicit_stage_kernel, dim3((g.nic + 255) / 256, g.njc), dim3(256), 0, ctx->stream, g, ctx->nvar,
                     q0, R, ctx->dsmin.p, ctx->ph.mu_inf, cfl, denom, out);
  return hipGetLastError();
}

int ensure_buffers(sg_ctx* ctx) {
  const size_t n = ctx->nsoa();
  for (DevBuf* b : {&ctx->ex_a, &ctx->ex_b, &ctx->ex_r})
    if (b->n != n) SG_TRY(b->alloc(n), "explicit work buffers");
  return SG_OK;
}

// Enqueues one step; the norms of the last stage's residual land in ctx->l2.
int enqueue_step(sg_ctx* ctx, double* q, int scheme, double cfl) {
  unsigned long long* counter = reinterpret_cast<unsigned long long*>(ctx->counter.p);
  double* R = ctx->ex_r.p;
  auto residual = [&](const double* in) {
    return launch_residual(ctx->g, ctx->ph, ctx->bc, ctx->sa, in, R, ctx->partial.p, counter, ctx->l2.p, ctx->stream,
                           ctx->clo, 0, &ctx->norm_parity);
  };
  if (scheme == SG_SCHEME_FORWARD_EULER) {
    SG_TRY(residual(q), "residual");
    SG_TRY(launch_stage(ctx, q, R, cfl, 1.0, q), "update");
    return SG_OK;
  }
  // RK4-Jameson: the stage input starts at q (q_tmp = q at the start of
  // every step of the reference), stages ping-pong between two buffers and
  // the last one writes q itself.
  const double* in = q;
  double* stage_buf[2] = {ctx->ex_a.p, ctx->ex_b.p};
  for (int s = 0; s < 4; ++s) {
    SG_TRY(residual(in), "residual");
    double* out = (s == 3) ? q : stage_buf[s & 1];
    SG_TRY(launch_stage(ctx, q, R, cfl, 4.0 - s, out), "update");
    in = out;
  }
  return SG_OK;
}

int check_args(sg_ctx* ctx, int scheme, double cfl) {
  if (scheme != SG_SCHEME_FORWARD_EULER && scheme != SG_SCHEME_RK4_JAMESON)
    return sg_fail(ctx, SG_ERR_ARG, "scheme must be SG_SCHEME_FORWARD_EULER or SG_SCHEME_RK4_JAMESON");
  if (!(cfl > 0.0) || !std::isfinite(cfl)) return sg_fail(ctx, SG_ERR_ARG, "cfl must be positive and finite");
  return SG_OK;
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" {

int sg_step_explicit_dev(sg_ctx* ctx, double* q_dev, int scheme, double cfl, double* l2) {
  if (!ctx || !q_dev) return SG_ERR_ARG;
  int rc;
  if ((rc = check_args(ctx, scheme, cfl))) return rc;
  if ((rc = ensure_buffers(ctx))) return rc;
  if ((rc = enqueue_step(ctx, q_dev, scheme, cfl))) return rc;
  double h[10];
  SG_TRY(hipMemcpyAsync(h, ctx->l2.p, ctx->nvar * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "D2H");
  SG_TRY(hipStreamSynchronize(ctx->stream), "explicit step");
  if (l2) std::copy(h, h + ctx->nvar, l2);
  for (int k = 0; k < ctx->nvar; ++k)
    if (!std::isfinite(h[k])) return sg_fail(ctx, SG_ERR_NAN, "non-finite residual norm in an explicit step");
  return SG_OK;
}

int sg_solve_explicit(sg_ctx* ctx, double* q_aos, const sg_explicit_opts* o, sg_solve_report* rep, double* l2_hist) {
  if (!ctx || !q_aos || !o) return SG_ERR_ARG;
  int rc;
  if ((rc = check_args(ctx, o->scheme, o->cfl))) return rc;
  if (o->iteration_max < 0) return sg_fail(ctx, SG_ERR_ARG, "iteration_max must be >= 0");
  if ((rc = ensure_buffers(ctx))) return rc;
  const auto t0 = std::chrono::steady_clock::now();
  const int nv = ctx->nvar;
  const size_t ncell = ctx->ncell();
  double* q = ctx->q.p;
  SG_TRY(hipMemcpyAsync(ctx->aos_a.p, q_aos, ncell * nv * sizeof(double), hipMemcpyHostToDevice, ctx->stream), "H2D");
  SG_TRY(launch_aos_to_soa(ctx->geo.nic, ctx->geo.njc, ctx->g.ldi, nv, ctx->aos_a.p, q, ctx->stream), "aos->soa");
  sg_solve_report R{};
  double cfl = o->cfl, l2[10];
  int counter = 0;
  rc = SG_OK;
  for (;; ++counter) {
    if ((rc = enqueue_step(ctx, q, o->scheme, cfl))) return rc;
    SG_TRY(hipMemcpyAsync(l2, ctx->l2.p, nv * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "D2H");
    SG_TRY(hipStreamSynchronize(ctx->stream), "explicit step");
    if (counter == 0) std::copy(l2, l2 + nv, R.l2_initial);
    if (l2_hist) std::copy(l2, l2 + nv, l2_hist + (size_t)counter * nv);
    if (o->verbose && counter % o->verbose == 0)
      std::fprintf(stderr, "explicit %6d  cfl %.3e  |R| %.3e %.3e %.3e %.3e\n", counter, cfl, l2[0], l2[1], l2[2], l2[3]);
    bool finite = true, below = o->tolerance > 0.0;
    for (int k = 0; k < nv; ++k) {
      finite = finite && std::isfinite(l2[k]);
      below = below && l2[k] < o->tolerance;
    }
    if (!finite) { rc = sg_fail(ctx, SG_ERR_NAN, "non-finite residual norm in the explicit loop"); break; }
    if (below) { R.converged = 1; break; }
    if (counter > o->iteration_max) break;  // solver.cpp:136 (after the update)
    if (o->cfl_ramp && counter + 1 > o->cfl_ramp_iteration) cfl = std::min(std::pow(cfl, o->cfl_ramp_exponent), 1e12);
  }
  std::copy(l2, l2 + nv, R.l2_final);
  R.iterations = counter + 1;
  R.cfl_final = cfl;
  SG_TRY(launch_soa_to_aos(ctx->geo.nic, ctx->geo.njc, ctx->g.ldi, nv, q, ctx->aos_b.p, ctx->stream), "soa->aos");
  SG_TRY(hipMemcpyAsync(q_aos, ctx->aos_b.p, ncell * nv * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "D2H");
  SG_TRY(hipStreamSynchronize(ctx->stream), "download");
  R.wall_s = R.residual_s = seconds_since(t0);
  if (rep) *rep = R;
  return rc;
}

}  // extern "C"
