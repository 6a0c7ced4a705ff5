/*
 * structured_gpu_adapter.hpp -- C++ classes with the member signatures the
 * reference's Newton loop calls, over the C ABI of structured_gpu.h.
 * Header-only; templated on the caller's mesh and configuration types, whose
 * members are read by the reference's names (Mesh: ni, nj, nic, njc, xv, yv,
 * solution; Solution: q, dt, rhs, nq, ntrans; Config: freestream, solver,
 * geometry -- src/utils/mesh.h, src/solver/solution.h, src/utils/config.h).
 *
 *   TransportEquationGpu   the Spalart-Allmaras model as the reference's one
 *                          transport equation: the commented-out
 *                          TransportEquation<Tx,Tad>(mesh, config) of
 *                          src/model/eulerequation.cpp:98, with ntrans = 1
 *                          (src/solver/solution.cpp:9)
 *   EulerEquationGpu       EulerEquation<Tx,Tad>::calc_residual / calc_dt /
 *                          initialize (src/model/eulerequation.h:53-55,
 *                          eulerequation.cpp:262-291)
 *   LinearSolverGpu        preallocate / set_lhs / set_rhs / solve /
 *                          solve_and_update (src/linearsolver/ls_eigen.h:26-31)
 *
 * Errors are thrown as std::runtime_error with the library's message (the C
 * ABI itself never throws).
 */
#ifndef STRUCTURED_GPU_ADAPTER_HPP
#define STRUCTURED_GPU_ADAPTER_HPP

#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "structured_gpu.h"

namespace sg_adapter {

inline void check(int rc, const sg_ctx* ctx, const char* what) {
  if (rc != SG_OK) throw std::runtime_error(std::string(what) + ": " + sg_last_error(ctx));
}

// The SA model as the reference's transport-equation hook.  The solution
// must hold one transport unknown (Solution::ntrans = 1, the hook at
// src/solver/solution.cpp:9, so that q, dt and rhs have nq + 1 = 5 unknowns
// per cell, rho nu~ fifth); nu~_inf = nut_ratio * mu_inf / rho_inf.
template <class Mesh, class Config>
class TransportEquationGpu {
 public:
  TransportEquationGpu(std::shared_ptr<Mesh> mesh, std::shared_ptr<Config> config, double nut_ratio = 3.0)
      : nut_inf_(nut_ratio * config->freestream->mu_inf / config->freestream->rho_inf) {
    if (!(config->freestream->mu_inf > 1e-15)) throw std::runtime_error("the SA model needs a viscous freestream");
    const auto& sol = *mesh->solution;
    if (sol.ntrans != 1 || sol.q.size() != mesh->nic * mesh->njc * (sol.nq + 1))
      throw std::runtime_error("the SA model needs Solution::ntrans = 1 (q with nq + 1 unknowns per cell)");
  }
  double nut_inf() const { return nut_inf_; }
  static constexpr int ntrans = 1;

 private:
  double nut_inf_;
};

template <class Mesh, class Config>
class EulerEquationGpu {
 public:
  // bcs: the [[boundary]] list in file order (BoundaryContainer,
  // src/model/bc.cpp:465-526); transport: the SA model, or null (laminar).
  EulerEquationGpu(std::shared_ptr<Mesh> mesh, std::shared_ptr<Config> config, const std::vector<sg_bc>& bcs,
                   const TransportEquationGpu<Mesh, Config>* transport = nullptr)
      : mesh_(mesh) {
    const auto& f = *config->freestream;
    sg_freestream fs{f.rho_inf, f.u_inf, f.v_inf, f.p_inf, f.T_inf, f.mu_inf, f.pr_inf,
                     f.aoa * 180.0 / M_PI,  // Config keeps radians (src/utils/config.cpp:40)
                     transport ? transport->nut_inf() : 0.0};
    const auto& s = *config->solver;
    sg_opts o{(int)s.order, transport ? 1 : 0, (int)config->geometry->tail,
              s.flux == "ausm" ? SG_FLUX_AUSM : SG_FLUX_ROE, s.dpdx, s.dpdy};
    check(sg_create((int)mesh->ni, (int)mesh->nj, mesh->xv.data(), mesh->yv.data(), &fs, bcs.data(), (int)bcs.size(),
                    &o, &ctx_),
          nullptr, "sg_create");
  }
  ~EulerEquationGpu() { sg_destroy(ctx_); }
  EulerEquationGpu(const EulerEquationGpu&) = delete;
  EulerEquationGpu& operator=(const EulerEquationGpu&) = delete;

  // eulerequation.h:53 -- AoS q[nic][njc][nq+ntrans] in, R = dq/dt out
  template <class ArrayQ, class ArrayR>
  void calc_residual(const ArrayQ& a_q, ArrayR& a_rhs, bool /*lhs*/ = false) {
    check(sg_residual(ctx_, a_q.data(), a_rhs.data(), nullptr), ctx_, "sg_residual");
  }
  // eulerequation.h:54 -- local time steps into Solution::dt
  void calc_dt(double cfl) {
    auto& sol = *mesh_->solution;
    check(sg_dt(ctx_, sol.q.data(), cfl, sol.dt.data()), ctx_, "sg_dt");
  }
  // eulerequation.cpp:262-277 -- freestream state into Solution::q
  void initialize() { check(sg_initialize(ctx_, mesh_->solution->q.data()), ctx_, "sg_initialize"); }
  sg_ctx* handle() const { return ctx_; }

 private:
  std::shared_ptr<Mesh> mesh_;
  sg_ctx* ctx_ = nullptr;
};

// The linear-solver interface, matrix-free: set_lhs linearises at the
// solution's current q with its dt (the COO arguments of the reference's
// ADOL-C Jacobian are not needed and may be null).
template <class Mesh>
class LinearSolverGpu {
 public:
  LinearSolverGpu(std::shared_ptr<Mesh> mesh, sg_ctx* ctx, int precond = 6, int restart = 30, int maxit = 60,
                  double rtol = 1e-3, int lhs_order = 1, int pc_width = 32)
      : mesh_(mesh), ctx_(ctx), precond_(precond), restart_(restart), maxit_(maxit), rtol_(rtol),
        lhs_order_(lhs_order), pc_width_(pc_width) {}
  void preallocate(int /*nnz*/) {
    check(sg_ls_preallocate(ctx_, restart_, precond_, maxit_, rtol_), ctx_, "sg_ls_preallocate");
    check(sg_ls_set_pc(ctx_, lhs_order_, 1.0, pc_width_), ctx_, "sg_ls_set_pc");
  }
  void set_lhs(int /*nnz*/, unsigned* /*rind*/, unsigned* /*cind*/, double* /*values*/) {
    auto& sol = *mesh_->solution;
    check(sg_ls_set_lhs(ctx_, sol.q.data(), sol.dt.data()), ctx_, "sg_ls_set_lhs");
  }
  void set_rhs(double* rhs) { check(sg_ls_set_rhs(ctx_, rhs), ctx_, "sg_ls_set_rhs"); }
  void solve() {}  // solve_and_update solves
  // ls_eigen.cpp:65-71: q <- q + UR dq; a Krylov solve that stops short of
  // its tolerance still updates q (as an inexact Newton step)
  void solve_and_update(double* q, double under_relaxation) {
    const int rc = sg_ls_solve_and_update(ctx_, q, under_relaxation);
    if (rc != SG_ERR_NOCONV) check(rc, ctx_, "sg_ls_solve_and_update");
  }

 private:
  std::shared_ptr<Mesh> mesh_;
  sg_ctx* ctx_;
  int precond_, restart_, maxit_;
  double rtol_;
  int lhs_order_, pc_width_;
};

}  // namespace sg_adapter

#endif  // STRUCTURED_GPU_ADAPTER_HPP
