// Adjoint-driven field inversion of the SA production multiplier beta:
// limited-memory BFGS (Nocedal 1980, two-loop recursion) with a projected
// Armijo backtracking line search on
//   J(beta) = F(q(beta)) + lambda * sum_c (beta_c - 1)^2,
// F the wall skin-friction mismatch of sg_objective_cf.  Every trial point
// re-converges the primal by Newton-Krylov from the last accepted state; every
// accepted point costs one adjoint solve for dF/dbeta.  Host driver over the
// public entry points only (sg_set_beta, sg_solve, sg_objective_cf,
// sg_adjoint_gradient).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <vector>

#include "sg_context.hpp"

namespace {

struct Clock {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double s() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (size_t k = 0; k < a.size(); ++k) s += a[k] * b[k];
  return s;
}

}  // namespace

extern "C" int sg_invert(sg_ctx* ctx, double* q_aos, const double* cf_target, double* beta,
                         const sg_inversion_opts* o, sg_inversion_report* rep, double* J_hist) {
  if (!ctx || !q_aos || !cf_target || !beta || !o) return SG_ERR_ARG;
  if (!ctx->sa) return sg_fail(ctx, SG_ERR_ARG, "field inversion needs the SA model");
  if (ctx->clo.kind != sg::CLOSURE_FIELD) return sg_fail(ctx, SG_ERR_ARG, "field inversion needs the field closure");
  if (o->iterations < 0 || o->memory < 1 || o->max_backtracks < 0 || !(o->beta_max > o->beta_min))
    return sg_fail(ctx, SG_ERR_ARG, "bad inversion options");
  Clock wall;
  double t_primal = 0.0, t_adjoint = 0.0;
  const size_t nc = ctx->ncell(), nq = nc * ctx->nvar;
  const double lam = o->reg;
  sg_inversion_report R{};
  std::vector<double> x(beta, beta + nc), q(q_aos, q_aos + nq), g(nc), ga(nc);
  for (double& b : x) b = std::min(o->beta_max, std::max(o->beta_min, b));

  auto prior = [&](const std::vector<double>& b) {
    double s = 0.0;
    for (double v : b) s += (v - 1.0) * (v - 1.0);
    return lam * s;
  };
  // primal at beta b from state qs (in place); returns rc of sg_solve
  auto primal = [&](const std::vector<double>& b, std::vector<double>& qs, sg_solve_report* sr) {
    Clock c;
    int rc = sg_set_beta(ctx, b.data());
    if (rc == SG_OK) rc = sg_solve(ctx, qs.data(), &o->newton, sr);
    // a solve that stalled (stall_window) sits at its roundoff floor, within
    // 1000x of the tolerance: a converged state for the objective
    if (rc == SG_ERR_NOCONV && sr->stalled) rc = SG_OK;
    t_primal += c.s();
    return rc;
  };
  auto gradient = [&](const std::vector<double>& qs, const std::vector<double>& b, double* F) {
    Clock c;
    sg_adjoint_report ar{};
    int rc = sg_adjoint_gradient(ctx, qs.data(), cf_target, ga.data(), F, &o->adjoint, &ar);
    t_adjoint += c.s();
    if (rc != SG_OK && rc != SG_ERR_NOCONV) return rc;
    R.adjoint_unconverged += rc == SG_ERR_NOCONV ? 1 : 0;
    R.adjoint_relres_max = std::max(R.adjoint_relres_max, ar.relres);
    R.adjoint_iterations += ar.iterations;
    if (o->newton.verbose)
      std::printf("adjoint   %d iterations, relres %.3e%s, F %.6e\n", ar.iterations, ar.relres,
                  rc == SG_ERR_NOCONV ? " (not converged)" : "", *F);
    for (size_t k = 0; k < nc; ++k) g[k] = ga[k] + 2.0 * lam * (b[k] - 1.0);
    return SG_OK;
  };

  sg_solve_report sr{};
  int rc = primal(x, q, &sr);
  ++R.evaluations;
  if (rc != SG_OK) return sg_fail(ctx, rc, "primal solve at the initial beta did not converge");
  double F = 0.0;
  if ((rc = gradient(q, x, &F))) return rc;
  double J = F + prior(x);
  R.F0 = F;
  R.J0 = J;
  if (J_hist) J_hist[0] = J;

  std::deque<std::vector<double>> S, Y;
  std::deque<double> RHO;
  std::vector<double> d(nc), xn(nc), qn(nq), alpha_k;
  int it = 0;
  for (; it < o->iterations; ++it) {
    // two-loop recursion: d = -H g
    d = g;
    alpha_k.assign(S.size(), 0.0);
    for (int m = (int)S.size() - 1; m >= 0; --m) {
      alpha_k[m] = RHO[m] * dot(S[m], d);
      for (size_t k = 0; k < nc; ++k) d[k] -= alpha_k[m] * Y[m][k];
    }
    double gamma;
    if (!S.empty()) {
      gamma = dot(S.back(), Y.back()) / dot(Y.back(), Y.back());
    } else {
      double gmax = 0.0;
      for (double v : g) gmax = std::max(gmax, std::fabs(v));
      gamma = gmax > 0.0 ? o->step0 / gmax : 0.0;
    }
    for (double& v : d) v *= gamma;
    for (size_t m = 0; m < S.size(); ++m) {
      const double b = RHO[m] * dot(Y[m], d);
      for (size_t k = 0; k < nc; ++k) d[k] += (alpha_k[m] - b) * S[m][k];
    }
    for (double& v : d) v = -v;
    double gd = dot(g, d);
    if (!(gd < 0.0)) {  // not a descent direction: restart from scaled steepest descent
      S.clear(); Y.clear(); RHO.clear();
      double gmax = 0.0;
      for (double v : g) gmax = std::max(gmax, std::fabs(v));
      if (gmax == 0.0) break;
      for (size_t k = 0; k < nc; ++k) d[k] = -g[k] * (o->step0 / gmax);
      gd = dot(g, d);
    }
    // projected Armijo backtracking; if the L-BFGS direction finds no
    // acceptable point, the search is repeated once along scaled -g with the
    // correction pairs dropped, and only a failure along -g ends the run
    double Jn = 0.0, Fn = 0.0;
    bool ok = false;
    for (int attempt = 0; attempt < 2 && !ok; ++attempt) {
      if (attempt == 1) {
        if (S.empty()) break;  // the direction was -g already
        S.clear(); Y.clear(); RHO.clear();
        double gmax = 0.0;
        for (double v : g) gmax = std::max(gmax, std::fabs(v));
        for (size_t k = 0; k < nc; ++k) d[k] = -g[k] * (o->step0 / gmax);
        ++R.steepest_restarts;
      }
      double a = 1.0;
      for (int bt = 0; bt <= o->max_backtracks; ++bt, a *= 0.25) {
        for (size_t k = 0; k < nc; ++k) xn[k] = std::min(o->beta_max, std::max(o->beta_min, x[k] + a * d[k]));
        qn = q;
        int pr = primal(xn, qn, &sr);
        ++R.evaluations;
        if (pr == SG_ERR_NAN || pr == SG_ERR_NOCONV) { ++R.failed_solves; continue; }
        if (pr != SG_OK) return pr;
        if ((rc = sg_objective_cf(ctx, qn.data(), cf_target, &Fn, nullptr))) return rc;
        Jn = Fn + prior(xn);
        double gs = 0.0;
        for (size_t k = 0; k < nc; ++k) gs += g[k] * (xn[k] - x[k]);
        if (o->newton.verbose)
          std::printf("linesearch step %.3e  J %.9e  (J0 %.9e, Armijo bound %.9e)\n", a, Jn, J, J + 1e-4 * gs);
        if (Jn <= J + 1e-4 * gs) { ok = true; break; }
      }
    }
    if (!ok) break;
    std::vector<double> gold = g;
    double Fg = 0.0;
    if ((rc = gradient(qn, xn, &Fg))) return rc;
    std::vector<double> s(nc), y(nc);
    for (size_t k = 0; k < nc; ++k) { s[k] = xn[k] - x[k]; y[k] = g[k] - gold[k]; }
    const double sy = dot(s, y);
    if (sy > 1e-12 * std::sqrt(dot(s, s) * dot(y, y))) {
      S.push_back(std::move(s)); Y.push_back(std::move(y)); RHO.push_back(1.0 / sy);
      if ((int)S.size() > o->memory) { S.pop_front(); Y.pop_front(); RHO.pop_front(); }
    }
    x = xn;
    q = qn;
    J = Jn;
    F = Fn;
    if (J_hist) J_hist[it + 1] = J;
  }
  // leave the context at the final beta
  if ((rc = sg_set_beta(ctx, x.data()))) return rc;
  std::copy(x.begin(), x.end(), beta);
  std::copy(q.begin(), q.end(), q_aos);
  R.iterations = it;
  R.J = J;
  R.F = F;
  R.wall_s = wall.s();
  R.primal_s = t_primal;
  R.adjoint_s = t_adjoint;
  if (rep) *rep = R;
  return SG_OK;
}
