// CPU ORACLE — test infrastructure only (see sg_oracle.hpp).  extern "C"
// wrappers so that tests/ and the benchmark's CPU-baseline leg can call the
// restatement through ctypes.  It shares only the plain-data structs of
// include/structured_gpu.h with the product.
#include <omp.h>
#include <chrono>
#include <memory>

#include "../include/structured_gpu.h"
#include "sg_oracle.hpp"

using namespace oracle;

namespace {

struct Setup {
  Geometry geo;
  Params prm;
  std::vector<Bc> bcs;
};

long resolve_end(long end, int face, size_t nic, size_t njc) {
  if (end >= 0) return end;
  return (face == FACE_BOTTOM || face == FACE_TOP) ? (long)nic + 2 + end : (long)njc + 2 + end;
}

void make_setup(Setup& s, int ni, int nj, const double* xv, const double* yv, const sg_freestream* fs,
                const sg_bc* bcs, int nbc, const sg_opts* o) {
  s.geo.build((size_t)ni, (size_t)nj, xv, yv);
  s.prm.rho_inf = fs->rho_inf; s.prm.u_inf = fs->u_inf; s.prm.v_inf = fs->v_inf;
  s.prm.p_inf = fs->p_inf; s.prm.T_inf = fs->T_inf; s.prm.mu_inf = fs->mu_inf;
  s.prm.pr_inf = fs->pr_inf; s.prm.aoa = fs->aoa_deg * M_PI / 180.0; s.prm.nut_inf = fs->nut_inf;
  s.prm.dpdx = o->dpdx; s.prm.dpdy = o->dpdy; s.prm.order = o->order; s.prm.sa = o->turbulence ? 1 : 0;
  s.prm.flux = o->flux;
  s.bcs.clear();
  for (int b = 0; b < nbc; ++b) {
    Bc c;
    c.type = bcs[b].type; c.face = bcs[b].face; c.start = (long)bcs[b].start;
    c.end = resolve_end((long)bcs[b].end, c.face, s.geo.nic, s.geo.njc);
    c.u = bcs[b].u; c.v = bcs[b].v; c.T = bcs[b].T;
    s.bcs.push_back(c);
  }
}

Closure g_closure;  // set by og_set_closure, used by every following call

template <class T>
void init_res(Residual<T>& r, const Setup& s, const double* wd, const double* beta) {
  r.init(&s.geo, s.prm, s.bcs);
  r.clo = g_closure;
  const size_t n = s.geo.nic * s.geo.njc;
  if (wd) r.wall_dist.assign(wd, wd + n);
  else { r.wall_dist.assign(n, 1.0); if (s.prm.sa) wall_distance(s.geo, s.bcs, r.wall_dist.data()); }
  if (beta) r.beta.assign(beta, beta + n);
}

}  // namespace

extern "C" {

// Closure of beta for the following calls: kind 0 field, 1 analytic (coef[7]),
// 2 MLP (params: SG_MLP_NPARAM floats, W1|b1|W2|b2|W3|b3).
int og_set_closure(int kind, const double* coef, const float* params) {
  Closure c;
  c.kind = kind;
  if (kind == 1) std::copy(coef, coef + 7, c.coef);
  if (kind == 2) {
    c.W1.assign(params, params + 320);
    c.b1.assign(params + 320, params + 384);
    c.W2.resize(4096);
    for (int k = 0; k < 4096; ++k) c.W2[k] = bf16_round((double)params[384 + k]);
    c.b2.assign(params + 4480, params + 4544);
    c.W3.assign(params + 4544, params + 4608);
    c.b3 = params[4608];
  }
  g_closure = c;
  return 0;
}

int og_metrics(int ni, int nj, const double* xv, const double* yv, double* nchi, double* neta, double* vol,
               double* xc, double* yc, double* ds_chi, double* ds_eta) {
  Geometry g;
  g.build((size_t)ni, (size_t)nj, xv, yv);
  std::copy(g.normal_chi.a.begin(), g.normal_chi.a.end(), nchi);
  std::copy(g.normal_eta.a.begin(), g.normal_eta.a.end(), neta);
  std::copy(g.volume.a.begin(), g.volume.a.end(), vol);
  std::copy(g.xc.a.begin(), g.xc.a.end(), xc);
  std::copy(g.yc.a.begin(), g.yc.a.end(), yc);
  std::copy(g.ds_chi.a.begin(), g.ds_chi.a.end(), ds_chi);
  std::copy(g.ds_eta.a.begin(), g.ds_eta.a.end(), ds_eta);
  return 0;
}

// R(q) in the reference layout; l2 (nvar entries) may be NULL.  wd / beta may
// be NULL (computed wall distance, beta = 1).
int og_residual(int ni, int nj, const double* xv, const double* yv, const sg_freestream* fs, const sg_bc* bcs,
                int nbc, const sg_opts* o, const double* wd, const double* beta, const double* q, double* r,
                double* l2) {
  Setup s;
  make_setup(s, ni, nj, xv, yv, fs, bcs, nbc, o);
  Residual<double> res;
  init_res(res, s, wd, beta);
  res.evaluate(q, r);
  if (l2) l2norm(s.geo.nic * s.geo.njc, (size_t)s.prm.nvar(), r, l2);
  return 0;
}

// Ghosted primitives after primvars and the boundary list (src/model/fluid.cpp:50-67,
// src/model/bc.cpp:430-433): out[6][nic+2][njc+2] = rho, u, v, p, T, nu~
// (nu~ zero for the laminar model); cells no boundary writes are NaN.
int og_ghosts(int ni, int nj, const double* xv, const double* yv, const sg_freestream* fs, const sg_bc* bcs, int nbc,
              const sg_opts* o, const double* q, double* out) {
  Setup s;
  make_setup(s, ni, nj, xv, yv, fs, bcs, nbc, o);
  Residual<double> res;
  init_res(res, s, nullptr, nullptr);
  const double nan = std::numeric_limits<double>::quiet_NaN();
  for (Arr2<double>* a : {&res.rho, &res.u, &res.v, &res.p, &res.Tt, &res.nt}) std::fill(a->a.begin(), a->a.end(), nan);
  res.primvars(q);
  res.apply_bcs();
  const size_t m = (s.geo.nic + 2) * (s.geo.njc + 2);
  int k = 0;
  for (Arr2<double>* a : {&res.rho, &res.u, &res.v, &res.p, &res.Tt, &res.nt}) {
    if (a == &res.nt && !s.prm.sa) std::fill(out + k * m, out + (k + 1) * m, 0.0);
    else std::copy(a->a.begin(), a->a.end(), out + k * m);
    ++k;
  }
  return 0;
}

// Features and beta per cell at q: out[nic][njc][6].
int og_closure_eval(int ni, int nj, const double* xv, const double* yv, const sg_freestream* fs, const sg_bc* bcs,
                    int nbc, const sg_opts* o, const double* wd, const double* beta, const double* q, double* out) {
  Setup s;
  make_setup(s, ni, nj, xv, yv, fs, bcs, nbc, o);
  Residual<double> res;
  init_res(res, s, wd, beta);
  res.aux = out;
  std::vector<double> r(s.geo.nic * s.geo.njc * s.prm.nvar());
  res.evaluate(q, r.data());
  return 0;
}

// J.v by one forward dual sweep of the same residual.
int og_jv(int ni, int nj, const double* xv, const double* yv, const sg_freestream* fs, const sg_bc* bcs, int nbc,
          const sg_opts* o, const double* wd, const double* beta, const double* q, const double* v, double* r,
          double* jv) {
  Setup s;
  make_setup(s, ni, nj, xv, yv, fs, bcs, nbc, o);
  Residual<Dual<1>> res;
  init_res(res, s, wd, beta);
  const size_t n = s.geo.nic * s.geo.njc * (size_t)s.prm.nvar();
  std::vector<Dual<1>> qd(n), rd(n);
  for (size_t m = 0; m < n; ++m) { qd[m].v = q[m]; qd[m].d[0] = v[m]; }
  res.evaluate(qd.data(), rd.data());
  for (size_t m = 0; m < n; ++m) { if (r) r[m] = rd[m].v; jv[m] = rd[m].d[0]; }
  return 0;
}

// dR/dbeta . w (the beta-direction derivative), by a dual in the beta field.
int og_dbeta(int ni, int nj, const double* xv, const double* yv, const sg_freestream* fs, const sg_bc* bcs, int nbc,
             const sg_opts* o, const double* wd, const double* beta, const double* q, const double* w, double* out) {
  Setup s;
  make_setup(s, ni, nj, xv, yv, fs, bcs, nbc, o);
  const size_t nc = s.geo.nic * s.geo.njc, nv = (size_t)s.prm.nvar();
  // beta enters only the SA production, linearly: dR/dbeta.w = P_c/beta_c*w_c/V_c
  // is evaluated as a finite combination of two residuals, which is exact for
  // a linear dependence up to rounding.
  std::vector<double> b0(nc, 1.0), b1(nc);
  if (beta) b0.assign(beta, beta + nc);
  for (size_t c = 0; c < nc; ++c) b1[c] = b0[c] + w[c];
  std::vector<double> r0(nc * nv), r1(nc * nv);
  Residual<double> res;
  init_res(res, s, wd, b0.data());
  res.evaluate(q, r0.data());
  init_res(res, s, wd, b1.data());
  res.evaluate(q, r1.data());
  for (size_t m = 0; m < nc * nv; ++m) out[m] = r1[m] - r0[m];
  return 0;
}

// Dense Jacobian (row-major, n x n) by n forward dual sweeps; tiny grids only.
int og_jacobian(int ni, int nj, const double* xv, const double* yv, const sg_freestream* fs, const sg_bc* bcs,
                int nbc, const sg_opts* o, const double* wd, const double* beta, const double* q, double* J) {
  Setup s;
  make_setup(s, ni, nj, xv, yv, fs, bcs, nbc, o);
  Residual<Dual<1>> res;
  init_res(res, s, wd, beta);
  const size_t n = s.geo.nic * s.geo.njc * (size_t)s.prm.nvar();
  std::vector<Dual<1>> qd(n), rd(n);
  for (size_t c = 0; c < n; ++c) {
    for (size_t m = 0; m < n; ++m) { qd[m].v = q[m]; qd[m].d[0] = (m == c) ? 1.0 : 0.0; }
    res.evaluate(qd.data(), rd.data());
    for (size_t m = 0; m < n; ++m) J[m * n + c] = rd[m].d[0];
  }
  return 0;
}

int og_dt(int ni, int nj, const double* xv, const double* yv, const sg_freestream* fs, const sg_bc* bcs, int nbc,
          const sg_opts* o, const double* q, double cfl, double* dt) {
  Setup s;
  make_setup(s, ni, nj, xv, yv, fs, bcs, nbc, o);
  local_dt(s.geo, s.prm, q, cfl, dt);
  return 0;
}

int og_surface(int ni, int nj, const double* xv, const double* yv, const sg_freestream* fs, const sg_bc* bcs,
               int nbc, const sg_opts* o, const double* wd, const double* beta, const double* q, double* x,
               double* cp, double* cf, double* forces) {
  Setup s;
  make_setup(s, ni, nj, xv, yv, fs, bcs, nbc, o);
  Residual<double> res;
  init_res(res, s, wd, beta);
  std::vector<double> r(s.geo.nic * s.geo.njc * (size_t)s.prm.nvar());
  res.evaluate(q, r.data());
  surface(s.geo, s.prm, res, q, (long)o->tail, x, cp, cf, forces);
  return 0;
}

int og_wall_distance(int ni, int nj, const double* xv, const double* yv, const sg_freestream* fs, const sg_bc* bcs,
                     int nbc, const sg_opts* o, double* d) {
  Setup s;
  make_setup(s, ni, nj, xv, yv, fs, bcs, nbc, o);
  wall_distance(s.geo, s.bcs, d);
  return 0;
}

// CPU baseline of the adjoint gradient (a bounded sample, for the benchmark
// only), on `threads` OpenMP threads:
//   out[0]: seconds to extract the Jacobian's 13 stencil blocks per cell,
//           by 5 x 5 cell colouring of dual sweeps of this residual (all nv
//           directions of a colour in one sweep);
//   out[1]: seconds per iteration of the transposed Krylov loop's work, the
//           average of `iters` iterations: J^T psi gathered from the blocks,
//           a block-Jacobi preconditioner apply and CGS2 orthogonalisation
//           (two multi-dots and two multi-axpys) against `basis` vectors.
int og_adjoint_sample(int ni, int nj, const double* xv, const double* yv, const sg_freestream* fs, const sg_bc* bcs,
                      int nbc, const sg_opts* o, const double* q, int threads, int iters, int basis, double* out) {
  using clk = std::chrono::steady_clock;
  if (threads > 0) omp_set_num_threads(threads);
  Setup s;
  make_setup(s, ni, nj, xv, yv, fs, bcs, nbc, o);
  const long nic = (long)s.geo.nic, njc = (long)s.geo.njc;
  const int nv = s.prm.nvar();
  const size_t nc = (size_t)nic * njc, n = nc * nv;
  static const int DI[13] = {0, 0, 0, -1, 1, -1, 1, -1, 1, 0, 0, -2, 2};
  static const int DJ[13] = {0, -1, 1, 0, 0, -1, -1, 1, 1, -2, 2, 0, 0};
  // (1) blocks[c][s][k][m] = dR_(c,k) / dq_(c + off_s, m)
  std::vector<double> blk(nc * 13 * nv * nv, 0.0);
  auto t0 = clk::now();
  {
    Residual<Dual<5>> res;
    init_res(res, s, nullptr, nullptr);
    std::vector<Dual<5>> qd(n), rd(n);
    for (int a = 0; a < 5; ++a)
      for (int b = 0; b < 5; ++b) {
        for (long i = 0; i < nic; ++i)
          for (long j = 0; j < njc; ++j) {
            const bool seed = i % 5 == a && j % 5 == b;
            for (int m = 0; m < nv; ++m) {
              Dual<5>& x = qd[((size_t)i * njc + j) * nv + m];
              x = Dual<5>(q[((size_t)i * njc + j) * nv + m]);
              if (seed) x.d[m] = 1.0;
            }
          }
        res.evaluate(qd.data(), rd.data());
#pragma omp parallel for
        for (long i = 0; i < nic; ++i)
          for (long j = 0; j < njc; ++j)
            for (int sl = 0; sl < 13; ++sl) {
              const long ii = i + DI[sl], jj = j + DJ[sl];
              if (ii < 0 || ii >= nic || jj < 0 || jj >= njc || ii % 5 != a || jj % 5 != b) continue;
              const size_t c = (size_t)i * njc + j;
              for (int k = 0; k < nv; ++k)
                for (int m = 0; m < nv; ++m) blk[((c * 13 + sl) * nv + k) * nv + m] = rd[c * nv + k].d[m];
            }
      }
  }
  out[0] = std::chrono::duration<double>(clk::now() - t0).count();
  // block-Jacobi inverse of the diagonal blocks (set-up, not timed)
  std::vector<double> dinv(nc * nv * nv);
#pragma omp parallel for
  for (long c = 0; c < (long)nc; ++c) {
    double a[5][10];
    for (int r = 0; r < nv; ++r)
      for (int m = 0; m < 2 * nv; ++m) a[r][m] = m < nv ? -blk[(((size_t)c * 13) * nv + r) * nv + m] + (r == m ? 1e3 : 0.0)
                                                        : (m - nv == r ? 1.0 : 0.0);
    for (int p = 0; p < nv; ++p) {
      const double ip = 1.0 / a[p][p];
      for (int m = 0; m < 2 * nv; ++m) a[p][m] *= ip;
      for (int r = 0; r < nv; ++r)
        if (r != p) {
          const double f = a[r][p];
          for (int m = 0; m < 2 * nv; ++m) a[r][m] -= f * a[p][m];
        }
    }
    for (int r = 0; r < nv; ++r)
      for (int m = 0; m < nv; ++m) dinv[((size_t)c * nv + r) * nv + m] = a[r][nv + m];
  }
  // (2) Krylov iterations' work
  const int nb = basis > 0 ? basis : 1;
  std::vector<double> V((size_t)(nb + 1) * n), z(n), w(n), h(nb), h2(nb);
  for (size_t m = 0; m < V.size(); ++m) V[m] = std::sin(0.001 * (double)m);
  t0 = clk::now();
  for (int it = 0; it < iters; ++it) {
    const double* v = V.data() + (size_t)(it % nb) * n;
#pragma omp parallel for
    for (long c = 0; c < (long)nc; ++c)
      for (int r = 0; r < nv; ++r) {
        double acc = 0.0;
        for (int m = 0; m < nv; ++m) acc += dinv[((size_t)c * nv + r) * nv + m] * v[(size_t)c * nv + m];
        z[(size_t)c * nv + r] = acc;
      }
    // w = (J^T z)_col: the rows of column cell c are c - off_s, block s of that row
#pragma omp parallel for
    for (long i = 0; i < nic; ++i)
      for (long j = 0; j < njc; ++j) {
        const size_t c = (size_t)i * njc + j;
        double acc[5] = {0, 0, 0, 0, 0};
        for (int sl = 0; sl < 13; ++sl) {
          const long ri = i - DI[sl], rj = j - DJ[sl];
          if (ri < 0 || ri >= nic || rj < 0 || rj >= njc) continue;
          const size_t r = (size_t)ri * njc + rj;
          for (int k = 0; k < nv; ++k)
            for (int m = 0; m < nv; ++m) acc[m] += blk[((r * 13 + sl) * nv + k) * nv + m] * z[r * nv + k];
        }
        for (int m = 0; m < nv; ++m) w[c * nv + m] = acc[m];
      }
    for (int pass = 0; pass < 2; ++pass) {
      double* hh = pass ? h2.data() : h.data();
      for (int b = 0; b < nb; ++b) {
        const double* vb = V.data() + (size_t)b * n;
        double acc = 0.0;
#pragma omp parallel for reduction(+ : acc)
        for (long m = 0; m < (long)n; ++m) acc += vb[m] * w[m];
        hh[b] = acc;
      }
#pragma omp parallel for
      for (long m = 0; m < (long)n; ++m) {
        double acc = w[m];
        for (int b = 0; b < nb; ++b) acc -= hh[b] * V[(size_t)b * n + m];
        w[m] = acc;
      }
    }
  }
  out[1] = std::chrono::duration<double>(clk::now() - t0).count() / (iters > 0 ? iters : 1);
  out[2] = w[n / 2] + h[0] + h2[0];  // keeps the work observable
  return 0;
}

// CPU baseline timing: median wall time (seconds) of `reps` residual
// evaluations after `warm` warm-ups, with `threads` OpenMP threads.
double og_time_residual(int ni, int nj, const double* xv, const double* yv, const sg_freestream* fs,
                        const sg_bc* bcs, int nbc, const sg_opts* o, const double* q, int warm, int reps,
                        int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  Setup s;
  make_setup(s, ni, nj, xv, yv, fs, bcs, nbc, o);
  Residual<double> res;
  init_res(res, s, nullptr, nullptr);
  std::vector<double> r(s.geo.nic * s.geo.njc * (size_t)s.prm.nvar());
  for (int w = 0; w < warm; ++w) res.evaluate(q, r.data());
  std::vector<double> t(reps > 0 ? reps : 1);
  for (int k = 0; k < reps; ++k) {
    auto t0 = std::chrono::steady_clock::now();
    res.evaluate(q, r.data());
    t[k] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  std::sort(t.begin(), t.end());
  return t[t.size() / 2];
}

int og_max_threads() { return omp_get_max_threads(); }

}  // extern "C"
