// CPU ORACLE — test infrastructure only.
//
// A plain C++ restatement of the reference solver's residual path
// (anandpratap/structured, src/model + src/utils/mesh.cpp), kept in the
// reference's own pass structure: primitive recovery into ghosted arrays,
// boundary conditions applied in list order, MUSCL per direction, mu/k over
// every ghosted cell, face Green-Gauss gradients, face averages, Roe or AUSM
// flux, Green-Gauss viscous flux, assembly, division by the cell volume.  It is
// templated on the scalar so that a forward dual number gives the exact
// directional derivative J.v of this same discretisation.
//
// Only tests/, the smoke check and the CPU-baseline leg of the benchmark may
// load this code, and only as the checker.  The product path (the HIP
// library) never calls it.
//
// Parity status per function (how each is pinned):
//   metrics, primvars, BCs (freestream/outflow/slipwall/wall/isothermal/
//   periodic/wake), MUSCL 1st/2nd order, Roe, AUSM, mu/k, gradients, face
//   averages, viscous flux, assembly, dt, L2 norm, surface Cp/Cf/forces:
//       pinned against the compiled reference (oracle/_ref) on the same
//       grids and states (golden fixtures, including the reference's own
//       channel and NACA 0012 examples), and by closed forms in
//       tests/test_oracle_*.py.  The explicit steps composed from these
//       (tests/oracle_ffi.py explicit_solve) are pinned against the
//       reference's own Solver::solve.
//   Spalart-Allmaras terms, beta field, wall objective, MLP:
//       the reference has none of these, so no reference output exists to
//       compare with.  Each is pinned by value instead: the published
//       SA-noft2 definition in closed form (production/destruction, the
//       diffusion, c_b2 and density-gradient terms on linear profiles, the
//       eddy viscosity and conductivity in the viscous fluxes, the ghost
//       values per boundary type: tests/test_oracle_sa.py and
//       tests/test_oracle_pins.py), the wall distance (y_c on the flat plate,
//       brute-force point-to-segment distance on the bump), the closure
//       features on a uniform shear, the MLP against a numpy model, and the
//       objective F against its definition (tests/test_gpu_contract.py);
//       central finite differences check the tangents.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <limits>
#include <vector>

namespace oracle {

constexpr double kGamma = 1.4;  // src/common.h:40

// ---------------------------------------------------------------- scalars --
template <int N>
struct Dual {
  double v;
  double d[N];
  Dual() : v(0.0) { for (int k = 0; k < N; ++k) d[k] = 0.0; }
  Dual(double x) : v(x) { for (int k = 0; k < N; ++k) d[k] = 0.0; }
};
template <int N> inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) { Dual<N> r(a.v + b.v); for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k]; return r; }
template <int N> inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) { Dual<N> r(a.v - b.v); for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k]; return r; }
template <int N> inline Dual<N> operator-(const Dual<N>& a) { Dual<N> r(-a.v); for (int k = 0; k < N; ++k) r.d[k] = -a.d[k]; return r; }
template <int N> inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) { Dual<N> r(a.v * b.v); for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k]; return r; }
template <int N> inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) { Dual<N> r(a.v / b.v); for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v; return r; }
template <int N> inline Dual<N> operator+(const Dual<N>& a, double b) { return a + Dual<N>(b); }
template <int N> inline Dual<N> operator+(double a, const Dual<N>& b) { return Dual<N>(a) + b; }
template <int N> inline Dual<N> operator-(const Dual<N>& a, double b) { return a - Dual<N>(b); }
template <int N> inline Dual<N> operator-(double a, const Dual<N>& b) { return Dual<N>(a) - b; }
template <int N> inline Dual<N> operator*(const Dual<N>& a, double b) { Dual<N> r(a.v * b); for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b; return r; }
template <int N> inline Dual<N> operator*(double a, const Dual<N>& b) { return b * a; }
template <int N> inline Dual<N> operator/(const Dual<N>& a, double b) { Dual<N> r(a.v / b); for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / b; return r; }
template <int N> inline Dual<N> operator/(double a, const Dual<N>& b) { return Dual<N>(a) / b; }
template <int N> inline Dual<N>& operator+=(Dual<N>& a, const Dual<N>& b) { a = a + b; return a; }
template <int N> inline Dual<N>& operator+=(Dual<N>& a, double b) { a.v += b; return a; }
template <int N> inline Dual<N>& operator-=(Dual<N>& a, const Dual<N>& b) { a = a - b; return a; }
template <int N> inline Dual<N>& operator/=(Dual<N>& a, const Dual<N>& b) { a = a / b; return a; }
template <int N> inline Dual<N>& operator/=(Dual<N>& a, double b) { a = a / b; return a; }
template <int N> inline bool operator<(const Dual<N>& a, const Dual<N>& b) { return a.v < b.v; }
template <int N> inline bool operator>(const Dual<N>& a, const Dual<N>& b) { return a.v > b.v; }
template <int N> inline bool operator<(const Dual<N>& a, double b) { return a.v < b; }
template <int N> inline bool operator>(const Dual<N>& a, double b) { return a.v > b; }
template <int N> inline bool operator>=(const Dual<N>& a, double b) { return a.v >= b; }
template <int N> inline bool operator<=(const Dual<N>& a, double b) { return a.v <= b; }
template <int N> inline Dual<N> sqrt(const Dual<N>& a) { Dual<N> r(std::sqrt(a.v)); for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / (2.0 * r.v); return r; }
template <int N> inline Dual<N> fabs(const Dual<N>& a) { return a.v < 0.0 ? -a : a; }
template <int N> inline Dual<N> pow(const Dual<N>& a, double e) { Dual<N> r(std::pow(a.v, e)); double dd = e * std::pow(a.v, e - 1.0); for (int k = 0; k < N; ++k) r.d[k] = dd * a.d[k]; return r; }
template <int N> inline Dual<N> exp(const Dual<N>& a) { Dual<N> r(std::exp(a.v)); for (int k = 0; k < N; ++k) r.d[k] = r.v * a.d[k]; return r; }
template <int N> inline Dual<N> tanh(const Dual<N>& a) { Dual<N> r(std::tanh(a.v)); double s = 1.0 - r.v * r.v; for (int k = 0; k < N; ++k) r.d[k] = s * a.d[k]; return r; }
template <int N> inline Dual<N> fmin(const Dual<N>& a, double b) { return a.v < b ? a : Dual<N>(b); }
template <int N> inline Dual<N> fmax(const Dual<N>& a, double b) { return a.v > b ? a : Dual<N>(b); }
inline double value(double x) { return x; }
template <int N> inline double value(const Dual<N>& x) { return x.v; }
using std::sqrt; using std::fabs; using std::pow; using std::exp; using std::tanh; using std::fmin; using std::fmax;

// ----------------------------------------------------------------- arrays --
template <class T>
struct Arr2 {
  size_t n0 = 0, n1 = 0;
  std::vector<T> a;
  void alloc(size_t m0, size_t m1, T fill) { n0 = m0; n1 = m1; a.assign(m0 * m1, fill); }
  T& operator()(size_t i, size_t j) { return a[i * n1 + j]; }
  const T& operator()(size_t i, size_t j) const { return a[i * n1 + j]; }
};
template <class T>
struct Arr3 {
  size_t n0 = 0, n1 = 0, n2 = 0;
  std::vector<T> a;
  void alloc(size_t m0, size_t m1, size_t m2, T fill) { n0 = m0; n1 = m1; n2 = m2; a.assign(m0 * m1 * m2, fill); }
  T& operator()(size_t i, size_t j, size_t k) { return a[(i * n1 + j) * n2 + k]; }
  const T& operator()(size_t i, size_t j, size_t k) const { return a[(i * n1 + j) * n2 + k]; }
};

// --------------------------------------------------------------- geometry --
// Mesh::calc_metrics, src/utils/mesh.cpp:171-205 (the chi_x.. transforms at
// :207-278 feed only the unused cell gradient and are not restated).
struct Geometry {
  size_t ni = 0, nj = 0, nic = 0, njc = 0;
  Arr2<double> xv, yv;                 // [ni][nj]
  Arr3<double> normal_chi, normal_eta; // [ni][njc][2], [nic][nj][2]
  Arr2<double> volume, xc, yc, ds_chi, ds_eta;  // [nic][njc]

  void build(size_t ni_, size_t nj_, const double* x, const double* y) {
    ni = ni_; nj = nj_; nic = ni - 1; njc = nj - 1;
    xv.alloc(ni, nj, 0.0); yv.alloc(ni, nj, 0.0);
    for (size_t k = 0; k < ni * nj; ++k) { xv.a[k] = x[k]; yv.a[k] = y[k]; }
    normal_chi.alloc(ni, njc, 2, 0.0);
    normal_eta.alloc(nic, nj, 2, 0.0);
    volume.alloc(nic, njc, 0.0); xc.alloc(nic, njc, 0.0); yc.alloc(nic, njc, 0.0);
    ds_chi.alloc(nic, njc, 0.0); ds_eta.alloc(nic, njc, 0.0);
    Arr3<double> reta, rchi;
    reta.alloc(nic, nj, 2, 0.0); rchi.alloc(ni, njc, 2, 0.0);
    for (size_t i = 0; i < nic; ++i)
      for (size_t j = 0; j < nj; ++j) {
        reta(i, j, 0) = xv(i + 1, j) - xv(i, j);
        reta(i, j, 1) = yv(i + 1, j) - yv(i, j);
        normal_eta(i, j, 0) = -reta(i, j, 1);
        normal_eta(i, j, 1) = reta(i, j, 0);
      }
    for (size_t i = 0; i < ni; ++i)
      for (size_t j = 0; j < njc; ++j) {
        rchi(i, j, 0) = xv(i, j + 1) - xv(i, j);
        rchi(i, j, 1) = yv(i, j + 1) - yv(i, j);
        normal_chi(i, j, 0) = rchi(i, j, 1);
        normal_chi(i, j, 1) = -rchi(i, j, 0);
      }
    for (size_t i = 0; i < nic; ++i)
      for (size_t j = 0; j < njc; ++j) {
        volume(i, j) = 0.5 * (reta(i, j, 0) * rchi(i, j, 1) - rchi(i, j, 0) * reta(i, j, 1) +
                              reta(i, j + 1, 0) * rchi(i + 1, j, 1) - rchi(i + 1, j, 0) * reta(i, j + 1, 1));
        xc(i, j) = 0.25 * (xv(i, j) + xv(i + 1, j) + xv(i, j + 1) + xv(i + 1, j + 1));
        yc(i, j) = 0.25 * (yv(i, j) + yv(i + 1, j) + yv(i, j + 1) + yv(i + 1, j + 1));
        ds_eta(i, j) = std::sqrt(normal_eta(i, j, 0) * normal_eta(i, j, 0) + normal_eta(i, j, 1) * normal_eta(i, j, 1));
        ds_chi(i, j) = std::sqrt(normal_chi(i, j, 0) * normal_chi(i, j, 0) + normal_chi(i, j, 1) * normal_chi(i, j, 1));
      }
  }
};

// -------------------------------------------------------------- parameters --
enum BcType { BC_FREESTREAM = 0, BC_OUTFLOW = 1, BC_SLIPWALL = 2, BC_WALL = 3,
              BC_ISOTHERMAL = 4, BC_PERIODIC = 5, BC_WAKE = 6 };
enum Face { FACE_BOTTOM = 0, FACE_RIGHT = 1, FACE_TOP = 2, FACE_LEFT = 3 };  // src/model/bc.h:8-11

struct Bc {
  int type, face;
  long start, end;  // ghosted indices, inclusive; negative end resolved by the caller
  double u, v, T;
};

// SA-noft2 constants (Spalart & Allmaras 1994, as catalogued by NASA TMR).
struct SaConst {
  double cb1 = 0.1355, sigma = 2.0 / 3.0, cb2 = 0.622, kappa = 0.41;
  double cw2 = 0.3, cw3 = 2.0, cv1 = 7.1, cv2 = 0.7, cv3 = 0.9, prt = 0.9;
  double cw1() const { return cb1 / (kappa * kappa) + (1.0 + cb2) / sigma; }
};

// Closure of the SA production multiplier beta (the reference has none):
// 0 the per-cell field, 1 beta = 1 + amp tanh(a . f + a0), 2 beta = 1 +
// MLP(f).  The MLP is evaluated in FP64 with the rounding points of the
// device inference kept: features to FP32, hidden layer 1 (and its tangent)
// and W2 to BF16; everything else exact.
struct Closure {
  int kind = 0;
  double coef[7] = {0, 0, 0, 0, 0, 0, 0};
  std::vector<double> W1, b1, W2, b2, W3;  // W2 already rounded to BF16
  double b3 = 0.0;
};

// Round to the nearest bfloat16 (8-bit significand), ties to even, via FP32.
inline double bf16_round(double x) {
  float f = (float)x;
  unsigned u;
  std::memcpy(&u, &f, 4);
  u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
  std::memcpy(&f, &u, 4);
  return (double)f;
}
inline double round_bf16(double x) { return bf16_round(x); }
template <int N> inline Dual<N> round_bf16(const Dual<N>& a) {
  Dual<N> r(bf16_round(a.v));
  for (int k = 0; k < N; ++k) r.d[k] = bf16_round(a.d[k]);
  return r;
}
inline double round_f32(double x) { return (double)(float)x; }
template <int N> inline Dual<N> round_f32(const Dual<N>& a) {
  Dual<N> r((double)(float)a.v);
  for (int k = 0; k < N; ++k) r.d[k] = (double)(float)a.d[k];
  return r;
}
template <int N> inline Dual<N> log1p(const Dual<N>& a) { Dual<N> r(std::log1p(a.v)); for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / (1.0 + a.v); return r; }
template <int N> inline Dual<N> asinh(const Dual<N>& a) { Dual<N> r(std::asinh(a.v)); double s = 1.0 / std::sqrt(1.0 + a.v * a.v); for (int k = 0; k < N; ++k) r.d[k] = s * a.d[k]; return r; }
using std::log1p; using std::asinh;

struct Params {
  double rho_inf = 1.0, u_inf = 0.0, v_inf = 0.0, p_inf = 1.0 / 1.4, T_inf = 1.0 / 1.4;
  double mu_inf = 0.0, pr_inf = 0.7, aoa = 0.0;  // aoa in radians
  double nut_inf = 0.0;                           // freestream nu~ (SA only)
  double dpdx = 0.0, dpdy = 0.0;
  int order = 2;      // MUSCL order: 1 or 2
  int sa = 0;         // 0: laminar, nq=4; 1: Spalart-Allmaras, 5 unknowns
  int flux = 0;       // convective flux: 0 Roe, 1 AUSM (solver.flux)
  bool viscous() const { return mu_inf > 1e-15; }   // src/utils/config.cpp:41
  int nvar() const { return sa ? 5 : 4; }
};

}  // namespace oracle

namespace oracle {

// One residual evaluation with the reference's intermediate arrays.  The
// ghosted primitive arrays are (nic+2) x (njc+2) with the interior at
// shift (1,1) (src/model/eulerequation.cpp:41-47, fluid.cpp:50-67); they are
// filled with NaN before every call so that a ghost cell no boundary
// condition writes shows up as NaN instead of stale memory.
template <class T>
struct Residual {
  const Geometry* g = nullptr;
  Params prm;
  SaConst sa;
  std::vector<Bc> bcs;
  std::vector<double> wall_dist;  // [nic][njc], SA only
  std::vector<double> beta;       // [nic][njc], SA only (1 = plain SA)
  Closure clo;                    // beta closure (kind 0: the beta field)
  double* aux = nullptr;          // optional [nic][njc][6]: features and beta

  Arr2<T> rho, u, v, p, Tt, mu, kk, nt, eta, rnt;   // ghosted
  Arr2<T> ql[5], qr[5];   // chi faces [ni][njc]: rho,u,v,p,nt
  Arr2<T> el[5], er[5];   // eta faces [nic][nj]
  Arr3<T> gu_c, gv_c, gT_c, gn_c, gr_c;  // chi-face gradients [ni][njc][2]
  Arr3<T> gu_e, gv_e, gT_e, gn_e, gr_e;  // eta-face gradients [nic][nj][2]
  Arr2<T> ub_c, vb_c, Tb_c, mub_c, kb_c, etab_c;  // chi-face averages
  Arr2<T> ub_e, vb_e, Tb_e, mub_e, kb_e, etab_e;  // eta-face averages
  Arr3<T> fc, fe;        // convective/diffusive face fluxes [*][*][nvar]

  double R() const { return prm.p_inf / prm.rho_inf / prm.T_inf; }  // fluid.cpp:12
  double cp() const { return kGamma * R() / (kGamma - 1.0); }

  void init(const Geometry* geo, const Params& p_, const std::vector<Bc>& b) {
    g = geo; prm = p_; bcs = b;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    size_t nic = g->nic, njc = g->njc, ni = g->ni, nj = g->nj;
    for (Arr2<T>* a : {&rho, &u, &v, &p, &Tt, &mu, &kk, &nt, &eta, &rnt}) a->alloc(nic + 2, njc + 2, T(nan));
    for (int k = 0; k < 5; ++k) { ql[k].alloc(ni, njc, T(0.0)); qr[k].alloc(ni, njc, T(0.0)); el[k].alloc(nic, nj, T(0.0)); er[k].alloc(nic, nj, T(0.0)); }
    for (Arr3<T>* a : {&gu_c, &gv_c, &gT_c, &gn_c, &gr_c}) a->alloc(ni, njc, 2, T(0.0));
    for (Arr3<T>* a : {&gu_e, &gv_e, &gT_e, &gn_e, &gr_e}) a->alloc(nic, nj, 2, T(0.0));
    for (Arr2<T>* a : {&ub_c, &vb_c, &Tb_c, &mub_c, &kb_c, &etab_c}) a->alloc(ni, njc, T(0.0));
    for (Arr2<T>* a : {&ub_e, &vb_e, &Tb_e, &mub_e, &kb_e, &etab_e}) a->alloc(nic, nj, T(0.0));
    fc.alloc(ni, njc, 5, T(0.0)); fe.alloc(nic, nj, 5, T(0.0));
    wall_dist.assign(nic * njc, 1.0);
    beta.assign(nic * njc, 1.0);
  }

  // FluidModel::primvars, src/model/fluid.cpp:50-67 (interior at shift 1,1).
  void primvars(const T* q) {
    const size_t nic = g->nic, njc = g->njc, nv = prm.nvar();
    const double Rg = R();
#pragma omp parallel for
    for (size_t i = 0; i < nic; ++i)
      for (size_t j = 0; j < njc; ++j) {
        const T* c = q + (i * njc + j) * nv;
        T r = c[0], uu = c[1] / r, vv = c[2] / r;
        rho(i + 1, j + 1) = r; u(i + 1, j + 1) = uu; v(i + 1, j + 1) = vv;
        T pp = (c[3] - 0.5 * r * (uu * uu + vv * vv)) * (kGamma - 1.0);
        p(i + 1, j + 1) = pp;
        Tt(i + 1, j + 1) = pp / r / Rg;
        if (prm.sa) nt(i + 1, j + 1) = c[4] / r;
      }
  }

  // Boundary conditions, src/model/bc.cpp; applied in list order (bc.cpp:430-433).
  void apply_bcs() {
    const long nic = (long)g->nic, njc = (long)g->njc;
    const double Rg = R();
    auto set = [&](long i, long j, T r, T uu, T vv, T pp, T TT, T n) {
      rho(i, j) = r; u(i, j) = uu; v(i, j) = vv; p(i, j) = pp; Tt(i, j) = TT; if (prm.sa) nt(i, j) = n;
    };
    for (const Bc& b : bcs) {
      const bool horiz = (b.face == FACE_BOTTOM || b.face == FACE_TOP);
      switch (b.type) {
        case BC_FREESTREAM: {  // bc.cpp:25-60
          for (long s = b.start; s <= b.end; ++s) {
            long i = horiz ? s : (b.face == FACE_LEFT ? 0 : nic + 1);
            long j = horiz ? (b.face == FACE_BOTTOM ? 0 : njc + 1) : s;
            T r(prm.rho_inf), pp(prm.p_inf);
            set(i, j, r, T(prm.u_inf), T(prm.v_inf), pp, pp / r / Rg, T(prm.nut_inf));
          }
        } break;
        case BC_SLIPWALL: {  // bc.cpp:77-128 (left/right not implemented there)
          if (!horiz) break;
          for (long i = b.start; i <= b.end; ++i) {
            const bool bot = b.face == FACE_BOTTOM;
            long j0 = bot ? 0 : njc + 1, j1 = bot ? 1 : njc, j2 = bot ? 2 : njc - 1;
            double nx = g->normal_eta(i - 1, bot ? 0 : njc, 0), ny = g->normal_eta(i - 1, bot ? 0 : njc, 1);
            double ds = nx * nx + ny * ny;
            T pp = 1.5 * p(i, j1) - 0.5 * p(i, j2);
            T r = 1.5 * rho(i, j1) - 0.5 * rho(i, j2);
            T un = u(i, j1) * nx + v(i, j1) * ny;
            T uu = u(i, j1) - 2.0 * un * nx / ds;
            T vv = v(i, j1) - 2.0 * un * ny / ds;
            T n = prm.sa ? T(1.5 * nt(i, j1) - 0.5 * nt(i, j2)) : T(0.0);
            set(i, j0, r, uu, vv, pp, pp / r / Rg, n);
          }
        } break;
        case BC_WALL:         // bc.cpp:149-204
        case BC_ISOTHERMAL: { // bc.cpp:387-427 (left/right not implemented there)
          const bool iso = b.type == BC_ISOTHERMAL;
          if (iso && !horiz) break;
          for (long s = b.start; s <= b.end; ++s) {
            long i0, j0, i1, j1, i2, j2;
            if (horiz) { i0 = i1 = i2 = s; if (b.face == FACE_BOTTOM) { j0 = 0; j1 = 1; j2 = 2; } else { j0 = njc + 1; j1 = njc; j2 = njc - 1; } }
            else { j0 = j1 = j2 = s; if (b.face == FACE_LEFT) { i0 = 0; i1 = 1; i2 = 2; } else { i0 = nic + 1; i1 = nic; i2 = nic - 1; } }
            T uu = 2.0 * b.u - (1.5 * u(i1, j1) - 0.5 * u(i2, j2));
            T vv = 2.0 * b.v - (1.5 * v(i1, j1) - 0.5 * v(i2, j2));
            T n = prm.sa ? T(-(1.5 * nt(i1, j1) - 0.5 * nt(i2, j2))) : T(0.0);
            if (!iso) {
              T TT = 1.5 * Tt(i1, j1) - 0.5 * Tt(i2, j2);
              T r = 1.5 * rho(i1, j1) - 0.5 * rho(i2, j2);
              set(i0, j0, r, uu, vv, r * Rg * TT, TT, n);
            } else {
              T pp = 1.5 * p(i1, j1) - 0.5 * p(i2, j2);
              T TT(b.T);
              set(i0, j0, pp / TT / Rg, uu, vv, pp, TT, n);
            }
          }
        } break;
        case BC_OUTFLOW: {  // bc.cpp:283-311 (only left/right exist there)
          if (horiz) break;
          long i0 = b.face == FACE_LEFT ? 0 : nic + 1;
          for (long j = b.start; j <= b.end; ++j) {
            long is = i0 - 1;
            T r = rho(is, j), pp(prm.p_inf);
            set(i0, j, r, u(is, j), v(is, j), pp, pp / r / Rg, prm.sa ? nt(is, j) : T(0.0));
          }
        } break;
        case BC_PERIODIC: {  // bc.cpp:328-365
          if (horiz) {
            for (long i = b.start; i <= b.end; ++i) {
              set(i, 0, rho(i, njc), u(i, njc), v(i, njc), p(i, njc), Tt(i, njc), prm.sa ? nt(i, njc) : T(0.0));
              set(i, njc + 1, rho(i, 1), u(i, 1), v(i, 1), p(i, 1), Tt(i, 1), prm.sa ? nt(i, 1) : T(0.0));
            }
          } else {
            for (long j = b.start; j <= b.end; ++j) {
              set(0, j, rho(nic, j), u(nic, j), v(nic, j), p(nic, j), Tt(nic, j), prm.sa ? nt(nic, j) : T(0.0));
              set(nic + 1, j, rho(1, j), u(1, j), v(1, j), p(1, j), Tt(1, j), prm.sa ? nt(1, j) : T(0.0));
            }
          }
        } break;
        case BC_WAKE: {  // bc.cpp:223-263: two sequential loops, both read row 1
          if (!horiz) break;
          long j0 = b.face == FACE_BOTTOM ? 0 : njc + 1;
          for (long i = b.start; i <= b.end; ++i) {
            long m = nic + 1 - i;
            set(i, j0, rho(m, 1), u(m, 1), v(m, 1), p(m, 1), Tt(m, 1), prm.sa ? nt(m, 1) : T(0.0));
          }
          for (long i = b.start; i <= b.end; ++i) {
            long m = nic + 1 - i;
            set(m, j0, rho(i, 1), u(i, 1), v(i, 1), p(i, 1), Tt(i, 1), prm.sa ? nt(i, 1) : T(0.0));
          }
        } break;
        default: break;
      }
    }
  }

  // ReconstructionFirstOrder/SecondOrder, src/model/reconstruction.cpp:15-154.
  void reconstruct_chi(const Arr2<T>& q, Arr2<T>& L, Arr2<T>& Rr, bool first_order) {
    const size_t ni = g->ni, nic = g->nic, njc = g->njc;
#pragma omp parallel for
    for (size_t i = 0; i < ni; ++i)
      for (size_t j = 0; j < njc; ++j) { L(i, j) = q(i, j + 1); Rr(i, j) = q(i + 1, j + 1); }
    if (first_order) return;
    const double eps = std::pow(10.0 / (double)nic, 3);
    const double thm = 2.0 / 3.0, thp = 4.0 / 3.0;
    Arr2<T> d; d.alloc(ni, njc, T(0.0));
#pragma omp parallel for
    for (size_t i = 0; i < ni; ++i)
      for (size_t j = 0; j < njc; ++j) d(i, j) = q(i + 1, j + 1) - q(i, j + 1);
#pragma omp parallel for
    for (size_t i = 0; i < nic; ++i)
      for (size_t j = 0; j < njc; ++j) {
        T dm = d(i, j), dp = d(i + 1, j);
        T a1 = 3.0 * dp * dm;
        T a2 = 2.0 * (dp - dm) * (dp - dm) + a1;
        T s = 0.25 * (a1 + eps) / (a2 + eps);
        L(i + 1, j) = L(i + 1, j) + s * (thm * dm + thp * dp);
        Rr(i, j) = Rr(i, j) - s * (thp * dm + thm * dp);
      }
  }
  void reconstruct_eta(const Arr2<T>& q, Arr2<T>& L, Arr2<T>& Rr, bool first_order) {
    const size_t nj = g->nj, nic = g->nic, njc = g->njc;
#pragma omp parallel for
    for (size_t i = 0; i < nic; ++i)
      for (size_t j = 0; j < nj; ++j) { L(i, j) = q(i + 1, j); Rr(i, j) = q(i + 1, j + 1); }
    if (first_order) return;
    const double eps = std::pow(10.0 / (double)njc, 3);
    const double thm = 2.0 / 3.0, thp = 4.0 / 3.0;
    Arr2<T> d; d.alloc(nic, nj, T(0.0));
#pragma omp parallel for
    for (size_t i = 0; i < nic; ++i)
      for (size_t j = 0; j < nj; ++j) d(i, j) = q(i + 1, j + 1) - q(i + 1, j);
#pragma omp parallel for
    for (size_t i = 0; i < nic; ++i)
      for (size_t j = 0; j < njc; ++j) {
        T dm = d(i, j), dp = d(i, j + 1);
        T a1 = 3.0 * dp * dm;
        T a2 = 2.0 * (dp - dm) * (dp - dm) + a1;
        T s = 0.25 * (a1 + eps) / (a2 + eps);
        L(i, j + 1) = L(i, j + 1) + s * (thm * dm + thp * dp);
        Rr(i, j) = Rr(i, j) - s * (thp * dm + thm * dp);
      }
  }

  // Mesh::calc_face, src/utils/mesh.cpp:8-32.
  void face_avg(const Arr2<T>& q, Arr2<T>& qc, Arr2<T>& qe) {
    const size_t ni = g->ni, nj = g->nj, nic = g->nic, njc = g->njc;
    for (size_t i = 0; i < ni; ++i)
      for (size_t j = 0; j < njc; ++j) {
        T a = q(i, j + 1), b = q(i + 1, j + 1);
        T top = 0.25 * (a + b + q(i, j + 2) + q(i + 1, j + 2));
        T bot = 0.25 * (a + b + q(i, j) + q(i + 1, j));
        qc(i, j) = 0.25 * (a + b + top + bot);
      }
    for (size_t i = 0; i < nic; ++i)
      for (size_t j = 0; j < nj; ++j) {
        T t = q(i + 1, j + 1), bt = q(i + 1, j);
        T l = 0.25 * (t + bt + q(i, j + 1) + q(i, j));
        T r = 0.25 * (t + bt + q(i + 2, j + 1) + q(i + 2, j));
        qe(i, j) = 0.25 * (l + r + t + bt);
      }
  }

  // Mesh::calc_gradient (face form), src/utils/mesh.cpp:34-131, including
  // the one-sided metric branches at i=0, i=ni-1 and j=0, j=nj-1.
  void face_grad(const Arr2<T>& q, Arr3<T>& gc, Arr3<T>& ge) {
    const size_t ni = g->ni, nj = g->nj, nic = g->nic, njc = g->njc;
    const Arr3<double>& nc = g->normal_chi; const Arr3<double>& ne = g->normal_eta; const Arr2<double>& V = g->volume;
    double nt_[2], nb_[2], nl_[2], nr_[2], vol;
    for (size_t i = 0; i < ni; ++i)
      for (size_t j = 0; j < njc; ++j) {
        T ql_ = q(i, j + 1), qr_ = q(i + 1, j + 1);
        T qt = 0.25 * (ql_ + qr_ + q(i, j + 2) + q(i + 1, j + 2));
        T qb = 0.25 * (ql_ + qr_ + q(i, j) + q(i + 1, j));
        for (int k = 0; k < 2; ++k) {
          if (i == 0) {
            nt_[k] = ne(i, j + 1, k); nb_[k] = ne(i, j, k);
            nr_[k] = 0.5 * (nc(i, j, k) + nc(i + 1, j, k)); nl_[k] = nc(i, j, k);
          } else if (i == ni - 1) {
            nt_[k] = ne(i - 1, j + 1, k); nb_[k] = ne(i - 1, j, k);
            nr_[k] = nc(i, j, k); nl_[k] = 0.5 * (nc(i, j, k) + nc(i - 1, j, k));
          } else {
            nt_[k] = 0.5 * (ne(i, j + 1, k) + ne(i - 1, j + 1, k)); nb_[k] = 0.5 * (ne(i, j, k) + ne(i - 1, j, k));
            nr_[k] = 0.5 * (nc(i, j, k) + nc(i + 1, j, k)); nl_[k] = 0.5 * (nc(i, j, k) + nc(i - 1, j, k));
          }
        }
        vol = (i == 0) ? V(i, j) : (i == ni - 1) ? V(i - 1, j) : 0.5 * (V(i, j) + V(i - 1, j));
        for (int k = 0; k < 2; ++k)
          gc(i, j, k) = (nt_[k] * qt - nb_[k] * qb + nr_[k] * qr_ - nl_[k] * ql_) / vol;
      }
    for (size_t i = 0; i < nic; ++i)
      for (size_t j = 0; j < nj; ++j) {
        T qt = q(i + 1, j + 1), qb = q(i + 1, j);
        T ql_ = 0.25 * (qt + qb + q(i, j + 1) + q(i, j));
        T qr_ = 0.25 * (qt + qb + q(i + 2, j + 1) + q(i + 2, j));
        for (int k = 0; k < 2; ++k) {
          if (j == 0) {
            nt_[k] = 0.5 * (ne(i, j, k) + ne(i, j + 1, k)); nb_[k] = ne(i, j, k);
            nl_[k] = nc(i, j, k); nr_[k] = nc(i + 1, j, k);
          } else if (j == nj - 1) {
            nt_[k] = ne(i, j, k); nb_[k] = 0.5 * (ne(i, j, k) + ne(i, j - 1, k));
            nl_[k] = nc(i, j - 1, k); nr_[k] = nc(i + 1, j - 1, k);
          } else {
            nt_[k] = 0.5 * (ne(i, j, k) + ne(i, j + 1, k)); nb_[k] = 0.5 * (ne(i, j, k) + ne(i, j - 1, k));
            nl_[k] = 0.5 * (nc(i, j, k) + nc(i, j - 1, k)); nr_[k] = 0.5 * (nc(i + 1, j, k) + nc(i + 1, j - 1, k));
          }
        }
        vol = (j == 0) ? V(i, j) : (j == nj - 1) ? V(i, j - 1) : 0.5 * (V(i, j) + V(i, j - 1));
        for (int k = 0; k < 2; ++k)
          ge(i, j, k) = (nt_[k] * qt - nb_[k] * qb + nr_[k] * qr_ - nl_[k] * ql_) / vol;
      }
  }

  // ConvectiveFluxRoe::evaluate, src/model/flux.cpp:51-146 (no entropy fix).
  // Writes the four mean-flow components, and with SA the first-order upwind
  // transport of rho*nu~ by the Roe mass flux.
  static void roe(double nx, double ny, T rl, T ul, T vl, T pl, T rr, T ur, T vr, T pr, T ntl, T ntr, bool with_sa, T* f) {
    const double gm1 = kGamma - 1.0, ogm1 = 1.0 / gm1;
    T rli = 1.0 / rl, rul = rl * ul, rvl = rl * vl, uvl = 0.5 * (ul * ul + vl * vl);
    T el_ = pl * ogm1 + rl * uvl, hl = (el_ + pl) * rli;
    T rri = 1.0 / rr, rur = rr * ur, rvr = rr * vr, uvr = 0.5 * (ur * ur + vr * vr);
    T er_ = pr * ogm1 + rr * uvr, hr = (er_ + pr) * rri;
    T rat = sqrt(rr * rli), rati = 1.0 / (rat + 1.0);
    T rav = rat * rl, uav = (rat * ur + ul) * rati, vav = (rat * vr + vl) * rati, hav = (rat * hr + hl) * rati;
    T uv = 0.5 * (uav * uav + vav * vav);
    T cav = sqrt(gm1 * (hav - uv));
    T d1 = rr - rl, d2 = ur - ul, d3 = vr - vl, d4 = pr - pl;
    double dr = std::sqrt(nx * nx + ny * ny), r1 = nx / dr, r2 = ny / dr;
    T uu = r1 * uav + r2 * vav, c2 = cav * cav, c2i = 1.0 / c2;
    T auu = fabs(uu), aupc = fabs(uu + cav), aumc = fabs(uu - cav);
    T uul = r1 * ul + r2 * vl, uur = r1 * ur + r2 * vr;
    T rcav = rav * cav, aquu = uur - uul, c2ih = 0.5 * c2i, ruuav = auu * rav;
    T b1 = auu * (d1 - c2i * d4);
    T b2 = c2ih * aupc * (d4 + rcav * aquu);
    T b3 = c2ih * aumc * (d4 - rcav * aquu);
    T b4 = b1 + b2 + b3, b5 = cav * (b2 - b3);
    T b6 = ruuav * (d2 - r1 * aquu), b7 = ruuav * (d3 - r2 * aquu);
    T a1 = b4, a2 = uav * b4 + r1 * b5 + b6, a3 = vav * b4 + r2 * b5 + b7;
    T a4 = hav * b4 + uu * b5 + uav * b6 + vav * b7 - c2 * b1 * ogm1;
    double aj = 0.5 * dr;
    T plar = pl + pr, epl = el_ + pl, epr = er_ + pr;
    f[0] = aj * (rl * uul + rr * uur - a1);
    f[1] = aj * (rul * uul + rur * uur + r1 * plar - a2);
    f[2] = aj * (rvl * uul + rvr * uur + r2 * plar - a3);
    f[3] = aj * (epl * uul + epr * uur - a4);
    if (with_sa) f[4] = 0.5 * (f[0] * (ntl + ntr) - fabs(f[0]) * (ntr - ntl));
  }

  // ConvectiveFluxAUSM::evaluate, src/model/flux.cpp:149-224 (Liou-Steffen
  // AUSM: split Mach numbers and pressures, upwinded on the interface Mach
  // number).  With SA, rho*nu~ is carried first-order upwind by the AUSM mass
  // flux, as for Roe.
  static T split_mach(T M, double sgn) {  // M+ (sgn=+1) / M- (sgn=-1)
    if (fabs(M) <= 1.0) return sgn * 0.25 * (M + sgn) * (M + sgn);
    return 0.5 * (M + sgn * fabs(M));
  }
  static T split_pres(T M, T p, double sgn) {  // P+ / P-
    if (fabs(M) <= 1.0) return 0.25 * p * (M + sgn) * (M + sgn) * (2.0 - sgn * M);
    return 0.5 * p * (M + sgn * fabs(M)) / M;
  }
  static void ausm(double nx, double ny, T rl, T ul, T vl, T pl, T rr, T ur, T vr, T pr, T ntl, T ntr, bool with_sa, T* f) {
    const double ogm1 = 1.0 / (kGamma - 1.0);
    const double ds = std::sqrt(nx * nx + ny * ny);
    T al = sqrt(kGamma * pl / rl), ar = sqrt(kGamma * pr / rr);
    T Ml = ((ul * nx + vl * ny) / ds) / al, Mr = ((ur * nx + vr * ny) / ds) / ar;
    T hl = (pl * ogm1 + rl * 0.5 * (ul * ul + vl * vl) + pl) / rl;
    T hr = (pr * ogm1 + rr * 0.5 * (ur * ur + vr * vr) + pr) / rr;
    T mh = split_mach(Ml, 1.0) + split_mach(Mr, -1.0);
    T ph = split_pres(Ml, pl, 1.0) + split_pres(Mr, pr, -1.0);
    const bool up = mh >= 0.0;
    T m = (up ? rl * al : rr * ar) * mh * ds;
    f[0] = m;
    f[1] = m * (up ? ul : ur) + ph * nx;
    f[2] = m * (up ? vl : vr) + ph * ny;
    f[3] = m * (up ? hl : hr);
    if (with_sa) f[4] = 0.5 * (f[0] * (ntl + ntr) - fabs(f[0]) * (ntr - ntl));
  }

  void conv_flux(double nx, double ny, T rl, T ul, T vl, T pl, T rr, T ur, T vr, T pr, T ntl, T ntr, T* f) const {
    if (prm.flux == 1) ausm(nx, ny, rl, ul, vl, pl, rr, ur, vr, pr, ntl, ntr, prm.sa, f);
    else roe(nx, ny, rl, ul, vl, pl, rr, ur, vr, pr, ntl, ntr, prm.sa, f);
  }

  // DiffusiveFluxGreenGauss::evaluate, src/model/flux.cpp:12-48; with SA the
  // nu~ diffusion (1/sigma) * eta_bar * grad(nu~).n is added as component 4.
  void viscous_flux(double nx, double ny, const T* gu, const T* gv, const T* gT, T ub, T vb, T m, T k, const T* gn, T et, T* f) const {
    T txy = m * (gu[1] + gv[0]);
    T txx = m * (2.0 * gu[0] - 2.0 / 3.0 * (gu[0] + gv[1]));
    T tyy = m * (2.0 * gv[1] - 2.0 / 3.0 * (gu[0] + gv[1]));
    T qx = -k * gT[0], qy = -k * gT[1];
    f[0] = T(0.0);
    f[1] = txx * nx + txy * ny;
    f[2] = txy * nx + tyy * ny;
    f[3] = (ub * txx + vb * txy - qx) * nx + (ub * txy + vb * tyy - qy) * ny;
    if (prm.sa) f[4] = (1.0 / sa.sigma) * et * (gn[0] * nx + gn[1] * ny);
  }

  // Laminar viscosity (power law) and conductivity, src/model/fluid.cpp:38-46;
  // with SA the eddy viscosity rho*nu~*fv1 is added (zero for nu~ < 0) and
  // eta = mu + rho*nu~*fn is the nu~ diffusion coefficient (SA-neg fn).
  void transport_props() {
    const size_t n = rho.a.size();
    const double c_p = cp(), cv1_3 = sa.cv1 * sa.cv1 * sa.cv1, cn1 = 16.0;
    for (size_t k = 0; k < n; ++k) {
      T ml = prm.mu_inf * pow(Tt.a[k] / prm.T_inf, 2.0 / 3.0);
      if (!prm.sa) { mu.a[k] = ml; kk.a[k] = ml * c_p / prm.pr_inf; continue; }
      T rn = rho.a[k] * nt.a[k];
      T chi = rn / ml, chi3 = chi * chi * chi;
      T mut(0.0), fn(1.0);
      if (value(nt.a[k]) >= 0.0) mut = rn * (chi3 / (chi3 + cv1_3));
      else fn = (cn1 + chi3) / (cn1 - chi3);
      mu.a[k] = ml + mut;
      kk.a[k] = c_p * (ml / prm.pr_inf + mut / sa.prt);
      eta.a[k] = ml + rn * fn;
    }
  }

  // SA-noft2 source at one interior cell, multiplied by the cell volume by the
  // caller.  Cell gradients are Green-Gauss over the cell's four faces with
  // the face value the arithmetic mean of the two adjacent cells (ghost cells
  // across boundary faces) and outward area-weighted normals.
  void cell_grad(const Arr2<T>& q, size_t i, size_t j, T* out) const {
    const size_t gi = i + 1, gj = j + 1;
    const Arr3<double>& nc = g->normal_chi; const Arr3<double>& ne = g->normal_eta;
    T c = q(gi, gj);
    T fe = 0.5 * (c + q(gi + 1, gj)), fw = 0.5 * (c + q(gi - 1, gj));
    T fn = 0.5 * (c + q(gi, gj + 1)), fs = 0.5 * (c + q(gi, gj - 1));
    const double V = g->volume(i, j);
    for (int k = 0; k < 2; ++k)
      out[k] = (fe * nc(i + 1, j, k) - fw * nc(i, j, k) + fn * ne(i, j + 1, k) - fs * ne(i, j, k)) / V;
  }
  // Local features of the closure (all O(1)): vorticity/strain ratio,
  // log wall-distance Reynolds number, asinh eddy-viscosity ratio,
  // streamwise pressure-gradient indicator, production/destruction balance.
  void features(size_t i, size_t j, T r, T uu, T vv, T n, T nu, const T* g_u, const T* g_v, const T* g_p, double d,
                T P1, T D, T* f) const {
    const double U2 = prm.u_inf * prm.u_inf + prm.v_inf * prm.v_inf, U = std::sqrt(U2);
    T om = fabs(g_v[0] - g_u[1]);
    T s12 = g_u[1] + g_v[0];
    T smag = sqrt(2.0 * (g_u[0] * g_u[0] + g_v[1] * g_v[1]) + s12 * s12 + 1e-12 * U2);
    f[0] = om / (om + smag);
    f[1] = log1p(om * (d * d) / nu) * 0.1;
    f[2] = asinh(n / nu) * 0.125;
    T q2 = uu * uu + vv * vv;
    T umag = sqrt(q2 + 1e-6 * U2);
    T qd = 0.5 * r * q2 + 0.025 * prm.rho_inf * U2;
    f[3] = tanh(d * (uu * g_p[0] + vv * g_p[1]) / (umag * qd));
    f[4] = P1 / (fabs(P1) + fabs(D) + 1e-8 * prm.rho_inf * U2 * U);
    (void)i; (void)j;
  }
  T mlp(const T* f) const {
    T x[5];
    for (int c = 0; c < 5; ++c) x[c] = round_f32(f[c]);
    T h1[64];
    for (int k = 0; k < 64; ++k) {
      T a(clo.b1[k]);
      for (int c = 0; c < 5; ++c) a = a + clo.W1[k * 5 + c] * x[c];
      h1[k] = round_bf16(tanh(a));
    }
    T y(clo.b3);
    for (int jj = 0; jj < 64; ++jj) {
      T a(0.0);
      for (int k = 0; k < 64; ++k) a = a + clo.W2[jj * 64 + k] * h1[k];
      y = y + clo.W3[jj] * tanh(a + clo.b2[jj]);
    }
    return y;
  }

  // SA-noft2 source: beta P1 - D + cb2 and density-gradient diffusion terms.
  T sa_source(size_t i, size_t j) const {
    const size_t gi = i + 1, gj = j + 1;
    T r = rho(gi, gj), n = nt(gi, gj);
    T ml = prm.mu_inf * pow(Tt(gi, gj) / prm.T_inf, 2.0 / 3.0);
    T nu = ml / r;
    T g_u[2], g_v[2], g_n[2], g_r[2];
    cell_grad(u, i, j, g_u); cell_grad(v, i, j, g_v); cell_grad(nt, i, j, g_n); cell_grad(rho, i, j, g_r);
    T omega = fabs(g_v[0] - g_u[1]);
    const double d = wall_dist[i * g->njc + j];
    const double k2d2 = sa.kappa * sa.kappa * d * d;
    T P1, D;
    if (value(n) >= 0.0) {
      T chi = n / nu, chi3 = chi * chi * chi;
      T fv1 = chi3 / (chi3 + sa.cv1 * sa.cv1 * sa.cv1);
      T fv2 = 1.0 - chi / (1.0 + chi * fv1);
      T sbar = n * fv2 / k2d2;
      T stil = (value(sbar) >= -sa.cv2 * value(omega))
                   ? omega + sbar
                   : omega + omega * (sa.cv2 * sa.cv2 * omega + sa.cv3 * sbar) / ((sa.cv3 - 2.0 * sa.cv2) * omega - sbar);
      T rr = (value(stil) > 0.0) ? fmin(n / (stil * k2d2), 10.0) : T(10.0);
      T r2 = rr * rr, r6 = r2 * r2 * r2;
      T gg = rr + sa.cw2 * (r6 - rr);
      T g2 = gg * gg, g6 = g2 * g2 * g2;
      const double cw3_6 = std::pow(sa.cw3, 6);
      T fw = gg * pow((1.0 + cw3_6) / (g6 + cw3_6), 1.0 / 6.0);
      P1 = sa.cb1 * stil * r * n;
      D = sa.cw1() * fw * r * (n / d) * (n / d);
    } else {
      P1 = sa.cb1 * (1.0 - 1.2) * omega * r * n;
      D = -sa.cw1() * r * (n / d) * (n / d);
    }
    T cb2t = (sa.cb2 / sa.sigma) * r * (g_n[0] * g_n[0] + g_n[1] * g_n[1]);
    T cross = -(1.0 / sa.sigma) * (nu + n) * (g_r[0] * g_n[0] + g_r[1] * g_n[1]);
    T b(beta[i * g->njc + j]);
    T f[5];
    if (clo.kind != 0 || aux) {
      T g_p[2];
      cell_grad(p, i, j, g_p);
      features(i, j, r, u(gi, gj), v(gi, gj), n, nu, g_u, g_v, g_p, d, P1, D, f);
    }
    if (clo.kind == 1) {
      T a(clo.coef[5]);
      for (int k = 0; k < 5; ++k) a = a + clo.coef[k] * f[k];
      b = 1.0 + clo.coef[6] * tanh(a);
    } else if (clo.kind == 2) {
      b = 1.0 + mlp(f);
    }
    if (aux) {
      double* o = aux + (i * g->njc + j) * 6;
      for (int k = 0; k < 5; ++k) o[k] = value(f[k]);
      o[5] = value(b);
    }
    return b * P1 - D + (cb2t + cross);
  }

  // EulerEquation::calc_residual, src/model/eulerequation.cpp:202-232.
  void evaluate(const T* q, T* rhs) {
    const size_t ni = g->ni, nj = g->nj, nic = g->nic, njc = g->njc;
    const size_t nv = prm.nvar();
    const double nan = std::numeric_limits<double>::quiet_NaN();
    for (Arr2<T>* a : {&rho, &u, &v, &p, &Tt, &mu, &kk, &nt, &eta, &rnt}) std::fill(a->a.begin(), a->a.end(), T(nan));
    primvars(q);
    apply_bcs();
    const bool fo = prm.order == 1;
    Arr2<T>* prim[4] = {&rho, &u, &v, &p};
    for (int k = 0; k < 4; ++k) { reconstruct_chi(*prim[k], ql[k], qr[k], fo); reconstruct_eta(*prim[k], el[k], er[k], fo); }
    if (prm.sa) { reconstruct_chi(nt, ql[4], qr[4], true); reconstruct_eta(nt, el[4], er[4], true); }
    const bool visc = prm.viscous();
    if (visc) {
      transport_props();
      face_grad(u, gu_c, gu_e); face_grad(v, gv_c, gv_e); face_grad(Tt, gT_c, gT_e);
      face_avg(u, ub_c, ub_e); face_avg(v, vb_c, vb_e); face_avg(Tt, Tb_c, Tb_e);
      face_avg(mu, mub_c, mub_e); face_avg(kk, kb_c, kb_e);
      if (prm.sa) { face_grad(nt, gn_c, gn_e); face_avg(eta, etab_c, etab_e); }
    }
    const Arr3<double>& nc = g->normal_chi; const Arr3<double>& ne = g->normal_eta;
    // convective fluxes
#pragma omp parallel for
    for (size_t i = 0; i < ni; ++i)
      for (size_t j = 0; j < njc; ++j) {
        T f[5];
        conv_flux(nc(i, j, 0), nc(i, j, 1), ql[0](i, j), ql[1](i, j), ql[2](i, j), ql[3](i, j),
                  qr[0](i, j), qr[1](i, j), qr[2](i, j), qr[3](i, j), ql[4](i, j), qr[4](i, j), f);
        for (size_t k = 0; k < nv; ++k) fc(i, j, k) = f[k];
      }
#pragma omp parallel for
    for (size_t i = 0; i < nic; ++i)
      for (size_t j = 0; j < nj; ++j) {
        T f[5];
        conv_flux(ne(i, j, 0), ne(i, j, 1), el[0](i, j), el[1](i, j), el[2](i, j), el[3](i, j),
                  er[0](i, j), er[1](i, j), er[2](i, j), er[3](i, j), el[4](i, j), er[4](i, j), f);
        for (size_t k = 0; k < nv; ++k) fe(i, j, k) = f[k];
      }
    for (size_t m = 0; m < nic * njc * nv; ++m) rhs[m] = T(0.0);
#pragma omp parallel for
    for (size_t i = 0; i < nic; ++i)
      for (size_t j = 0; j < njc; ++j)
        for (size_t k = 0; k < nv; ++k) {
          T& r = rhs[(i * njc + j) * nv + k];
          r -= (fe(i, j + 1, k) - fe(i, j, k));
          r -= (fc(i + 1, j, k) - fc(i, j, k));
        }
    if (visc) {
      for (size_t i = 0; i < ni; ++i)
        for (size_t j = 0; j < njc; ++j) {
          T f[5];
          T gu[2] = {gu_c(i, j, 0), gu_c(i, j, 1)}, gv[2] = {gv_c(i, j, 0), gv_c(i, j, 1)}, gT[2] = {gT_c(i, j, 0), gT_c(i, j, 1)};
          T gn[2] = {gn_c(i, j, 0), gn_c(i, j, 1)};
          viscous_flux(nc(i, j, 0), nc(i, j, 1), gu, gv, gT, ub_c(i, j), vb_c(i, j), mub_c(i, j), kb_c(i, j), gn, etab_c(i, j), f);
          for (size_t k = 0; k < nv; ++k) fc(i, j, k) = f[k];
        }
      for (size_t i = 0; i < nic; ++i)
        for (size_t j = 0; j < nj; ++j) {
          T f[5];
          T gu[2] = {gu_e(i, j, 0), gu_e(i, j, 1)}, gv[2] = {gv_e(i, j, 0), gv_e(i, j, 1)}, gT[2] = {gT_e(i, j, 0), gT_e(i, j, 1)};
          T gn[2] = {gn_e(i, j, 0), gn_e(i, j, 1)};
          viscous_flux(ne(i, j, 0), ne(i, j, 1), gu, gv, gT, ub_e(i, j), vb_e(i, j), mub_e(i, j), kb_e(i, j), gn, etab_e(i, j), f);
          for (size_t k = 0; k < nv; ++k) fe(i, j, k) = f[k];
        }
      for (size_t i = 0; i < nic; ++i)
        for (size_t j = 0; j < njc; ++j)
          for (size_t k = 1; k < nv; ++k) {
            T& r = rhs[(i * njc + j) * nv + k];
            r += (fe(i, j + 1, k) - fe(i, j, k));
            r += (fc(i + 1, j, k) - fc(i, j, k));
          }
    }
    const Arr2<double>& V = g->volume;
    for (size_t i = 0; i < nic; ++i)
      for (size_t j = 0; j < njc; ++j) {
        rhs[(i * njc + j) * nv + 1] += -prm.dpdx * V(i, j);
        rhs[(i * njc + j) * nv + 2] += -prm.dpdy * V(i, j);
        if (prm.sa && visc) rhs[(i * njc + j) * nv + 4] += V(i, j) * sa_source(i, j);
      }
#pragma omp parallel for
    for (size_t i = 0; i < nic; ++i)
      for (size_t j = 0; j < njc; ++j)
        for (size_t k = 0; k < nv; ++k) rhs[(i * njc + j) * nv + k] /= V(i, j);
  }
};

// EulerEquation::calc_dt, src/model/eulerequation.cpp:237-259 (mu_inf, not
// local mu; the same dt for every equation of a cell).
inline void local_dt(const Geometry& g, const Params& prm, const double* q, double cfl, double* dt) {
  const size_t nic = g.nic, njc = g.njc, nv = prm.nvar();
#pragma omp parallel for
  for (size_t i = 0; i < nic; ++i)
    for (size_t j = 0; j < njc; ++j) {
      const double* c = q + (i * njc + j) * nv;
      double r = c[0], uu = c[1] / r, vv = c[2] / r;
      double pp = (c[3] - 0.5 * r * (uu * uu + vv * vv)) * (kGamma - 1.0);
      double lam = std::sqrt(kGamma * pp / r) + std::fabs(uu) + std::fabs(vv);
      double len = std::min(g.ds_eta(i, j), g.ds_chi(i, j));
      double val = cfl / (lam / len + 2.0 * prm.mu_inf / len / len);
      for (size_t k = 0; k < nv; ++k) dt[(i * njc + j) * nv + k] = val;
    }
}

// Residual L2 norm per equation, src/solver/solver.cpp:125-134 (unnormalised).
inline void l2norm(size_t ncell, size_t nv, const double* r, double* out) {
  for (size_t k = 0; k < nv; ++k) {
    double s = 0.0;
    for (size_t c = 0; c < ncell; ++c) s += r[c * nv + k] * r[c * nv + k];
    out[k] = std::sqrt(s);
  }
}

// IOManager::write_surface, src/utils/io.cpp:181-255: wall cells
// i in [tail-1, tail-1+nb), nb = ni - 2*tail + 1 (src/utils/mesh.cpp:349-350).
// forces = {cl_p, cd_p, cl_v, cd_v, cl, cd}.  gu_e/gv_e are the eta-face
// gradients of the last residual evaluation at the same state.
template <class T>
void surface(const Geometry& g, const Params& prm, const Residual<T>& res, const double* q, long tail,
             double* xw, double* cpo, double* cfo, double* forces) {
  const size_t njc = g.njc, nv = prm.nvar();
  const long j1 = tail - 1, nb = (long)g.ni - 2 * tail + 1;
  double Fnp = 0.0, Fcp = 0.0, Fnv = 0.0, Fcv = 0.0;
  const double qinf = 0.5 * prm.rho_inf * (prm.u_inf * prm.u_inf + prm.v_inf * prm.v_inf);
  auto pres = [&](size_t i, size_t j) {
    const double* c = q + (i * njc + j) * nv;
    double r = c[0], uu = c[1] / r, vv = c[2] / r;
    return (c[3] - 0.5 * r * (uu * uu + vv * vv)) * (kGamma - 1.0);
  };
  for (long n = 0; n < nb; ++n) {
    size_t i = (size_t)(j1 + n);
    xw[n] = g.xc(i, 0);
    double cp = (0.5 * (pres(i, 0) + pres(i, 1)) - prm.p_inf) / qinf;
    double ux = value(res.gu_e(i, 0, 0)), uy = value(res.gu_e(i, 0, 1));
    double vx = value(res.gv_e(i, 0, 0)), vy = value(res.gv_e(i, 0, 1));
    double tau = prm.mu_inf * (uy - vx) / qinf;
    cpo[n] = cp; cfo[n] = tau;
    double dx = g.xv(i + 1, 0) - g.xv(i, 0), dy = g.yv(i + 1, 0) - g.yv(i, 0);
    Fnp = Fnp - cp * dx;
    Fcp = Fcp + cp * dy;
    double sfdiv = 2.0 / 3.0 * (ux + vy);
    double sxx = prm.mu_inf * (2.0 * ux - sfdiv) / qinf;
    double syy = prm.mu_inf * (2.0 * vy - sfdiv) / qinf;
    Fnv = Fnv - tau * dy + syy * dx;
    Fcv = Fcv + tau * dx - sxx * dy;
  }
  double ca = std::cos(prm.aoa), sa_ = std::sin(prm.aoa);
  forces[0] = -Fcp * sa_ + Fnp * ca; forces[1] = Fcp * ca + Fnp * sa_;
  forces[2] = -Fcv * sa_ + Fnv * ca; forces[3] = Fcv * ca + Fnv * sa_;
  forces[4] = forces[0] + forces[2]; forces[5] = forces[1] + forces[3];
}

// Wall distance for SA: distance from each cell centre to the nearest no-slip
// wall face ("wall" / "isothermalwall" on the bottom or top face).
inline void wall_distance(const Geometry& g, const std::vector<Bc>& bcs, double* d) {
  std::vector<double> sx0, sy0, sx1, sy1;
  for (const Bc& b : bcs) {
    if (b.type != BC_WALL && b.type != BC_ISOTHERMAL) continue;
    if (b.face != FACE_BOTTOM && b.face != FACE_TOP) continue;
    size_t j = b.face == FACE_BOTTOM ? 0 : g.nj - 1;
    for (long s = std::max<long>(b.start, 1); s <= std::min<long>(b.end, (long)g.nic); ++s) {
      size_t i = (size_t)(s - 1);
      sx0.push_back(g.xv(i, j)); sy0.push_back(g.yv(i, j)); sx1.push_back(g.xv(i + 1, j)); sy1.push_back(g.yv(i + 1, j));
    }
  }
  for (size_t i = 0; i < g.nic; ++i)
    for (size_t j = 0; j < g.njc; ++j) {
      double px = g.xc(i, j), py = g.yc(i, j), best = std::numeric_limits<double>::infinity();
      for (size_t s = 0; s < sx0.size(); ++s) {
        double ex = sx1[s] - sx0[s], ey = sy1[s] - sy0[s];
        double t = ((px - sx0[s]) * ex + (py - sy0[s]) * ey) / (ex * ex + ey * ey);
        t = std::min(1.0, std::max(0.0, t));
        double dx = px - (sx0[s] + t * ex), dy = py - (sy0[s] + t * ey);
        best = std::min(best, std::sqrt(dx * dx + dy * dy));
      }
      d[i * g.njc + j] = best;
    }
}

}  // namespace oracle
