// Fused finite-volume residual on LDS tiles (one pass per tile):
// primitive recovery + ghost boundary cells -> power-law mu/k (+ SA eddy
// viscosity) -> MUSCL/van-Albada-type reconstruction -> Roe or AUSM face flux
// -> Green-Gauss face gradients and viscous flux -> SA source -> assembly /V,
// with the per-block partial sums of R^2 for the L2 norm.
//
// Follows EulerEquation::calc_residual (src/model/eulerequation.cpp:202-232)
// and the kernels it calls: FluidModel::primvars (src/model/fluid.cpp:50-67),
// BoundaryContainer::apply (src/model/bc.cpp:430-433), ReconstructionSecondOrder
// (src/model/reconstruction.cpp:55-154), ConvectiveFluxRoe (src/model/flux.cpp:51-146),
// ConvectiveFluxAUSM (:149-224), Mesh::calc_gradient / calc_face
// (src/utils/mesh.cpp:8-131) and DiffusiveFluxGreenGauss (src/model/flux.cpp:12-48).
//
// Tiling.  A 256-thread block is a 32 x 8 grid of face owners: thread
// (tx, ty) computes the chi face i0+tx on row j0+ty and the eta face j0+ty on
// column i0+tx, so every thread evaluates exactly two faces.  The block then
// assembles the 31 x 7 cells whose four faces it owns; tiles advance by
// (31, 7), so one column and one row of faces per tile are computed twice.
// The cell data of the tile plus its halo (two cells along each face normal
// for MUSCL, one cell around for the viscous stencil) is staged once in LDS
// as primitives and transport properties.
#pragma once
#include "sg_closure.hpp"
#include "sg_device.hpp"
#include "sg_scalar.hpp"

namespace sg {

constexpr int TX = 32, TY = 8, NT = TX * TY;
constexpr int RW = TX + 3, RH = TY + 3, NPOS = RW * RH;  // LDS region, origin (i0-2, j0-2)
constexpr int OX = TX - 1, OY = TY - 1;                    // output cells per tile

template <class S>
struct Prim {
  S r, u, v, p, T, n;
};

// Primitives from the conservative values (rho, rho u, rho v, rho E[, rho nu~]).
template <class S, bool SA>
__device__ __forceinline__ Prim<S> prim_from_cons(const S* qc, const DevPhys& ph) {
  Prim<S> P;
  P.r = qc[0];
  const S ir = srcp(P.r);
  P.u = qc[1] * ir;
  P.v = qc[2] * ir;
  P.p = (qc[3] - 0.5 * P.r * (P.u * P.u + P.v * P.v)) * (kGamma - 1.0);
  P.T = P.p * ir * ph.iR;
  if (SA) P.n = qc[4] * ir;
  else P.n = S(0.0);
  return P;
}

template <class S, bool SA>
__device__ __forceinline__ Prim<S> prim_interior(const Field<S>& q, const DevGeom& g, const DevPhys& ph, int ci, int cj) {
  const size_t idx = (size_t)cj * g.ldi + ci;
  S qc[5];
#pragma unroll
  for (int k = 0; k < (SA ? 5 : 4); ++k) qc[k] = q[k * g.plane + idx];
  return prim_from_cons<S, SA>(qc, ph);
}

// src/model/bc.cpp: freestream :25-60, slipwall :77-128, wall :149-204,
// outflow :283-311, periodic/wake copies :223-263,328-365, isothermal :387-427.
template <class S>
__device__ __forceinline__ Prim<S> apply_op(const BcNode& nd, const Prim<S>& A, const Prim<S>& B, const DevPhys& ph) {
  Prim<S> o;
  switch (nd.op) {
    case OP_FREESTREAM:
      o.r = S(ph.rho_inf); o.u = S(ph.u_inf); o.v = S(ph.v_inf); o.p = S(ph.p_inf);
      o.T = sdiv(o.p, o.r) * ph.iR; o.n = S(ph.nut_inf);
      break;
    case OP_SLIP: {
      const double nx = nd.c0, ny = nd.c1, ids = frcp(nx * nx + ny * ny);
      o.p = 1.5 * A.p - 0.5 * B.p;
      o.r = 1.5 * A.r - 0.5 * B.r;
      S un = A.u * nx + A.v * ny;
      o.u = A.u - 2.0 * un * nx * ids;
      o.v = A.v - 2.0 * un * ny * ids;
      o.T = sdiv(o.p, o.r) * ph.iR;
      o.n = 1.5 * A.n - 0.5 * B.n;
    } break;
    case OP_WALL:
      o.T = 1.5 * A.T - 0.5 * B.T;
      o.r = 1.5 * A.r - 0.5 * B.r;
      o.u = 2.0 * nd.c0 - (1.5 * A.u - 0.5 * B.u);
      o.v = 2.0 * nd.c1 - (1.5 * A.v - 0.5 * B.v);
      o.p = o.r * ph.R * o.T;
      o.n = -(1.5 * A.n - 0.5 * B.n);
      break;
    case OP_ISO:
      o.p = 1.5 * A.p - 0.5 * B.p;
      o.u = 2.0 * nd.c0 - (1.5 * A.u - 0.5 * B.u);
      o.v = 2.0 * nd.c1 - (1.5 * A.v - 0.5 * B.v);
      o.T = S(nd.c2);
      o.r = sdiv(o.p, o.T) * ph.iR;
      o.n = -(1.5 * A.n - 0.5 * B.n);
      break;
    case OP_OUTFLOW:
      o.r = A.r; o.u = A.u; o.v = A.v; o.p = S(ph.p_inf);
      o.T = sdiv(o.p, o.r) * ph.iR; o.n = A.n;
      break;
    default:  // OP_COPY
      o = A;
      break;
  }
  return o;
}

template <class S, bool SA>
__device__ __forceinline__ Prim<S> ref_leaf(int ref, const Field<S>& q, const DevGeom& g, const DevPhys& ph) {
  return prim_interior<S, SA>(q, g, ph, ref % g.nic, ref / g.nic);
}

// Node whose operands are interior cells.
template <class S, bool SA>
__device__ Prim<S> eval_leaf_node(const BcNode& nd, const Field<S>& q, const DevGeom& g, const DevPhys& ph) {
  Prim<S> A, B;
  if (nd.op != OP_FREESTREAM) A = ref_leaf<S, SA>(nd.ra, q, g, ph);
  if (nd.op == OP_SLIP || nd.op == OP_WALL || nd.op == OP_ISO) B = ref_leaf<S, SA>(nd.rb, q, g, ph);
  return apply_op<S>(nd, A, B, ph);
}

// Node whose operands are interior cells or leaf nodes (depth <= 2 is
// guaranteed by the host-side boundary compiler).
template <class S, bool SA>
__device__ Prim<S> eval_node(int id, const DevBc& bc, const Field<S>& q, const DevGeom& g, const DevPhys& ph) {
  const BcNode nd = bc.nodes[id];
  Prim<S> A, B;
  if (nd.op != OP_FREESTREAM) {
    A = nd.ra >= 0 ? ref_leaf<S, SA>(nd.ra, q, g, ph) : eval_leaf_node<S, SA>(bc.nodes[-(nd.ra + 1)], q, g, ph);
  }
  if (nd.op == OP_SLIP || nd.op == OP_WALL || nd.op == OP_ISO) {
    B = nd.rb >= 0 ? ref_leaf<S, SA>(nd.rb, q, g, ph) : eval_leaf_node<S, SA>(bc.nodes[-(nd.rb + 1)], q, g, ph);
  }
  return apply_op<S>(nd, A, B, ph);
}

__device__ __forceinline__ int ghost_index(int ci, int cj, int nic, int njc) {
  const int gi = ci + 1, gj = cj + 1;
  if (gj == 0) return gi;
  if (gj == njc + 1) return (nic + 2) + gi;
  if (gi == 0) return 2 * (nic + 2) + gj - 1;
  return 2 * (nic + 2) + njc + gj - 1;
}

// A pure tangent (value 0) for the dual, nothing for plain doubles.
template <class S>
__device__ __forceinline__ S beta_tangent(double t);
template <>
__device__ __forceinline__ double beta_tangent<double>(double) { return 0.0; }
template <>
__device__ __forceinline__ Dual beta_tangent<Dual>(double t) { return Dual(0.0, t); }

// LDS slots of the staged cell data.
enum { L_R = 0, L_U, L_V, L_P, L_T, L_MU, L_K, L_N, L_ETA, L_ML, L_MAX };

// MUSCL left/right states, src/model/reconstruction.cpp:142-219.
// Left state of a face from cell c with neighbours cm (behind) and cp (ahead):
//   dm = c - cm, dp = cp - c, s = 0.25 (3 dp dm + eps) / (2 (dp - dm)^2 + 3 dp dm + eps)
//   left  = c + s (2/3 dm + 4/3 dp);   right = c - s (4/3 dm + 2/3 dp)
template <class S>
__device__ __forceinline__ S muscl_left(S cm, S c, S cp, double eps) {
  const double thm = 2.0 / 3.0, thp = 4.0 / 3.0;
  S dm = c - cm, dp = cp - c;
  S a1 = 3.0 * dp * dm;
  S a2 = 2.0 * (dp - dm) * (dp - dm) + a1;
  S s = 0.25 * sdiv(a1 + eps, a2 + eps);
  return c + s * (thm * dm + thp * dp);
}
template <class S>
__device__ __forceinline__ S muscl_right(S cm, S c, S cp, double eps) {
  const double thm = 2.0 / 3.0, thp = 4.0 / 3.0;
  S dm = c - cm, dp = cp - c;
  S a1 = 3.0 * dp * dm;
  S a2 = 2.0 * (dp - dm) * (dp - dm) + a1;
  S s = 0.25 * sdiv(a1 + eps, a2 + eps);
  return c - s * (thp * dm + thm * dp);
}

// Roe flux without entropy fix, src/model/flux.cpp:51-146; with SA the
// rho*nu~ transport is first-order upwind on the Roe mass flux.
template <class S, bool SA>
__device__ __forceinline__ void roe_flux(double nx, double ny, S rl, S ul, S vl, S pl, S rr, S ur, S vr, S pr, S ntl,
                                         S ntr, S* f) {
  const double gm1 = kGamma - 1.0, ogm1 = 1.0 / gm1;
  S rli = srcp(rl), rul = rl * ul, rvl = rl * vl, uvl = 0.5 * (ul * ul + vl * vl);
  S el = pl * ogm1 + rl * uvl, hl = (el + pl) * rli;
  S rri = srcp(rr), rur = rr * ur, rvr = rr * vr, uvr = 0.5 * (ur * ur + vr * vr);
  S er = pr * ogm1 + rr * uvr, hr = (er + pr) * rri;
  S rat = ssqrt(rr * rli), rati = srcp(rat + 1.0);
  S rav = rat * rl, uav = (rat * ur + ul) * rati, vav = (rat * vr + vl) * rati, hav = (rat * hr + hl) * rati;
  S uv = 0.5 * (uav * uav + vav * vav);
  S cav = ssqrt(gm1 * (hav - uv));
  S d1 = rr - rl, d2 = ur - ul, d3 = vr - vl, d4 = pr - pl;
  const double dr = fsqrt(nx * nx + ny * ny), idr = frcp(dr), r1 = nx * idr, r2 = ny * idr;
  S uu = r1 * uav + r2 * vav, c2 = cav * cav, c2i = srcp(c2);
  S auu = sabs(uu), aupc = sabs(uu + cav), aumc = sabs(uu - cav);
  S uul = r1 * ul + r2 * vl, uur = r1 * ur + r2 * vr;
  S rcav = rav * cav, aquu = uur - uul, c2ih = 0.5 * c2i, ruuav = auu * rav;
  S b1 = auu * (d1 - c2i * d4);
  S b2 = c2ih * aupc * (d4 + rcav * aquu);
  S b3 = c2ih * aumc * (d4 - rcav * aquu);
  S b4 = b1 + b2 + b3, b5 = cav * (b2 - b3);
  S b6 = ruuav * (d2 - r1 * aquu), b7 = ruuav * (d3 - r2 * aquu);
  S a2 = uav * b4 + r1 * b5 + b6, a3 = vav * b4 + r2 * b5 + b7;
  S a4 = hav * b4 + uu * b5 + uav * b6 + vav * b7 - c2 * b1 * ogm1;
  const double aj = 0.5 * dr;
  S plar = pl + pr, epl = el + pl, epr = er + pr;
  f[0] = aj * (rl * uul + rr * uur - b4);
  f[1] = aj * (rul * uul + rur * uur + r1 * plar - a2);
  f[2] = aj * (rvl * uul + rvr * uur + r2 * plar - a3);
  f[3] = aj * (epl * uul + epr * uur - a4);
  if (SA) f[4] = 0.5 * (f[0] * (ntl + ntr) - sabs(f[0]) * (ntr - ntl));
}

// AUSM flux (Liou-Steffen), src/model/flux.cpp:149-224: split Mach numbers
// M+-(M) = +-(M +- 1)^2 / 4 and pressures P+-(M) = p (M +- 1)^2 (2 -+ M) / 4
// for |M| <= 1, otherwise the upwind values (M+ = M, P+ = p for M > 1 and 0
// for M < -1; mirrored for M-P-); the interface Mach number picks the side
// whose rho a (1, u, v, h) is convected.  With SA, rho*nu~ is carried
// first-order upwind by the AUSM mass flux, as for Roe.
template <class S, bool SA>
__device__ __forceinline__ void ausm_flux(double nx, double ny, S rl, S ul, S vl, S pl, S rr, S ur, S vr, S pr, S ntl,
                                          S ntr, S* f) {
  const double ds = fsqrt(nx * nx + ny * ny), ids = frcp(ds), gg = kGamma / (kGamma - 1.0);
  S rli = srcp(rl), rri = srcp(rr);
  S al = ssqrt(kGamma * pl * rli), ar = ssqrt(kGamma * pr * rri);
  S Ml = (ul * nx + vl * ny) * ids * srcp(al), Mr = (ur * nx + vr * ny) * ids * srcp(ar);
  S mp, pp, mm, pm;
  if (fabs(val(Ml)) <= 1.0) {
    S a = Ml + 1.0, a2 = 0.25 * a * a;
    mp = a2; pp = pl * a2 * (2.0 - Ml);
  } else if (val(Ml) > 0.0) {
    mp = Ml; pp = pl;
  } else {
    mp = S(0.0); pp = S(0.0);
  }
  if (fabs(val(Mr)) <= 1.0) {
    S b = Mr - 1.0, b2 = 0.25 * b * b;
    mm = -b2; pm = pr * b2 * (2.0 + Mr);
  } else if (val(Mr) < 0.0) {
    mm = Mr; pm = pr;
  } else {
    mm = S(0.0); pm = S(0.0);
  }
  S mh = mp + mm, ph = pp + pm;
  const bool up = val(mh) >= 0.0;
  S r = up ? rl : rr, a = up ? al : ar, u = up ? ul : ur, v = up ? vl : vr;
  S h = gg * (up ? pl * rli : pr * rri) + 0.5 * (u * u + v * v);
  S m = r * a * mh * ds;
  f[0] = m;
  f[1] = m * u + ph * nx;
  f[2] = m * v + ph * ny;
  f[3] = m * h;
  if (SA) f[4] = 0.5 * (f[0] * (ntl + ntr) - sabs(f[0]) * (ntr - ntl));
}

template <class S, bool SA>
__device__ __forceinline__ void conv_flux(int kind, double nx, double ny, S rl, S ul, S vl, S pl, S rr, S ur, S vr,
                                          S pr, S ntl, S ntr, S* f) {
  if (kind == FLUX_AUSM) ausm_flux<S, SA>(nx, ny, rl, ul, vl, pl, rr, ur, vr, pr, ntl, ntr, f);
  else roe_flux<S, SA>(nx, ny, rl, ul, vl, pl, rr, ur, vr, pr, ntl, ntr, f);
}

template <class S>
struct FaceGeo {
  double tx, ty, bx, by, lx, ly, rx, ry, ivol;  // top/bottom/left/right normals, 1/auxiliary volume
};

// Green-Gauss face gradient (src/utils/mesh.cpp:34-131) of one variable from
// the face's left/right cells and its top/bottom 4-cell averages.
template <class S>
__device__ __forceinline__ void gg_grad(const FaceGeo<S>& G, S qt, S qb, S ql, S qr, S* out) {
  out[0] = (G.tx * qt - G.bx * qb + G.rx * qr - G.lx * ql) * G.ivol;
  out[1] = (G.ty * qt - G.by * qb + G.ry * qr - G.ly * ql) * G.ivol;
}

// DiffusiveFluxGreenGauss::evaluate, src/model/flux.cpp:12-48 (+ SA nu~ diffusion).
template <class S, bool SA>
__device__ __forceinline__ void visc_flux(double nx, double ny, const S* gu, const S* gv, const S* gT, S ub, S vb, S m,
                                          S k, const S* gn, S et, S* f) {
  S txy = m * (gu[1] + gv[0]);
  S txx = m * (2.0 * gu[0] - 2.0 / 3.0 * (gu[0] + gv[1]));
  S tyy = m * (2.0 * gv[1] - 2.0 / 3.0 * (gu[0] + gv[1]));
  S qx = -k * gT[0], qy = -k * gT[1];
  f[1] = txx * nx + txy * ny;
  f[2] = txy * nx + tyy * ny;
  f[3] = (ub * txx + vb * txy - qx) * nx + (ub * txy + vb * tyy - qy) * ny;
  if (SA) f[4] = (1.0 / kSigma) * et * (gn[0] * nx + gn[1] * ny);
}

template <class S, bool SA>
__device__ __forceinline__ void transport(const Prim<S>& P, const DevPhys& ph, S* slot) {
  S ml = ph.mu_inf * spow23(P.T * ph.iTinf);
  if (!SA) {
    slot[L_MU] = ml;
    slot[L_K] = ml * ph.cp * ph.ipr;
    return;
  }
  S rn = P.r * P.n;
  S chi = sdiv(rn, ml), chi3 = chi * chi * chi;
  S mut(0.0), fn(1.0);
  if (val(P.n) >= 0.0) mut = rn * sdiv(chi3, chi3 + kCv1 * kCv1 * kCv1);
  else fn = sdiv(kCn1 + chi3, kCn1 - chi3);
  slot[L_MU] = ml + mut;
  slot[L_K] = ph.cp * (ml * ph.ipr + mut * (1.0 / kPrt));
  slot[L_ETA] = ml + rn * fn;
  slot[L_ML] = ml;
}

// SA-noft2 source (production * beta - destruction + cb2 and density-gradient
// diffusion terms, NASA TMR compressible form) at one cell; the cell
// gradients are Green-Gauss over its four faces with arithmetic face means.
// Split as src = beta * P1 - D + rest so that a closure can supply beta.
template <class S>
struct SaTerms {
  S P1, D, rest;
};
template <class S>
__device__ __forceinline__ SaTerms<S> sa_terms(S r, S n, S ml, const S* gu, const S* gv, const S* gn, const S* gr,
                                               double d) {
  const double cw1 = kCb1 / (kKappa * kKappa) + (1.0 + kCb2) / kSigma;
  S nu = sdiv(ml, r);
  S omega = sabs(gv[0] - gu[1]);
  const double k2d2 = kKappa * kKappa * d * d, ik2d2 = frcp(k2d2), id = frcp(d);
  SaTerms<S> t;
  if (val(n) >= 0.0) {
    S chi = sdiv(n, nu), chi3 = chi * chi * chi;
    S fv1 = sdiv(chi3, chi3 + kCv1 * kCv1 * kCv1);
    S fv2 = 1.0 - sdiv(chi, 1.0 + chi * fv1);
    S sbar = n * fv2 * ik2d2;
    S stil = (val(sbar) >= -kCv2 * val(omega))
                 ? omega + sbar
                 : omega + sdiv(omega * (kCv2 * kCv2 * omega + kCv3 * sbar), (kCv3 - 2.0 * kCv2) * omega - sbar);
    S rr = (val(stil) > 0.0) ? smin(sdiv(n, stil * k2d2), 10.0) : S(10.0);
    S r2 = rr * rr, r6 = r2 * r2 * r2;
    S gg = rr + kCw2 * (r6 - rr);
    S g2 = gg * gg, g6 = g2 * g2 * g2;
    const double cw3_6 = kCw3 * kCw3 * kCw3 * kCw3 * kCw3 * kCw3;
    S fw = gg * spow16(sdiv(S(1.0 + cw3_6), g6 + cw3_6));
    t.P1 = kCb1 * stil * r * n;
    t.D = cw1 * fw * r * (n * id) * (n * id);
  } else {
    t.P1 = kCb1 * (1.0 - kCt3) * omega * r * n;
    t.D = -cw1 * r * (n * id) * (n * id);
  }
  S cb2t = (kCb2 / kSigma) * r * (gn[0] * gn[0] + gn[1] * gn[1]);
  S cross = -(1.0 / kSigma) * (nu + n) * (gr[0] * gn[0] + gr[1] * gn[1]);
  t.rest = cb2t + cross;
  return t;
}

// Fused norm output: with l2 set, the block that finishes last finalises the
// norms.  state holds the arrival counters (kNormStateWords words: one per
// tile group, then the count of groups done; all 0 between launches); parity
// alternates between consecutive fused launches on the same partials.
struct NormOut {
  unsigned long long* state;
  double* l2;         // ||R_k||_2 (k < NV), then ||V R_k||_2 (NV <= k < 2 NV)
  int count;          // entries of l2 written: NV (caller's buffer) or 2 NV
  unsigned parity;
};

// Spin budget of the finalising block per thread (polls beyond the first
// load of each partial); past it the norms are written as NaN.
constexpr int kNormSpinLimit = 1 << 22;

// CL: closure kind of beta (ClosureKind); CLOSURE_MLP stages the hidden
// layer of the network (and its tangent) in the same LDS area.
template <class S, bool SA, int CL>
__global__ __launch_bounds__(NT, sizeof(S) == sizeof(double) ? 1024 / NT : 1) void residual_kernel(DevGeom g, DevPhys ph, DevBc bc, Field<S> q, OutField<S> out,
                                                       double* partial, const double* beta_dot, DevClosure clo,
                                                       NormOut nout) {
  constexpr int NL = SA ? L_MAX : (L_K + 1);
  constexpr int NV = SA ? 5 : 4;
  constexpr int NLDS0 = NL * NPOS > 2 * NV * NT ? NL * NPOS : 2 * NV * NT;
  constexpr int NLDS_D = NLDS0 * (int)(sizeof(S) / sizeof(double));
  constexpr bool TAN = sizeof(S) > sizeof(double);
  constexpr int NMLP_D = (CL == CLOSURE_MLP) ? (TAN ? 2 : 1) * (NT / 64) * 64 * kHld / 4 : 0;
  __shared__ double lds_raw[NLDS_D > NMLP_D ? NLDS_D : NMLP_D];
  S* lds = reinterpret_cast<S*>(lds_raw);
  __shared__ double red[NT / 64][2 * NV];

  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  // XCD-aware tile order: work-groups are handed to the 8 XCDs round-robin in
  // linear order, so XCD x's k-th group takes the k-th tile of the x-th
  // contiguous run of tiles; neighbouring tiles then share their halo rows in
  // the same L2.  The map is a bijection on the tiles for any tile count.
  int tile = blockIdx.y * gridDim.x + blockIdx.x;
  {
    const int N = gridDim.x * gridDim.y, x = tile & 7, k = tile >> 3, qn = N >> 3, rn = N & 7;
    tile = (x < rn) ? x * (qn + 1) + k : rn * (qn + 1) + (x - rn) * qn + k;
  }
  const int i0 = (tile % gridDim.x) * OX, j0 = (tile / gridDim.x) * OY;
  const int nic = g.nic, njc = g.njc;
  const double nan = __builtin_nan("");
  const bool fused = partial && nout.l2;

  // ---- stage primitives (+ transport properties) of tile + halo in LDS.
  // A thread stages at most NPASS positions; the conservative values of its
  // interior positions are loaded first, all in flight together, and only
  // then converted.
  constexpr int NPASS = (NPOS + NT - 1) / NT;
  constexpr int NQ = SA ? 5 : 4;
  S qraw[NPASS][NQ];
#pragma unroll
  for (int pass = 0; pass < NPASS; ++pass) {
    const int p = tid + pass * NT;
    const int ci = i0 - 2 + p % RW, cj = j0 - 2 + p / RW;
    const bool interior = p < NPOS && ci >= 0 && ci < nic && cj >= 0 && cj < njc;
#pragma unroll
    for (int k = 0; k < NQ; ++k) qraw[pass][k] = S(0.0);
    if (interior) {
      const size_t idx = (size_t)cj * g.ldi + ci;
#pragma unroll
      for (int k = 0; k < NQ; ++k) qraw[pass][k] = q[k * g.plane + idx];
    }
  }
#pragma unroll
  for (int pass = 0; pass < NPASS; ++pass) {
    const int p = tid + pass * NT;
    if (p >= NPOS) break;
    const int lx = p % RW, ly = p / RW;
    const int ci = i0 - 2 + lx, cj = j0 - 2 + ly;
    S slot[L_MAX];
    const bool interior = ci >= 0 && ci < nic && cj >= 0 && cj < njc;
    const bool ring = !interior && ci >= -1 && ci <= nic && cj >= -1 && cj <= njc;
    int node = -1;
    if (ring) node = bc.ghost_map[ghost_index(ci, cj, nic, njc)];
    if (interior || node >= 0) {
      Prim<S> P = interior ? prim_from_cons<S, SA>(qraw[pass], ph) : eval_node<S, SA>(node, bc, q, g, ph);
      slot[L_R] = P.r; slot[L_U] = P.u; slot[L_V] = P.v; slot[L_P] = P.p; slot[L_T] = P.T;
      if (SA) slot[L_N] = P.n;
      const bool box = lx >= 1 && lx < RW - 1 && ly >= 1 && ly < RH - 1;
      if (ph.viscous && box) transport<S, SA>(P, ph, slot);
      else { slot[L_MU] = S(nan); slot[L_K] = S(nan); if (SA) { slot[L_ETA] = S(nan); slot[L_ML] = S(nan); } }
    } else {
#pragma unroll
      for (int k = 0; k < L_MAX; ++k) slot[k] = S(nan);
    }
#pragma unroll
    for (int k = 0; k < NL; ++k) lds[k * NPOS + p] = slot[k];
  }
  __syncthreads();

  auto at = [&](int k, int lx, int ly) -> S { return lds[k * NPOS + ly * RW + lx]; };
  auto ncx = [&](int fi, int cj) { return g.ncx[(size_t)cj * g.ldf + fi]; };
  auto ncy = [&](int fi, int cj) { return g.ncy[(size_t)cj * g.ldf + fi]; };
  auto nex = [&](int ci, int fj) { return g.nex[(size_t)fj * g.ldi + ci]; };
  auto ney = [&](int ci, int fj) { return g.ney[(size_t)fj * g.ldi + ci]; };
  auto vol = [&](int ci, int cj) { return g.vol[(size_t)cj * g.ldi + ci]; };

  S hc[NV], he[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) { hc[k] = S(0.0); he[k] = S(0.0); }

  // ---- chi face fi = i0+tx on row cj = j0+ty (between cells fi-1 and fi)
  {
    const int fi = i0 + tx, cj = j0 + ty;
    if (fi <= nic && cj < njc) {
      const int xl = tx + 1, xr = tx + 2, y = ty + 2;  // LDS coords of the left/right cells
      S st[2][5];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        S cl = at(L_R + k, xl, y), cr = at(L_R + k, xr, y);
        if (ph.order == 2) {
          st[0][k] = (fi >= 1) ? muscl_left(at(L_R + k, xl - 1, y), cl, cr, ph.eps_chi) : cl;
          st[1][k] = (fi <= nic - 1) ? muscl_right(cl, cr, at(L_R + k, xr + 1, y), ph.eps_chi) : cr;
        } else {
          st[0][k] = cl; st[1][k] = cr;
        }
      }
      S ntl(0.0), ntr(0.0);
      if (SA) { ntl = at(L_N, xl, y); ntr = at(L_N, xr, y); }
      const double nx = ncx(fi, cj), ny = ncy(fi, cj);
      conv_flux<S, SA>(ph.flux, nx, ny, st[0][0], st[0][1], st[0][2], st[0][3], st[1][0], st[1][1], st[1][2], st[1][3],
                       ntl, ntr, hc);
      if (ph.viscous) {
        FaceGeo<S> G;
        if (fi == 0) {
          G.tx = nex(0, cj + 1); G.ty = ney(0, cj + 1); G.bx = nex(0, cj); G.by = ney(0, cj);
          G.rx = 0.5 * (ncx(0, cj) + ncx(1, cj)); G.ry = 0.5 * (ncy(0, cj) + ncy(1, cj));
          G.lx = ncx(0, cj); G.ly = ncy(0, cj); G.ivol = frcp(vol(0, cj));
        } else if (fi == nic) {
          G.tx = nex(nic - 1, cj + 1); G.ty = ney(nic - 1, cj + 1); G.bx = nex(nic - 1, cj); G.by = ney(nic - 1, cj);
          G.rx = ncx(nic, cj); G.ry = ncy(nic, cj);
          G.lx = 0.5 * (ncx(nic, cj) + ncx(nic - 1, cj)); G.ly = 0.5 * (ncy(nic, cj) + ncy(nic - 1, cj));
          G.ivol = frcp(vol(nic - 1, cj));
        } else {
          G.tx = 0.5 * (nex(fi, cj + 1) + nex(fi - 1, cj + 1)); G.ty = 0.5 * (ney(fi, cj + 1) + ney(fi - 1, cj + 1));
          G.bx = 0.5 * (nex(fi, cj) + nex(fi - 1, cj)); G.by = 0.5 * (ney(fi, cj) + ney(fi - 1, cj));
          G.rx = 0.5 * (ncx(fi, cj) + ncx(fi + 1, cj)); G.ry = 0.5 * (ncy(fi, cj) + ncy(fi + 1, cj));
          G.lx = 0.5 * (ncx(fi, cj) + ncx(fi - 1, cj)); G.ly = 0.5 * (ncy(fi, cj) + ncy(fi - 1, cj));
          G.ivol = frcp(0.5 * (vol(fi, cj) + vol(fi - 1, cj)));
        }
        // left/right cells and their north/south neighbours
        auto grad_avg = [&](int k, S* gr, S* avg) {
          S a = at(k, xl, y), b = at(k, xr, y);
          S qt = 0.25 * (a + b + at(k, xl, y + 1) + at(k, xr, y + 1));
          S qb = 0.25 * (a + b + at(k, xl, y - 1) + at(k, xr, y - 1));
          if (gr) gg_grad(G, qt, qb, a, b, gr);
          if (avg) *avg = 0.25 * (a + b + qt + qb);
        };
        S gu[2], gv[2], gT[2], gn[2] = {S(0.0), S(0.0)}, ub, vb, mb, kb, eb(0.0);
        grad_avg(L_U, gu, &ub);
        grad_avg(L_V, gv, &vb);
        grad_avg(L_T, gT, nullptr);
        grad_avg(L_MU, nullptr, &mb);
        grad_avg(L_K, nullptr, &kb);
        if (SA) { grad_avg(L_N, gn, nullptr); grad_avg(L_ETA, nullptr, &eb); }
        S fv[5];
        visc_flux<S, SA>(nx, ny, gu, gv, gT, ub, vb, mb, kb, gn, eb, fv);
#pragma unroll
        for (int k = 1; k < NV; ++k) hc[k] = hc[k] - fv[k];
      }
    }
  }

  // ---- eta face fj = j0+ty on column ci = i0+tx (between cells fj-1 and fj)
  {
    const int fj = j0 + ty, ci = i0 + tx;
    if (fj <= njc && ci < nic) {
      const int x = tx + 2, ys = ty + 1, yn = ty + 2;
      S st[2][5];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        S cs = at(L_R + k, x, ys), cn = at(L_R + k, x, yn);
        if (ph.order == 2) {
          st[0][k] = (fj >= 1) ? muscl_left(at(L_R + k, x, ys - 1), cs, cn, ph.eps_eta) : cs;
          st[1][k] = (fj <= njc - 1) ? muscl_right(cs, cn, at(L_R + k, x, yn + 1), ph.eps_eta) : cn;
        } else {
          st[0][k] = cs; st[1][k] = cn;
        }
      }
      S nts(0.0), ntn(0.0);
      if (SA) { nts = at(L_N, x, ys); ntn = at(L_N, x, yn); }
      const double nx = nex(ci, fj), ny = ney(ci, fj);
      conv_flux<S, SA>(ph.flux, nx, ny, st[0][0], st[0][1], st[0][2], st[0][3], st[1][0], st[1][1], st[1][2], st[1][3],
                       nts, ntn, he);
      if (ph.viscous) {
        FaceGeo<S> G;
        if (fj == 0) {
          G.tx = 0.5 * (nex(ci, 0) + nex(ci, 1)); G.ty = 0.5 * (ney(ci, 0) + ney(ci, 1));
          G.bx = nex(ci, 0); G.by = ney(ci, 0);
          G.lx = ncx(ci, 0); G.ly = ncy(ci, 0); G.rx = ncx(ci + 1, 0); G.ry = ncy(ci + 1, 0);
          G.ivol = frcp(vol(ci, 0));
        } else if (fj == njc) {
          G.tx = nex(ci, njc); G.ty = ney(ci, njc);
          G.bx = 0.5 * (nex(ci, njc) + nex(ci, njc - 1)); G.by = 0.5 * (ney(ci, njc) + ney(ci, njc - 1));
          G.lx = ncx(ci, njc - 1); G.ly = ncy(ci, njc - 1); G.rx = ncx(ci + 1, njc - 1); G.ry = ncy(ci + 1, njc - 1);
          G.ivol = frcp(vol(ci, njc - 1));
        } else {
          G.tx = 0.5 * (nex(ci, fj) + nex(ci, fj + 1)); G.ty = 0.5 * (ney(ci, fj) + ney(ci, fj + 1));
          G.bx = 0.5 * (nex(ci, fj) + nex(ci, fj - 1)); G.by = 0.5 * (ney(ci, fj) + ney(ci, fj - 1));
          G.lx = 0.5 * (ncx(ci, fj) + ncx(ci, fj - 1)); G.ly = 0.5 * (ncy(ci, fj) + ncy(ci, fj - 1));
          G.rx = 0.5 * (ncx(ci + 1, fj) + ncx(ci + 1, fj - 1)); G.ry = 0.5 * (ncy(ci + 1, fj) + ncy(ci + 1, fj - 1));
          G.ivol = frcp(0.5 * (vol(ci, fj) + vol(ci, fj - 1)));
        }
        auto grad_avg = [&](int k, S* gr, S* avg) {
          S t = at(k, x, yn), b = at(k, x, ys);
          S l = 0.25 * (t + b + at(k, x - 1, yn) + at(k, x - 1, ys));
          S r = 0.25 * (t + b + at(k, x + 1, yn) + at(k, x + 1, ys));
          if (gr) gg_grad(G, t, b, l, r, gr);
          if (avg) *avg = 0.25 * (l + r + t + b);
        };
        S gu[2], gv[2], gT[2], gn[2] = {S(0.0), S(0.0)}, ub, vb, mb, kb, eb(0.0);
        grad_avg(L_U, gu, &ub);
        grad_avg(L_V, gv, &vb);
        grad_avg(L_T, gT, nullptr);
        grad_avg(L_MU, nullptr, &mb);
        grad_avg(L_K, nullptr, &kb);
        if (SA) { grad_avg(L_N, gn, nullptr); grad_avg(L_ETA, nullptr, &eb); }
        S fv[5];
        visc_flux<S, SA>(nx, ny, gu, gv, gT, ub, vb, mb, kb, gn, eb, fv);
#pragma unroll
        for (int k = 1; k < NV; ++k) he[k] = he[k] - fv[k];
      }
    }
  }

  // ---- SA source of the owned output cell (needs the staged cells)
  const int ci = i0 + tx, cj = j0 + ty;
  const bool owns = tx < OX && ty < OY && ci < nic && cj < njc;
  S src(0.0);
  if (SA && ph.viscous) {  // uniform over the block
    const size_t cidx = (size_t)cj * g.ldi + ci;
    SaTerms<S> st{S(0.0), S(0.0), S(0.0)};
    S feat[kNFeat];
    if (owns) {
      const int x = tx + 2, y = ty + 2;
      const double ex = ncx(ci + 1, cj), ey = ncy(ci + 1, cj), wx = ncx(ci, cj), wy = ncy(ci, cj);
      const double nx_ = nex(ci, cj + 1), ny_ = ney(ci, cj + 1), sx = nex(ci, cj), sy = ney(ci, cj);
      const double iV = frcp(vol(ci, cj));
      auto cgrad = [&](int k, S* o) {
        S c = at(k, x, y);
        S fe = 0.5 * (c + at(k, x + 1, y)), fw = 0.5 * (c + at(k, x - 1, y));
        S fn = 0.5 * (c + at(k, x, y + 1)), fs = 0.5 * (c + at(k, x, y - 1));
        o[0] = (fe * ex - fw * wx + fn * nx_ - fs * sx) * iV;
        o[1] = (fe * ey - fw * wy + fn * ny_ - fs * sy) * iV;
      };
      S gu[2], gv[2], gn[2], gr[2];
      cgrad(L_U, gu); cgrad(L_V, gv); cgrad(L_N, gn); cgrad(L_R, gr);
      const S r = at(L_R, x, y), n = at(L_N, x, y), ml = at(L_ML, x, y);
      st = sa_terms<S>(r, n, ml, gu, gv, gn, gr, g.wd[cidx]);
      if (CL != CLOSURE_FIELD || clo.aux) {
        S gp[2];
        cgrad(L_P, gp);
        sa_features<S>(r, at(L_U, x, y), at(L_V, x, y), n, sdiv(ml, r), gu, gv, gp, g.wd[cidx], st.P1, st.D, ph, feat);
      }
    }
    S bcell(1.0);
    if (CL == CLOSURE_FIELD) {
      if (owns) {
        bcell = S(g.beta[cidx]);
        if (beta_dot) bcell = bcell + beta_tangent<S>(beta_dot[cidx]);
      }
    } else if (CL == CLOSURE_ANALYTIC) {
      if (owns) bcell = closure_analytic<S>(clo, feat);
    } else {
      float xv[kNFeat], xt[kNFeat];
#pragma unroll
      for (int k = 0; k < kNFeat; ++k) {
        xv[k] = owns ? (float)val(feat[k]) : 0.f;
        xt[k] = owns ? (float)tan_(feat[k]) : 0.f;
      }
      __syncthreads();  // staged cells no longer read: reuse LDS for the hidden layer
      __bf16* hs = reinterpret_cast<__bf16*>(lds_raw) + (tid >> 6) * 64 * kHld;
      __bf16* ht = hs + (NT / 64) * 64 * kHld;
      float y, yt;
      mlp_wave<TAN>(clo.params, reinterpret_cast<const __bf16*>(clo.w2), xv, xt, hs, ht, y, yt);
      bcell = S(1.0 + (double)y);
      if (TAN) bcell = bcell + beta_tangent<S>((double)yt);
    }
    if (owns) {
      src = bcell * st.P1 - st.D + st.rest;
      if (clo.aux) {
#pragma unroll
        for (int k = 0; k < kNFeat; ++k) clo.aux[k * g.plane + cidx] = val(feat[k]);
        clo.aux[kNFeat * g.plane + cidx] = val(bcell);
      }
    }
  }

  // ---- exchange face fluxes through LDS (reusing the staging area)
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    lds[k * NT + tid] = hc[k];
    lds[(NV + k) * NT + tid] = he[k];
  }
  __syncthreads();

  double sq[2 * NV];  // R^2 (reference norm) and (V R)^2 (flux-balance norm)
  S res[NV];
  const size_t oidx = (size_t)cj * g.ldi + ci;
#pragma unroll
  for (int k = 0; k < 2 * NV; ++k) sq[k] = 0.0;
  if (owns) {
    const double V = vol(ci, cj), iV = frcp(V);
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      S hw = lds[k * NT + tid], hE = lds[k * NT + tid + 1];
      S hs = lds[(NV + k) * NT + tid], hn = lds[(NV + k) * NT + tid + TX];
      S r = S(0.0) - (hn - hs);
      r = r - (hE - hw);
      if (k == 1) r = r + (-ph.dpdx * V);
      if (k == 2) r = r + (-ph.dpdy * V);
      if (SA && k == 4) r = r + V * src;
      r = r * iV;
      res[k] = r;
      sq[k] = val(r) * val(r);
      sq[NV + k] = sq[k] * V * V;
    }
  }
  // Fused norm: count this block's arrival before its stores.  Vector memory
  // operations complete in issue order, so a wait for the count's return
  // issued after the stores would also wait for them to be acknowledged; the
  // parity protocol below makes the order of arrival and stores irrelevant.
  // Blocks count on the counter of their group (tile index modulo
  // kNormGroups, one 128-byte line each), so that arrivals on one word come
  // from one XCD's blocks only; the last of a group counts on the group
  // counter, and the last group's last block finalises.
  const int grp = (int)(blockIdx.y * gridDim.x + blockIdx.x) % kNormGroups;
  unsigned long long arrival = 0;
  if (fused && tid == 0)
    arrival = __hip_atomic_fetch_add(nout.state + grp * kNormLineWords, 1ull, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
  if (owns) {
#pragma unroll
    for (int k = 0; k < NV; ++k) out.put(k * g.plane + oidx, res[k]);
  }
  // ---- block partial sums of R^2 (fixed order: wavefront DPP sum, then waves)
  if (partial) {
    const int lane = tid & 63, wv = tid >> 6;
    const int nb = gridDim.x * gridDim.y, bid = blockIdx.y * gridDim.x + blockIdx.x;
    // only the norms the caller asked for (the first nout.count of R^2 then
    // (V R)^2) are reduced and summed
    const int nsum = nout.count;
#pragma unroll
    for (int k = 0; k < 2 * NV; ++k)
      if (k < nsum) sq[k] = wave_sum_lane63(sq[k]);
    if (lane == 63) {
#pragma unroll
      for (int k = 0; k < 2 * NV; ++k) red[wv][k] = sq[k];
    }
    __syncthreads();
    // Fused: the sign bit of each stored partial (a sum of squares, >= 0)
    // carries the launch parity -- set on launches of parity 0 -- so the
    // finalising block recognises this launch's value in every slot from
    // that slot alone.  A block therefore counts its arrival without waiting
    // for its partial stores, and no ordering between slots, or between a
    // slot and the counter, is needed: each slot is one 64-bit atomic store,
    // read by atomic loads of the same location, and launches on a stream
    // are ordered by the kernel boundary.  Slots start at +0.0 (parity 1), so
    // a slot never written in this context never passes for parity 0.
    const bool want_neg = fused && (nout.parity & 1u) == 0u;
    __shared__ bool last;
    // every slot is stored on every fused launch (zero beyond nsum), so that
    // a slot always holds this launch's parity or the previous launch's
    if (tid < 2 * NV) {
      double s = 0.0;
      if (tid < nsum)
        for (int w = 0; w < NT / 64; ++w) s += red[w][tid];
      if (want_neg) s = -s;
      __hip_atomic_store(&partial[(size_t)tid * nb + bid], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (fused && tid == 0) {
      // blocks of group grp: grp, grp + G, ... < nb; groups with a block: min(nb, G)
      const unsigned long long gsize = (unsigned long long)((nb - grp + kNormGroups - 1) / kNormGroups);
      bool l = false;
      if (arrival == gsize - 1) {
        const unsigned long long gdone = __hip_atomic_fetch_add(nout.state + kNormGroups * kNormLineWords, 1ull,
                                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        l = gdone == (unsigned long long)(min(nb, kNormGroups) - 1);
      }
      last = l;
    }
    if (fused) __syncthreads();
    if (fused && last) {
      // The block that arrives last sums the partials of all blocks in a
      // fixed order (so the norm does not depend on which block it is),
      // polling each slot until it holds this launch's value.  Every other
      // block has arrived (each group counter reached its group's size before
      // the group count reached the number of groups: single-location
      // read-modify-write order), so it is running and stores its partials,
      // and the wait ends.
      int spins = 0;
      double acc[2 * NV];
#pragma unroll
      for (int k = 0; k < 2 * NV; ++k) acc[k] = 0.0;
      // the plain norms, then (if asked for) the weighted ones; four rows of
      // blocks per round (one round up to 1,024 blocks), all loads of a round
      // in flight together
      constexpr int NR = 4;
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        if (half * NV >= nsum) break;
        for (int b0 = 0; b0 < nb; b0 += NR * NT) {
          double v[NR][NV];
#pragma unroll
          for (int u = 0; u < NR; ++u) {
            const int b = b0 + u * NT + tid;
#pragma unroll
            for (int kk = 0; kk < NV; ++kk) {
              const int k = half * NV + kk;
              v[u][kk] = (b < nb && k < nsum) ? __hip_atomic_load(&partial[(size_t)k * nb + b], __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_AGENT)
                                              : (want_neg ? -0.0 : 0.0);
            }
          }
#pragma unroll
          for (int u = 0; u < NR; ++u) {
            const int b = b0 + u * NT + tid;
#pragma unroll
            for (int kk = 0; kk < NV; ++kk) {
              const int k = half * NV + kk;
              while (__builtin_signbit(v[u][kk]) != want_neg) {
                if (++spins > kNormSpinLimit) { v[u][kk] = nan; break; }
                v[u][kk] = __hip_atomic_load(&partial[(size_t)k * nb + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              }
              acc[k] += fabs(v[u][kk]);
            }
          }
        }
      }
#pragma unroll
      for (int k = 0; k < 2 * NV; ++k) acc[k] = wave_sum_lane63(acc[k]);
      __syncthreads();  // red[] is reused below; every wave of this block reached here
      if (lane == 63) {
#pragma unroll
        for (int k = 0; k < 2 * NV; ++k) red[wv][k] = acc[k];
      }
      __syncthreads();
      if (tid < 2 * NV && tid < nout.count) {
        double s = 0.0;
        for (int w = 0; w < NT / 64; ++w) s += red[w][tid];
        nout.l2[tid] = sqrt(s);
      }
      // every arrival of this launch is in: reset the counters
      if (tid <= kNormGroups)
        __hip_atomic_store(nout.state + tid * kNormLineWords, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

}  // namespace sg
