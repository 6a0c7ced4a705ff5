// Plain-data descriptors shared by the host context and the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace sg {
// Arrival counters of the fused norm: one per group of residual tiles
// (kNormGroups, tile index modulo kNormGroups: the XCD the tile runs on) and
// one for the groups, each on its own 128-byte line; 64-bit words.
constexpr int kNormGroups = 8, kNormLineWords = 16;
constexpr int kNormStateWords = (kNormGroups + 1) * kNormLineWords;
}  // namespace sg

namespace sg {

// Device geometry, struct-of-arrays.  Cell arrays: a[j*ldi + i] (i < nic,
// j < njc).  Chi-face arrays (faces i = 0..ni-1 between cells i-1 and i):
// a[j*ldf + i].  Eta-face arrays (faces j = 0..nj-1): a[j*ldi + i].
struct DevGeom {
  int ni, nj, nic, njc;
  int ldi, ldf;
  size_t plane;              // njc * ldi, stride between state components
  const double* ncx;         // normal_chi x, [njc][ldf]
  const double* ncy;
  const double* nex;         // normal_eta x, [nj][ldi]
  const double* ney;
  const double* vol;         // [njc][ldi]
  const double* wd;          // wall distance (SA)
  const double* beta;        // SA production multiplier
};

struct DevPhys {
  double rho_inf, u_inf, v_inf, p_inf, T_inf, mu_inf, pr, nut_inf;
  double R, cp;              // gas constant p/(rho T) and cp = gamma R/(gamma-1)
  double iR, ipr, iTinf;     // 1/R, 1/Pr, 1/T_inf
  double eps_chi, eps_eta;   // limiter epsilons (10/nic)^3, (10/njc)^3
  double dpdx, dpdy;
  int order;                 // 1 or 2
  int viscous;               // mu_inf > 1e-15
  int flux;                  // FluxKind
  int pad_;
};

// Convective flux scheme (solver.flux).
enum FluxKind { FLUX_ROE = 0, FLUX_AUSM = 1 };

// One boundary-condition evaluation.  Operand refs: >= 0 is an interior
// cell (cj*nic + ci), < 0 is another node -(ref+1).
enum BcOp { OP_FREESTREAM = 0, OP_SLIP = 1, OP_WALL = 2, OP_ISO = 3, OP_OUTFLOW = 4, OP_COPY = 5 };
struct BcNode {
  int op, ra, rb, pad;
  double c0, c1, c2;         // (u_bc, v_bc, T_bc) or (nx, ny) of the wall face
};

// Ghost ring: perimeter index of ghosted cell (gi, gj):
//   bottom row gj=0          : gi                      (gi = 0..nic+1)
//   top row    gj=njc+1      : (nic+2) + gi
//   left col   gi=0,  gj=1..njc : 2(nic+2) + gj-1
//   right col  gi=nic+1       : 2(nic+2) + njc + gj-1
// ghost_map[perimeter] = final node id, or -1 if no boundary writes it.
struct DevBc {
  const BcNode* nodes;
  const int* ghost_map;
};

// Closure for the SA production multiplier beta:
//   CLOSURE_FIELD    beta from the per-cell field DevGeom::beta (default)
//   CLOSURE_ANALYTIC beta = 1 + amp tanh(a . f + a0), FP64, f the local features
//   CLOSURE_MLP      beta = 1 + MLP(f) on the matrix cores (BF16 operands,
//                    FP32 accumulation), W1/b1/b2/W3/b3 FP32, W2 BF16
enum ClosureKind { CLOSURE_FIELD = 0, CLOSURE_ANALYTIC = 1, CLOSURE_MLP = 2 };
struct DevClosure {
  int kind = CLOSURE_FIELD;
  double coef[7] = {0, 0, 0, 0, 0, 0, 0};  // analytic: a[5], a0, amp
  const float* params = nullptr;           // MLP packed FP32 parameters
  const void* w2 = nullptr;                // MLP W2 as BF16 [64 out][64 in]
  double* aux = nullptr;                   // optional output: features (5 planes) and beta (1 plane)
};

constexpr double kGamma = 1.4;

// SA-noft2 constants (Spalart & Allmaras 1994 as catalogued by NASA TMR),
// SA-neg branch for nu~ < 0 (c_t3 = 1.2, c_n1 = 16).
constexpr double kCb1 = 0.1355, kSigma = 2.0 / 3.0, kCb2 = 0.622, kKappa = 0.41;
constexpr double kCw2 = 0.3, kCw3 = 2.0, kCv1 = 7.1, kCv2 = 0.7, kCv3 = 0.9, kPrt = 0.9;
constexpr double kCt3 = 1.2, kCn1 = 16.0;

// Local time step of one cell, EulerEquation::calc_dt
// (src/model/eulerequation.cpp:237-259): CFL / (lambda/l + 2 mu_inf/l^2) with
// lambda = c + |u| + |v| and l the shorter face length; IEEE divisions and
// square root in the reference's order, so the value is the reference's.
__device__ __forceinline__ double cell_dt(double r, double ru, double rv, double rE, double len, double mu_inf,
                                          double cfl) {
  const double u = ru / r, v = rv / r;
  const double p = (rE - 0.5 * r * (u * u + v * v)) * (kGamma - 1.0);
  const double lam = sqrt(kGamma * p / r) + fabs(u) + fabs(v);
  return cfl / (lam / len + 2.0 * mu_inf / len / len);
}

// Wavefront sum in a fixed order with DPP lane moves (no LDS round trips):
// xor 1 and xor 2 inside quads, the mirrors inside 8 and 16 lanes, then the
// row broadcasts of lanes 15 and 31 carry the row sums up; lane 63 ends with
// the total.  Lanes a move does not write receive 0.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_sum_lane63(double v) {
  v += dpp_move<0xB1>(v);        // quad_perm [1,0,3,2]
  v += dpp_move<0x4E>(v);        // quad_perm [2,3,0,1]
  v += dpp_move<0x141>(v);       // row_half_mirror
  v += dpp_move<0x140>(v);       // row_mirror
  v += dpp_move<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_move<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}

}  // namespace sg
