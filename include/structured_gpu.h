/*
 * structured_gpu.h -- C ABI of the MI355X residual / tangent / adjoint library.
 *
 * Plain pointers and sizes only.  Host-buffer entry points take and return the
 * reference's array-of-structures layout q[i][j][k] (i over nic = ni-1 cells,
 * j over njc = nj-1 cells, k over the nvar unknowns per cell, k fastest;
 * src/solver/solution.cpp:28).  Device-buffer entry points (suffix _dev) take
 * device pointers in the library's struct-of-arrays layout:
 * a[k*plane + j*ldi + i] with ldi = sg_ldi(ctx) and plane = njc*ldi.
 *
 * Every function returns 0 (SG_OK) or a negative SG_ERR_* code; the text of
 * the last error of a context is available from sg_last_error().  One context
 * owns one HIP stream and its device buffers; a context is not thread-safe.
 * Host-buffer calls are synchronous; _dev calls are ordered on the context's
 * stream and are synchronous as well unless the name ends in _async.
 */
#ifndef STRUCTURED_GPU_H
#define STRUCTURED_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SG_OK 0
#define SG_ERR_ARG (-1)     /* bad argument or shape                         */
#define SG_ERR_BC (-2)      /* boundary list reads a ghost before writing it */
#define SG_ERR_NAN (-3)     /* non-finite residual / norm                    */
#define SG_ERR_HIP (-4)     /* HIP runtime error                             */
#define SG_ERR_NOCONV (-5)  /* linear or nonlinear solve did not converge    */
#define SG_ERR_IO (-6)      /* file missing, short or malformed             */
#define SG_ERR_STATE (-7)   /* call out of order (e.g. no state uploaded)    */

/* Boundary types and faces, src/model/bc.h:8-11 and bc.cpp:465-526. */
enum sg_bc_type {
  SG_BC_FREESTREAM = 0,
  SG_BC_OUTFLOW = 1,
  SG_BC_SLIPWALL = 2,
  SG_BC_WALL = 3,        /* adiabatic no-slip wall ("wall")            */
  SG_BC_ISOTHERMAL = 4,  /* "isothermalwall"                           */
  SG_BC_PERIODIC = 5,
  SG_BC_WAKE = 6
};
enum sg_face { SG_FACE_BOTTOM = 0, SG_FACE_RIGHT = 1, SG_FACE_TOP = 2, SG_FACE_LEFT = 3 };

/* [freestream] table, src/utils/config.cpp:31-43. */
typedef struct sg_freestream {
  double rho_inf, u_inf, v_inf, p_inf, T_inf;
  double mu_inf;   /* > 1e-15 switches the viscous terms on                 */
  double pr_inf;
  double aoa_deg;  /* used only to rotate the wall forces into Cl / Cd      */
  double nut_inf;  /* freestream nu~ for the SA model (absolute value)      */
} sg_freestream;

/* One [[boundary]] entry.  start/end are ghosted indices along the face,
 * inclusive; a negative end counts from the far end (end -> n+2+end),
 * src/model/bc.cpp:435-457. */
typedef struct sg_bc {
  int type;
  int face;
  long long start;
  long long end;
  double u, v, T;
} sg_bc;

/* Convective flux (src/model/flux.cpp): ConvectiveFluxRoe (:51-146),
 * ConvectiveFluxAUSM (:149-224). */
#define SG_FLUX_ROE 0
#define SG_FLUX_AUSM 1

typedef struct sg_opts {
  int order;       /* MUSCL order, 1 or 2 (solver.order)                    */
  int turbulence;  /* 0: laminar (4 unknowns), 1: Spalart-Allmaras (5)      */
  int tail;        /* geometry.tail: surface cells are [tail-1, ni-tail)    */
  int flux;        /* solver.flux: SG_FLUX_ROE (0) or SG_FLUX_AUSM (1); the
                      reference's config default is "ausm"
                      (src/utils/config.cpp:54, selected at
                      src/model/eulerequation.cpp:122-127)                   */
  double dpdx, dpdy;  /* [source] body force                                */
} sg_opts;

typedef struct sg_ctx sg_ctx;

/* ---- setup: Mesh::setup + EulerEquation ctor + BoundaryContainer,
 *      src/utils/mesh.cpp:391-414, src/model/eulerequation.cpp:31-133 ---- */

/* xv, yv: vertex coordinates [ni][nj] (nj fastest, as Mesh::xv). */
int sg_create(int ni, int nj, const double* xv, const double* yv, const sg_freestream* fs,
              const sg_bc* bcs, int nbc, const sg_opts* opts, sg_ctx** out);
void sg_destroy(sg_ctx* ctx);
const char* sg_last_error(const sg_ctx* ctx);
int sg_nvar(const sg_ctx* ctx);
int sg_ldi(const sg_ctx* ctx);
/* Per-cell SA production multiplier beta [nic][njc] (default 1). */
int sg_set_beta(sg_ctx* ctx, const double* beta);
/* Wall distance [nic][njc] used by the SA destruction term.  sg_create
 * computes it from the wall boundaries; this overrides it. */
int sg_set_walldist(sg_ctx* ctx, const double* d);
int sg_get_walldist(sg_ctx* ctx, double* d);
/* Metric arrays as the reference's Mesh holds them:
 * normal_chi [ni][njc][2], normal_eta [nic][nj][2], volume/xc/yc [nic][njc]. */
int sg_get_metrics(sg_ctx* ctx, double* normal_chi, double* normal_eta, double* volume, double* xc, double* yc);

/* ---- residual: EulerEquation::calc_residual, src/model/eulerequation.cpp:202-232
 *      (R = dq/dt, already divided by the cell volume), plus the per-equation
 *      L2 norm of Solver::step, src/solver/solver.cpp:125-134.  l2 may be NULL. */
int sg_residual(sg_ctx* ctx, const double* q_aos, double* r_aos, double* l2);

/* ---- initial state: EulerEquation::initialize, src/model/eulerequation.cpp:262-277.
 *      Fills q_aos [nic][njc][nvar] with the freestream conservative state
 *      (rho, rho u, rho v, p/(gamma-1) + rho |u|^2 / 2) and, with the SA
 *      model, rho nu~_inf as the fifth unknown (the ntrans hook of
 *      src/solver/solution.cpp:9). */
int sg_initialize(sg_ctx* ctx, double* q_aos);

/* ---- tangent: J.v with J = dR/dq at q (replaces the ADOL-C tape +
 *      sparse_jac forward sweep, src/solver/solver.cpp:69-90,153-158). */
int sg_jv(sg_ctx* ctx, const double* q_aos, const double* v_aos, double* jv_aos);

/* ---- adjoint: J^T.psi and (dR/dbeta)^T.psi (no counterpart in the
 *      reference, which has no reverse mode; dbeta is [nic][njc]).  J^T.psi is
 *      applied in gather form from the exact Jacobian blocks of the 13-point
 *      stencil (extracted at q by coloured tangent sweeps); dR/dbeta is
 *      diagonal (beta_c enters only the SA production of cell c) and comes
 *      from one tangent sweep in the beta direction. */
int sg_jtv(sg_ctx* ctx, const double* q_aos, const double* psi_aos, double* jtpsi_aos, double* dbeta);

typedef struct sg_adjoint_opts {
  int restart;     /* GMRES restart length                                  */
  int maxit;       /* total Krylov iterations                               */
  double rtol;     /* relative residual tolerance                           */
  int precond;     /* 0 none, 4 five-colour ILU(0), 5 natural-order ILU(0),
                      6 chunked line ILU(0) on the 13-point stencil, 7 the
                      same on the 3x3-box pattern                           */
  int lhs_order;   /* 1: the preconditioner factors the first-order Jacobian */
  int pc_width;    /* chunk width of precond 6 / 7 (8, 16, 24 or 32); 0 keeps the
                      context's current width                               */
  double pc_cfl;   /* the preconditioner factors I/dt - J at this CFL        */
} sg_adjoint_opts;
typedef struct sg_adjoint_report {
  int iterations;
  int reorthogonalised;  /* Arnoldi steps that took the second Gram-Schmidt
                            pass (selective re-orthogonalisation)           */
  double relres;
  double wall_s;
} sg_adjoint_report;

/* Solve J^T psi = rhs at q (all AoS). */
int sg_adjoint_solve(sg_ctx* ctx, const double* q_aos, const double* rhs_aos, double* psi_aos,
                     const sg_adjoint_opts* opts, sg_adjoint_report* report);
/* Wall objective F = sum_i ds_i (Cf_i - Cf*_i)^2 over the surface cells of
 * sg_wall (cf_target has nb entries) and its exact gradient dF/dq (AoS, may
 * be NULL). */
int sg_objective_cf(sg_ctx* ctx, const double* q_aos, const double* cf_target, double* F, double* dFdq_aos);
/* dF/dbeta [nic][njc] at a converged state q via one adjoint solve. */
int sg_adjoint_gradient(sg_ctx* ctx, const double* q_aos, const double* cf_target, double* grad_beta, double* F,
                        const sg_adjoint_opts* opts, sg_adjoint_report* report);

/* ---- local time step: EulerEquation::calc_dt, src/model/eulerequation.cpp:237-259 */
int sg_dt(sg_ctx* ctx, const double* q_aos, double cfl, double* dt_aos);

/* ---- wall quantities: IOManager::write_surface, src/utils/io.cpp:181-255.
 *      x, cp, cf have nb = ni-2*tail+1 entries; forces = {cl_p, cd_p, cl_v,
 *      cd_v, cl, cd}.  Evaluates the residual at q internally, because the
 *      wall Cf uses the eta-face gradients of that evaluation. */
int sg_wall(sg_ctx* ctx, const double* q_aos, double* x, double* cp, double* cf, double* forces);

/* ---- files: the Mesh grid loaders (src/utils/mesh.cpp:133-169) and the
 *      IOManager writers (src/utils/io.cpp:41-255), same layouts and text
 *      formatting as the reference ---- */
/* format "p3d" (plot3d_loader: "1 ni nj", then x, then y, j-outer i-inner) or
 * "simple" (simple_loader: x y pairs, j-outer i-inner).  xv, yv: [ni][nj].
 * A header that does not match ni, nj or a short file is SG_ERR_ARG (the
 * reference asserts); errors read back through sg_last_error(NULL). */
int sg_read_grid(const char* path, const char* format, int ni, int nj, double* xv, double* yv);
/* <label>.out: raw float64 q[i][j][k] (write_restart, io.cpp:134-152) */
int sg_write_restart(sg_ctx* ctx, const char* path, const double* q_aos);
/* read_restart (io.cpp:154-179): a file of any other size than
 * nic*njc*nvar doubles is SG_ERR_ARG (the reference asserts) */
int sg_read_restart(sg_ctx* ctx, const char* path, double* q_aos);
/* <label>.surface: "x cp cf" per wall cell (write_surface, io.cpp:181-255),
 * from sg_wall at q; forces (optional, 6 entries) as sg_wall */
int sg_write_surface(sg_ctx* ctx, const char* path, const double* q_aos, double* forces);
/* <label>.tec: Tecplot point file of xc, yc, q (+ psi_0 for SA) and the ten
 * q_aux columns (write_tecplot, io.cpp:49-99; the reference never fills
 * q_aux, written here as 0) */
int sg_write_tecplot(sg_ctx* ctx, const char* path, const double* q_aos);

/* ---- device-resident variants (SoA device pointers, see header comment) */
int sg_residual_dev(sg_ctx* ctx, const double* q_dev, double* r_dev, double* l2_host);
int sg_jv_dev(sg_ctx* ctx, const double* q_dev, const double* v_dev, double* jv_dev);
/* sg_jv_dev without synchronising (enqueued on the context's stream). */
int sg_jv_dev_async(sg_ctx* ctx, const double* q_dev, const double* v_dev, double* jv_dev);
int sg_jtv_dev(sg_ctx* ctx, const double* q_dev, const double* psi_dev, double* jtpsi_dev, double* dbeta_dev);
/* Enqueue the residual (and, if l2_dev is not NULL, its per-equation L2 norm
 * ||R_k||_2 into l2_dev[0..nvar-1], device memory: exactly nvar doubles are
 * written) without synchronising. */
int sg_residual_dev_async(sg_ctx* ctx, const double* q_dev, double* r_dev, double* l2_dev);
/* sg_residual_dev_async without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops).  No
 * counterpart in the reference.  The norm scratch (per-tile partial sums and
 * an arrival counter) lives in a static device array of each GPU, at most
 * SG_NOALLOC_MAX_TILES residual tiles of 31 x 7 cells (1.25 MiB, held until
 * the process ends); a larger grid is SG_ERR_ARG.  Calls that overlap on one
 * GPU share that array and race: serialise them. */
#define SG_NOALLOC_MAX_TILES 16384
int sg_residual_dev_async_noalloc(sg_ctx* ctx, const double* q_dev, double* r_dev, double* l2_dev);
int sg_synchronize(sg_ctx* ctx);
/* The context's HIP stream (a hipStream_t), so callers can order their own
 * work and events with the library's. */
void* sg_get_stream(sg_ctx* ctx);
/* ---- linear solver: the reference's interface (preallocate / set_lhs /
 *      set_rhs / solve / solve_and_update, src/linearsolver/ls_eigen.h:25-31),
 *      matrix-free.  set_lhs takes the linearisation state q and the local
 *      time steps dt (as sg_dt returns them) and sets A = I/dt - J(q); it
 *      also builds the preconditioner from the exact Jacobian's 13 stencil
 *      blocks per cell (precond 0: none, 1: block Jacobi, 2: j-line
 *      block-tridiagonal, 3: symmetric multicolour j-line Gauss-Seidel,
 *      4: five-colour block ILU(0) of I/dt - J on the 13-point stencil,
 *      5: block ILU(0) in natural j-major order, level-scheduled,
 *      6: chunked line block ILU(0): chunks of pc_width columns, each ordered
 *      column by column, couplings between chunks dropped, FP32 factors,
 *      7: precond 6 on the 3x3-box pattern: the +-2 blocks and their fill
 *      dropped, which with lhs_order 1 are fill only).
 *      The Krylov method is restarted FGMRES (restart length `restart`,
 *      at most `maxit` iterations, relative tolerance `rtol`). */
int sg_ls_preallocate(sg_ctx* ctx, int restart, int precond, int maxit, double rtol);
int sg_ls_set_lhs(sg_ctx* ctx, const double* q_aos, const double* dt_aos);
/* Preconditioner settings for the following set_lhs (after preallocate):
 * lhs_order as in sg_newton_opts, the factor (>= 1) by which the
 * preconditioner's 1/dt is larger than the operator's, and the chunk width of
 * precond 6 / 7 (8, 16, 24 or 32; 0 keeps the current one). */
int sg_ls_set_pc(sg_ctx* ctx, int lhs_order, double pc_dt_factor, int pc_width);
int sg_ls_set_rhs(sg_ctx* ctx, const double* rhs_aos);
int sg_ls_solve(sg_ctx* ctx, double* dq_aos, int* iters, double* relres);
int sg_ls_solve_and_update(sg_ctx* ctx, double* q_aos, double under_relaxation);
/* z = M^-1 v (transpose = 0) or M^-T v (transpose = 1, preconditioners 0,
 * 4, 5, 6 and 7) with the preconditioner built by the last set_lhs. */
int sg_ls_precond_apply(sg_ctx* ctx, const double* v_aos, double* z_aos, int transpose);
/* sg_ls_precond_apply on device vectors in the library's SoA layout (as
 * sg_residual_dev), enqueued on the context's stream without synchronising;
 * v_dev and z_dev must not overlap.  The preconditioner application inside
 * LinearSolver::solve (src/solver/linearsolver.cpp, the PETSc PC), exposed
 * for callers that keep their Krylov vectors on the device. */
int sg_ls_precond_apply_dev_async(sg_ctx* ctx, const double* v_dev, double* z_dev, int transpose);

/* ---- nonlinear solve: Solver::solve / Solver::step, src/solver/solver.cpp:52-236.
 *      Ramp semantics as the reference (after step `counter`, if
 *      counter > ramp_iteration: cfl = min(cfl^exponent, cfl_max), UR =
 *      min(UR^exponent, 10), solver.cpp:200-230); ser > 0 instead sets cfl =
 *      cfl0 * (|R0|/|R|)^ser_exponent.  Stops when every equation's L2 norm
 *      is below tolerance (SG_OK), or, like Solver::step (solver.cpp:136),
 *      at the residual evaluation whose counter exceeds iteration_max, before
 *      its linear solve: counters 0..iteration_max each solve, so
 *      iteration_max + 1 linear solves (SG_ERR_NOCONV).  report->iterations
 *      is the number of linear solves (Newton updates) made. */
typedef struct sg_newton_opts {
  double cfl;
  int cfl_ramp;
  int cfl_ramp_iteration;
  double cfl_ramp_exponent;
  double cfl_max;
  double under_relaxation;
  int under_relaxation_ramp;
  int under_relaxation_ramp_iteration;
  double under_relaxation_ramp_exponent;
  int iteration_max;
  double tolerance;
  int gmres_restart;
  int gmres_maxit;
  double gmres_rtol;
  int precond;
  int ser;
  double ser_exponent;
  int verbose;
  /* > 0: the preconditioner factors I/dt_pc - J with dt_pc taken at
   * min(cfl, pc_cfl_max) (a diagonal boost that keeps the incomplete
   * factorisation stable at large CFL); the Krylov operator keeps dt. */
  double pc_cfl_max;
  /* 0: stop on the reference's norm ||R_k||_2 (R = flux balance / V); 1: on
   * ||V R_k||_2, whose roundoff floor does not grow as cells shrink. */
  int norm_weighted;
  /* > 0: also stop when the stopping norm, already within 1000x of the
   * tolerance, has not fallen below half its best earlier value during the
   * last stall_window steps (a floor set by rounding or by a
   * piecewise-constant closure); the report then has stalled = 1. */
  int stall_window;
  /* solver.lhs_order (src/utils/config.cpp): 0 = the residual's order; 1 =
   * the preconditioner factors the Jacobian of the first-order residual
   * (src/solver/solver.cpp:92-101); the Krylov operator stays exact. */
  int lhs_order;
  int pc_width;  /* chunk width of precond 6 / 7: 8, 16, 24 or 32 (0: 16) */
} sg_newton_opts;

typedef struct sg_solve_report {
  int iterations;
  int converged;
  int linear_iterations;
  int stalled;
  double l2_initial[5];
  double l2_final[5];
  double l2w_final[5];   /* ||V R_k||_2 at the end */
  double cfl_final;
  double wall_s, residual_s, jacobian_s, linsolve_s;  /* Profiler buckets, src/utils/config.cpp:5-29 */
} sg_solve_report;

int sg_solve(sg_ctx* ctx, double* q_aos, const sg_newton_opts* opts, sg_solve_report* report);
/* Per-iteration L2 norms (nvar per entry) and CFL of the last sg_solve, one
 * entry per residual evaluation (report->iterations + 1 of them); returns the
 * number of entries recorded. */
int sg_solve_history(sg_ctx* ctx, double* l2_hist, double* cfl_hist, int max_entries);
/* The CFL and the under-relaxation factor in force at each residual
 * evaluation of the last sg_solve (the CFL, UNDER_RELAXATION pair that
 * Solver::step uses for that counter, src/solver/solver.cpp:63,173); returns
 * the number of entries recorded. */
int sg_solve_ramp_history(sg_ctx* ctx, double* cfl_hist, double* ur_hist, int max_entries);

/* ---- explicit pseudo-time stepping: the build of Solver::step without
 *      ADOL-C, src/solver/solver.cpp:103-121, with the update kernels
 *      update_forward_euler / update_rk4 (:12-26) ---- */
#define SG_SCHEME_FORWARD_EULER 0   /* solver.scheme = "forward_euler" */
#define SG_SCHEME_RK4_JAMESON 1     /* solver.scheme = "rk4_jameson"   */

/* One step on a device state (SoA, as sg_residual_dev): dt = calc_dt(q, cfl)
 * per cell (eulerequation.cpp:237-259), then
 *   forward Euler:  q <- q + R(q) dt;
 *   RK4-Jameson:    q_s <- q + R(q_{s-1}) dt / (4 - s), s = 0..3, q_{-1} = q,
 *                   and q <- q_3.
 * l2 (host, nvar entries, optional): ||R_k||_2 of the last stage's residual,
 * the norm Solver::step reports (solver.cpp:125-134).  SG_ERR_NAN when that
 * norm is not finite (q is updated all the same, as in the reference). */
int sg_step_explicit_dev(sg_ctx* ctx, double* q_dev, int scheme, double cfl, double* l2);

typedef struct sg_explicit_opts {
  int scheme;               /* SG_SCHEME_*                                      */
  int iteration_max;        /* solver.iteration_max                             */
  double cfl;               /* solver.cfl                                       */
  int cfl_ramp;             /* solver.cfl_ramp                                  */
  int cfl_ramp_iteration;   /* solver.cfl_ramp_iteration                        */
  double cfl_ramp_exponent; /* solver.cfl_ramp_exponent                         */
  double tolerance;         /* 0 (the reference's test is disabled,
                               solver.cpp:145); > 0 stops once every ||R_k||_2
                               is below it                                    */
  int verbose;
  int reserved;
} sg_explicit_opts;

/* Solver::solve with an explicit scheme (solver.cpp:197-236): steps
 * counter = 0, 1, ... each update q, and the loop ends after the step whose
 * counter exceeds iteration_max (iteration_max + 2 updates), the CFL ramping
 * as CFL <- min(CFL^e, 1e12) once counter > cfl_ramp_iteration.  q_aos: host
 * state, updated in place.  l2_hist (optional): nvar norms per step, room for
 * iteration_max + 2 steps.  report->iterations is the number of steps. */
int sg_solve_explicit(sg_ctx* ctx, double* q_aos, const sg_explicit_opts* opts, sg_solve_report* report,
                      double* l2_hist);

/* ---- field inversion (no counterpart in the reference): L-BFGS on
 *      J(beta) = F(q(beta)) + reg * sum (beta - 1)^2 with projected Armijo
 *      backtracking; each trial re-converges the primal by sg_solve from the
 *      last accepted state, each accepted point costs one adjoint gradient. */
typedef struct sg_inversion_opts {
  int iterations;         /* L-BFGS iterations                                */
  int memory;             /* stored correction pairs                          */
  double reg;             /* weight of the prior term sum (beta - 1)^2        */
  double step0;           /* max |delta beta| of the first (steepest) step    */
  double beta_min, beta_max;  /* box the iterates are projected onto          */
  int max_backtracks;     /* step halvings (factor 4) per line search         */
  int reserved;
  sg_newton_opts newton;  /* primal re-solves                                 */
  sg_adjoint_opts adjoint;
} sg_inversion_opts;
typedef struct sg_inversion_report {
  int iterations, evaluations, failed_solves;
  int adjoint_unconverged;  /* adjoint solves that stopped above rtol (their
                               gradient is used all the same)                */
  double J0, J, F0, F;
  double wall_s, primal_s, adjoint_s;
  double adjoint_relres_max;  /* largest final relres of the adjoint solves  */
  int adjoint_iterations;     /* Krylov iterations of all adjoint solves     */
  int steepest_restarts;      /* line searches retried along -g after the
                                 L-BFGS direction found no Armijo point      */
} sg_inversion_report;
/* q_aos: converged state at the initial beta (in), state at the final beta
 * (out); beta [nic][njc] in/out; J_hist (iterations + 1 entries, may be NULL)
 * receives J at every accepted iterate.  A primal re-solve that stops on
 * opts->newton.stall_window (at its roundoff floor, within 1000x of the
 * tolerance) counts as converged. */
int sg_invert(sg_ctx* ctx, double* q_aos, const double* cf_target, double* beta, const sg_inversion_opts* opts,
              sg_inversion_report* report, double* J_hist);

/* Layout conversion between the reference AoS and the library SoA. */
int sg_aos_to_soa_dev(sg_ctx* ctx, const double* aos_dev, double* soa_dev, int nvar);
int sg_soa_to_aos_dev(sg_ctx* ctx, const double* soa_dev, double* aos_dev, int nvar);

/* ---- closures for the SA production multiplier beta (no counterpart in the
 *      reference, whose SA source has no multiplier).  The closure is
 *      evaluated inside every residual and J.v launch of the context, so the
 *      Newton solve, its Jacobian blocks and J.v all see it.
 *
 * Local features f (five per cell, O(1)):
 *   f0 = |W| / (|W| + |S|), f1 = log(1 + d^2 |W| / nu) / 10,
 *   f2 = asinh(nu~ / nu) / 8, f3 = tanh(d (u . grad p) / (|u| q)),
 *   f4 = P / (|P| + |D| + eps)
 * (W vorticity, S strain rate, d wall distance, P and D the SA production
 * with beta = 1 and destruction). */
#define SG_CLOSURE_FIELD 0     /* beta from sg_set_beta (default)             */
#define SG_CLOSURE_ANALYTIC 1  /* beta = 1 + c6 tanh(c0 f0 + ... + c4 f4 + c5) */
#define SG_CLOSURE_MLP 2       /* beta = 1 + MLP(f), BF16 matrix cores         */
#define SG_MLP_NPARAM 4609     /* W1[64][5] b1[64] W2[64][64] b2[64] W3[64] b3 */
/* coef: 7 doubles for SG_CLOSURE_ANALYTIC; mlp_params: SG_MLP_NPARAM floats
 * (host) for SG_CLOSURE_MLP; unused arguments may be NULL. */
int sg_set_closure(sg_ctx* ctx, int kind, const double* coef, const float* mlp_params);
/* The same as sg_set_closure(ctx, SG_CLOSURE_MLP, NULL, params): loads the
 * network's SG_MLP_NPARAM weights (host floats) and selects the MLP closure. */
int sg_set_mlp_weights(sg_ctx* ctx, const float* params);
/* Features [nic][njc][5] and the beta the closure gives [nic][njc] at q
 * (either may be NULL).  With SG_CLOSURE_FIELD, beta is the field. */
int sg_closure_eval(sg_ctx* ctx, const double* q_aos, double* features, double* beta);

/* ---- closure network training (no counterpart in the reference).  A
 *      5 -> 64 -> 64 -> 1 tanh MLP; the 64 x 64 products of the forward and
 *      backward passes run on the matrix cores with BF16 operands and FP32
 *      accumulation; first/last layers, loss (mean squared error), gradient
 *      reductions and Adam (beta1 0.9, beta2 0.999, eps 1e-8) are FP32 and
 *      deterministic. */
typedef struct sg_mlp sg_mlp;
/* Glorot-uniform weights (seeded), zero biases; max_rows bounds the rows of
 * one forward/backward pass (minibatch size). */
int sg_mlp_create(sg_mlp** net, int max_rows, unsigned long long seed);
void sg_mlp_destroy(sg_mlp* net);
const char* sg_mlp_last_error(const sg_mlp* net);
int sg_mlp_set_params(sg_mlp* net, const float* params);
int sg_mlp_get_params(sg_mlp* net, float* params);
/* y[n] = MLP(x[n][5]) (host arrays; any n). */
int sg_mlp_forward(sg_mlp* net, int n, const float* x, float* y);
/* Mean squared error over x[n][5] -> t[n] (n <= max_rows) and its gradient
 * with respect to the SG_MLP_NPARAM parameters. */
int sg_mlp_loss_grad(sg_mlp* net, int n, const float* x, const float* t, float* loss, float* grad);
/* epochs of Adam over n samples (host arrays, copied to the device once) in
 * minibatches of `batch` rows drawn by a seeded shuffle per epoch (the last
 * partial minibatch is dropped when n > batch); loss_hist[epochs] receives
 * the mean minibatch loss of each epoch (may be NULL). */
int sg_mlp_train(sg_mlp* net, int n, const float* x, const float* t, int epochs, int batch, float lr,
                 unsigned long long seed, float* loss_hist);
/* Closure of a context from the current parameters of a network. */
int sg_set_closure_mlp(sg_ctx* ctx, const sg_mlp* net);

#ifdef __cplusplus
}
#endif
#endif /* STRUCTURED_GPU_H */
