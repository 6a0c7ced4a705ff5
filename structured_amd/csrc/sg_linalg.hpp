// Device linear algebra used by the Newton-Krylov driver (sg_linalg.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace sg {

constexpr int NSLOT = 13;  // Jacobian stencil blocks kept per cell

size_t multi_dot_partial_size(int kmax);
// out[i] = V_i . w for i < k (V_i = V + i*ldv); out is device memory.  With
// ga and gb (device doubles) the kernels run only if *ga < gf * *gb and
// otherwise write nothing: a conditional pass decided on the device.
hipError_t multi_dot(const double* V, size_t ldv, int k, const double* w, size_t n, double* partial, double* out,
                     bool accumulate, hipStream_t s, const double* ga = nullptr, const double* gb = nullptr,
                     double gf = 0.5);
// w -= sum_i c[i] V_i, c device memory; gated as multi_dot
hipError_t multi_axpy(const double* V, size_t ldv, int k, const double* c, double* w, size_t n, hipStream_t s,
                      const double* ga = nullptr, const double* gb = nullptr, double gf = 0.5);
// y (+)= sum_i c[i] X_i, c device memory
hipError_t lincomb(const double* X, size_t ldx, int k, const double* c_dev, double* y, size_t n, bool accumulate,
                   hipStream_t s);
hipError_t axpby(double a, const double* x, double b, double* y, size_t n, hipStream_t s);
// w = z * idt (per cell) - jz
hipError_t shift_minus(const double* z, const double* idt, const double* jz, double* w, size_t plane, int nv,
                       hipStream_t s);
// Coloured Jacobian extraction with colour period per (5: the 13-point
// stencil; 3: the 3x3 box)
hipError_t seed_colour(int nic, int njc, int ldi, int nv, int per, int a, int b, int m, double* v, hipStream_t s);
hipError_t extract_blocks(int nic, int njc, int ldi, int nv, int per, int a, int b, int m, const double* jv,
                          double* blocks, hipStream_t s);
hipError_t zero_fill(double* x, size_t n, hipStream_t s);
hipError_t line_factor(int nic, int njc, int ldi, int nv, const double* idt, const double* blocks, double* fac,
                       bool jacobi, hipStream_t s);
hipError_t line_solve(int nic, int njc, int ldi, int nv, const double* fac, const double* b, double* x, hipStream_t s);
// symmetric multicolour line Gauss-Seidel colour step (see sg_linalg.hip)
hipError_t line_gs(int nic, int njc, int ldi, int nv, int colour, const double* blocks, const double* fac,
                   const double* bvec, double* t, double* x, hipStream_t s);
// multicolour (5 colours) block ILU(0) of I/dt - J on the 13-point stencil
hipError_t ilu_factor(int nic, int njc, int ldi, int nv, const double* idt, const double* blocks, double* ilu,
                      hipStream_t s);
hipError_t ilu_apply(int nic, int njc, int ldi, int nv, const double* ilu, const double* b, double* z, hipStream_t s);
// level-scheduled block ILU(0) of I/dt - J in natural j-major order (one
// persistent work-group); apply with transpose = true solves with (LU)^T
hipError_t lilu_factor(int nic, int njc, int ldi, int nv, const double* idt, const double* blocks, double* ilu,
                       hipStream_t s);
hipError_t lilu_apply(int nic, int njc, int ldi, int nv, const double* ilu, const double* b, double* z, bool transpose,
                      hipStream_t s);
// chunked line block ILU(0) (sg_cilu.hip): chunks of w in {8, 16, 32} columns,
// on the 13-point stencil or (nine) the 3x3-box pattern; cilu_size is the
// storage in doubles that cilu_factor writes and cilu_apply reads: two FP32
// copies of the factors (plain and transposed), the apply's progress words
// (written by every apply), and the factorisation's operands gathered in
// level order
bool cilu_width_ok(int w);
size_t cilu_size(int nic, int njc, int nv, int w);
hipError_t cilu_factor(int nic, int njc, int ldi, int nv, int w, bool nine, const double* idt, const double* blocks,
                       double* rec, hipStream_t s);
hipError_t cilu_apply(int nic, int njc, int ldi, int nv, int w, bool nine, double* store, const double* b, double* z,
                      bool transpose, hipStream_t s);
// adjoint helpers
hipError_t jt_gather(int nic, int njc, int ldi, int nv, const double* blocks, const double* psi, double* out,
                     hipStream_t s);
hipError_t dbeta_apply(int nic, int njc, int ldi, const double* psi, const double* diag, double* out, hipStream_t s);
hipError_t ilu_apply_t(int nic, int njc, int ldi, int nv, const double* ilu, const double* b, double* z, hipStream_t s);
// per-block minima of the admissible update fraction (see sg_linalg.hip)
hipError_t update_limit(int nic, int njc, int ldi, const double* q, const double* dq, double f, double* block_min,
                        int nblocks, hipStream_t s);
hipError_t invert_dt(const double* dt, double* idt, int nic, int njc, int ldi, hipStream_t s);

}  // namespace sg
