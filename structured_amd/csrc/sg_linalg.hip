// Device linear algebra for the matrix-free Newton-Krylov solve:
// deterministic multi-dot products and multi-axpys for the Gram-Schmidt
// passes (with device-side gates for the selective second pass), scaling,
// colour seeding and extraction of Jacobian blocks from J.v products, the
// block preconditioners other than the chunked line ILU (j-line
// block-tridiagonal, line Gauss-Seidel, multicolour and natural-order block
// ILU(0)), and the adjoint products J^T psi and (dR/dbeta)^T psi from the
// blocks.  All vectors are SoA with zero padding (length nvar*plane).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sg_device.hpp"
#include "sg_linalg.hpp"

namespace sg {

constexpr int DOT_BLOCK = 256;

// Gate of the conditional kernels: with ga and gb (device) the kernel runs
// only when *ga < gf * *gb, else it returns at once (nothing written).
__device__ __forceinline__ bool gate_open(const double* ga, const double* gb, double gf) {
  return !ga || *ga < gf * *gb;
}

// out_partial[b*k + i] = sum over slice b of V_i . w.  The grid is
// DOT_SLICES x (groups of DOT_KC vectors): block (b, g) sums vectors
// g*DOT_KC .. g*DOT_KC + DOT_KC - 1 over the element pairs of slice b (pairs
// b*DOT_BLOCK + t + m*DOT_SLICES*DOT_BLOCK, read as 16-byte loads; an odd last
// element goes to slice 0, thread 0), with its accumulators in registers,
// then over the block in a fixed order (DPP wavefront sums, then the four
// waves).  All groups run at once, so a launch keeps every vector's loads in
// flight together; each group reads w once more, from L2 / MALL after the
// first.  Vectors start on 16-byte boundaries (the host checks).
constexpr int DOT_KC = 8;
constexpr int DOT_SLICES = 512;  // partial sums per dot product
__global__ __launch_bounds__(DOT_BLOCK) void multi_dot_kernel(const double* V, size_t ldv, int k, const double* w,
                                                              size_t n, double* partial, const double* ga,
                                                              const double* gb, double gf) {
  if (!gate_open(ga, gb, gf)) return;
  __shared__ double red[DOT_BLOCK / 64][DOT_KC];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i0 = (int)blockIdx.y * DOT_KC, kk = min(DOT_KC, k - i0);
  double acc[DOT_KC];
#pragma unroll
  for (int i = 0; i < DOT_KC; ++i) acc[i] = 0.0;
  const size_t n2 = n / 2;
  const double2* w2 = reinterpret_cast<const double2*>(w);
  for (size_t e = (size_t)blockIdx.x * DOT_BLOCK + threadIdx.x; e < n2; e += (size_t)DOT_SLICES * DOT_BLOCK) {
    const double2 we = w2[e];
#pragma unroll
    for (int i = 0; i < DOT_KC; ++i)
      if (i < kk) {
        const double2 v = reinterpret_cast<const double2*>(V + (size_t)(i0 + i) * ldv)[e];
        acc[i] += v.x * we.x;
        acc[i] += v.y * we.y;
      }
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < DOT_KC; ++i)
      if (i < kk) acc[i] += V[(size_t)(i0 + i) * ldv + n - 1] * w[n - 1];
  }
#pragma unroll
  for (int i = 0; i < DOT_KC; ++i) acc[i] = wave_sum_lane63(acc[i]);
  if (lane == 63) {
#pragma unroll
    for (int i = 0; i < DOT_KC; ++i) red[wv][i] = acc[i];
  }
  __syncthreads();
  if ((int)threadIdx.x < kk) {
    double t = 0.0;
    for (int q = 0; q < DOT_BLOCK / 64; ++q) t += red[q][threadIdx.x];
    partial[(size_t)blockIdx.x * k + i0 + threadIdx.x] = t;
  }
}

// out[i] = sum_b partial[b*k + i]  (+ optional add into out when accumulate)
__global__ __launch_bounds__(256) void dot_finalize_kernel(const double* partial, int nb, int k, double* out,
                                                           int accumulate, const double* ga, const double* gb,
                                                           double gf) {
  if (!gate_open(ga, gb, gf)) return;
  const int i = blockIdx.x;
  __shared__ double red[4];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) s += partial[(size_t)b * k + i];
  s = wave_sum_lane63(s);
  if ((threadIdx.x & 63) == 63) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = red[0] + red[1] + red[2] + red[3];
    out[i] = accumulate ? out[i] + t : t;
  }
}

// the kernels read vectors as 16-byte pairs: every vector (V_i = V + i*ldv and w)
// must start on a 16-byte boundary
static bool pairs_aligned(const double* V, size_t ldv, int k, const double* w) {
  const uintptr_t a = (uintptr_t)V | (uintptr_t)w;
  return (a & 15) == 0 && (k <= 1 || (ldv & 1) == 0);
}

hipError_t multi_dot(const double* V, size_t ldv, int k, const double* w, size_t n, double* partial, double* out,
                     bool accumulate, hipStream_t s, const double* ga, const double* gb, double gf) {
  if (!pairs_aligned(V, ldv, k, w)) return hipErrorInvalidValue;
  const dim3 grid(DOT_SLICES, (unsigned)((k + DOT_KC - 1) / DOT_KC));
  hipLaunchKernelGGL(multi_dot_kernel, grid, dim3(DOT_BLOCK), 0, s, V, ldv, k, w, n, partial, ga, gb, gf);
  hipLaunchKernelGGL(dot_finalize_kernel, dim3(k), dim3(256), 0, s, partial, DOT_SLICES, k, out, accumulate ? 1 : 0,
                     ga, gb, gf);
  return hipGetLastError();
}

size_t multi_dot_partial_size(int kmax) { return (size_t)DOT_SLICES * kmax; }

// w -= sum_i c[i] V_i   (c in device memory).  Element pairs (16-byte
// loads), the vectors taken eight at a time with their loads issued together
// ahead of the eight updates; the updates keep the order i = 0, 1, ..., so
// every element gets the arithmetic of a plain loop over i.
constexpr int AXPY_KC = 8;
__global__ void multi_axpy_kernel(const double* V, size_t ldv, int k, const double* c, double* w, size_t n,
                                  const double* ga, const double* gb, double gf) {
  if (!gate_open(ga, gb, gf)) return;
  const size_t n2 = n / 2;
  double2* w2 = reinterpret_cast<double2*>(w);
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n2; e += (size_t)gridDim.x * blockDim.x) {
    double2 s = w2[e];
    int i = 0;
    for (; i + AXPY_KC <= k; i += AXPY_KC) {
      double2 v[AXPY_KC];
#pragma unroll
      for (int u = 0; u < AXPY_KC; ++u) v[u] = reinterpret_cast<const double2*>(V + (size_t)(i + u) * ldv)[e];
#pragma unroll
      for (int u = 0; u < AXPY_KC; ++u) {
        s.x -= c[i + u] * v[u].x;
        s.y -= c[i + u] * v[u].y;
      }
    }
    for (; i < k; ++i) {
      const double2 v = reinterpret_cast<const double2*>(V + (size_t)i * ldv)[e];
      s.x -= c[i] * v.x;
      s.y -= c[i] * v.y;
    }
    w2[e] = s;
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    double t = w[n - 1];
    for (int i = 0; i < k; ++i) t -= c[i] * V[(size_t)i * ldv + n - 1];
    w[n - 1] = t;
  }
}

hipError_t multi_axpy(const double* V, size_t ldv, int k, const double* c, double* w, size_t n, hipStream_t s,
                      const double* ga, const double* gb, double gf) {
  if (!pairs_aligned(V, ldv, k, w)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(multi_axpy_kernel, dim3(2048), dim3(256), 0, s, V, ldv, k, c, w, n, ga, gb, gf);
  return hipGetLastError();
}

// y = sum_i c[i] X_i  (c in device memory), in pairs and groups as multi_axpy
__global__ void lincomb_kernel(const double* X, size_t ldx, int k, const double* c, double* y, size_t n, int accumulate) {
  const size_t n2 = n / 2;
  double2* y2 = reinterpret_cast<double2*>(y);
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n2; e += (size_t)gridDim.x * blockDim.x) {
    double2 s = accumulate ? y2[e] : make_double2(0.0, 0.0);
    int i = 0;
    for (; i + AXPY_KC <= k; i += AXPY_KC) {
      double2 v[AXPY_KC];
#pragma unroll
      for (int u = 0; u < AXPY_KC; ++u) v[u] = reinterpret_cast<const double2*>(X + (size_t)(i + u) * ldx)[e];
#pragma unroll
      for (int u = 0; u < AXPY_KC; ++u) {
        s.x += c[i + u] * v[u].x;
        s.y += c[i + u] * v[u].y;
      }
    }
    for (; i < k; ++i) {
      const double2 v = reinterpret_cast<const double2*>(X + (size_t)i * ldx)[e];
      s.x += c[i] * v.x;
      s.y += c[i] * v.y;
    }
    y2[e] = s;
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    double t = accumulate ? y[n - 1] : 0.0;
    for (int i = 0; i < k; ++i) t += c[i] * X[(size_t)i * ldx + n - 1];
    y[n - 1] = t;
  }
}

hipError_t lincomb(const double* X, size_t ldx, int k, const double* c_dev, double* y, size_t n, bool accumulate,
                   hipStream_t s) {
  if (!pairs_aligned(X, ldx, k, y)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(lincomb_kernel, dim3(1024), dim3(256), 0, s, X, ldx, k, c_dev, y, n, accumulate ? 1 : 0);
  return hipGetLastError();
}

// y = a*x + b*y
__global__ void axpby_kernel(double a, const double* x, double b, double* y, size_t n) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    y[e] = a * x[e] + b * y[e];
}
hipError_t axpby(double a, const double* x, double b, double* y, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(axpby_kernel, dim3(1024), dim3(256), 0, s, a, x, b, y, n);
  return hipGetLastError();
}

// y = x * scale[k*plane...] style: w = z * idt - jz  (operator I/dt - J)
__global__ void shift_minus_kernel(const double* z, const double* idt, const double* jz, double* w, size_t plane, int nv) {
  const size_t n = plane * nv;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    w[e] = z[e] * idt[e % plane] - jz[e];
}
hipError_t shift_minus(const double* z, const double* idt, const double* jz, double* w, size_t plane, int nv, hipStream_t s) {
  hipLaunchKernelGGL(shift_minus_kernel, dim3(1024), dim3(256), 0, s, z, idt, jz, w, plane, nv);
  return hipGetLastError();
}

// v = e_m on the cells of colour (a, b) = (i mod per, j mod per), 0
// elsewhere.  per = 5 separates the 13-point stencil (reach 2), per = 3 the
// 3x3 box (reach 1).
__global__ void seed_kernel(int nic, int njc, int ldi, int nv, int per, int a, int b, int m, double* v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= nic) return;
  const size_t plane = (size_t)njc * ldi, c = (size_t)j * ldi + i;
  const bool on = (i % per == a) && (j % per == b);
  for (int k = 0; k < nv; ++k) v[k * plane + c] = (on && k == m) ? 1.0 : 0.0;
}
hipError_t seed_colour(int nic, int njc, int ldi, int nv, int per, int a, int b, int m, double* v, hipStream_t s) {
  if (per != 3 && per != 5) return hipErrorInvalidValue;
  hipLaunchKernelGGL(seed_kernel, dim3((nic + 127) / 128, njc), dim3(128), 0, s, nic, njc, ldi, nv, per, a, b, m, v);
  return hipGetLastError();
}

// Stencil slots of the exact Jacobian of the residual: the 3x3 box plus
// (0,+-2) and (+-2,0) (MUSCL reach).  Slot s holds J[r, r + (di,dj)].
__constant__ int kSlotDi[NSLOT] = {0, 0, 0, -1, 1, -1, 1, -1, 1, 0, 0, -2, 2};
__constant__ int kSlotDj[NSLOT] = {0, -1, 1, 0, 0, -1, -1, 1, 1, -2, 2, 0, 0};

__device__ __forceinline__ int slot_of(int di, int dj) {
  for (int s = 0; s < NSLOT; ++s)
    if (kSlotDi[s] == di && kSlotDj[s] == dj) return s;
  return -1;
}

// From J.v of colour (a,b), variable m: row cell r holds column m of the
// block J[r, c] for the unique colour-(a,b) cell c in r's per x per window.
// Blocks are struct-of-arrays: blocks[((slot*nv + row)*nv + col)*plane + cell].
__global__ void extract_kernel(int nic, int njc, int ldi, int nv, int per, int a, int b, int m, const double* jv,
                               double* blocks) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= nic) return;
  const int h = per / 2;
  int di = ((a - i) % per + per) % per;
  if (di > h) di -= per;
  int dj = ((b - j) % per + per) % per;
  if (dj > h) dj -= per;
  const int slot = slot_of(di, dj);
  if (slot < 0) return;
  const size_t plane = (size_t)njc * ldi, c = (size_t)j * ldi + i;
  for (int k = 0; k < nv; ++k) blocks[((size_t)(slot * nv + k) * nv + m) * plane + c] = jv[k * plane + c];
}
hipError_t extract_blocks(int nic, int njc, int ldi, int nv, int per, int a, int b, int m, const double* jv,
                          double* blocks, hipStream_t s) {
  if (per != 3 && per != 5) return hipErrorInvalidValue;
  hipLaunchKernelGGL(extract_kernel, dim3((nic + 127) / 128, njc), dim3(128), 0, s, nic, njc, ldi, nv, per, a, b, m,
                     jv, blocks);
  return hipGetLastError();
}

// Zeroes n doubles (a kernel rather than a memset, so that the call may be
// captured into a graph).
__global__ void zero_kernel(double* x, size_t n) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    x[e] = 0.0;
}
hipError_t zero_fill(double* x, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(zero_kernel, dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)), dim3(256), 0, s, x, n);
  return hipGetLastError();
}

// ---- small dense helpers (row-major nv x nv, nv <= 5) ----
template <int NV>
__device__ void inv_small(const double* A, double* Ainv) {
  double M[NV][2 * NV];
  for (int r = 0; r < NV; ++r)
    for (int c = 0; c < NV; ++c) { M[r][c] = A[r * NV + c]; M[r][NV + c] = (r == c) ? 1.0 : 0.0; }
  for (int p = 0; p < NV; ++p) {
    int piv = p;
    double best = fabs(M[p][p]);
    for (int r = p + 1; r < NV; ++r)
      if (fabs(M[r][p]) > best) { best = fabs(M[r][p]); piv = r; }
    if (piv != p)
      for (int c = 0; c < 2 * NV; ++c) { double t = M[p][c]; M[p][c] = M[piv][c]; M[piv][c] = t; }
    const double ip = 1.0 / M[p][p];
    for (int c = 0; c < 2 * NV; ++c) M[p][c] *= ip;
    for (int r = 0; r < NV; ++r) {
      if (r == p) continue;
      const double f = M[r][p];
      if (f != 0.0)
        for (int c = 0; c < 2 * NV; ++c) M[r][c] -= f * M[p][c];
    }
  }
  for (int r = 0; r < NV; ++r)
    for (int c = 0; c < NV; ++c) Ainv[r * NV + c] = M[r][NV + c];
}

// Block-tridiagonal factor of each i-column of A = I/dt - J restricted to
// the (0,-1), (0,0), (0,+1) blocks (jacobi = 0) or its block diagonal only
// (jacobi = 1).  fac (SoA, 3 x NV*NV per cell): Dinv_j, C_j = Dinv_j U_j, L_j.
template <int NV>
__global__ void line_factor_kernel(int nic, int njc, int ldi, const double* idt, const double* blocks, double* fac,
                                   int jacobi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nic) return;
  const size_t plane = (size_t)njc * ldi;
  constexpr int B = NV * NV;
  double D[B], Dinv[B], Cprev[B], L[B];
  for (int j = 0; j < njc; ++j) {
    const size_t c = (size_t)j * ldi + i;
    for (int e = 0; e < B; ++e) D[e] = -blocks[(size_t)(0 * B + e) * plane + c];
    for (int k = 0; k < NV; ++k) D[k * NV + k] += idt[c];
    if (!jacobi && j > 0) {
      for (int e = 0; e < B; ++e) L[e] = -blocks[(size_t)(1 * B + e) * plane + c];
      for (int r = 0; r < NV; ++r)
        for (int cc = 0; cc < NV; ++cc) {
          double s = 0.0;
          for (int t = 0; t < NV; ++t) s += L[r * NV + t] * Cprev[t * NV + cc];
          D[r * NV + cc] -= s;
        }
    } else {
      for (int e = 0; e < B; ++e) L[e] = 0.0;
    }
    inv_small<NV>(D, Dinv);
    for (int e = 0; e < B; ++e) {
      fac[(size_t)(0 * B + e) * plane + c] = Dinv[e];
      fac[(size_t)(2 * B + e) * plane + c] = L[e];
    }
    for (int r = 0; r < NV; ++r)
      for (int cc = 0; cc < NV; ++cc) {
        double s = 0.0;
        if (!jacobi && j < njc - 1)
          for (int t = 0; t < NV; ++t) s += Dinv[r * NV + t] * (-blocks[(size_t)(2 * B + t * NV + cc) * plane + c]);
        Cprev[r * NV + cc] = s;
        fac[(size_t)(1 * B + r * NV + cc) * plane + c] = s;
      }
  }
}

// Solve one factored line: x_line = T^-1 rhs_line (rhs in `t`, SoA).
template <int NV>
__device__ void line_substitute(int i, int njc, int ldi, size_t plane, const double* fac, const double* t, double* x) {
  constexpr int B = NV * NV;
  double y[NV];
  for (int k = 0; k < NV; ++k) y[k] = 0.0;
  for (int j = 0; j < njc; ++j) {
    const size_t c = (size_t)j * ldi + i;
    double r[NV];
    for (int k = 0; k < NV; ++k) {
      double s = t[k * plane + c];
      for (int m = 0; m < NV; ++m) s -= fac[(size_t)(2 * B + k * NV + m) * plane + c] * y[m];
      r[k] = s;
    }
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int m = 0; m < NV; ++m) s += fac[(size_t)(0 * B + k * NV + m) * plane + c] * r[m];
      y[k] = s;
      x[k * plane + c] = s;
    }
  }
  for (int j = njc - 2; j >= 0; --j) {
    const size_t c = (size_t)j * ldi + i;
    double cur[NV];
    for (int k = 0; k < NV; ++k) {
      double s = x[k * plane + c];
      for (int m = 0; m < NV; ++m) s -= fac[(size_t)(1 * B + k * NV + m) * plane + c] * y[m];
      cur[k] = s;
    }
    for (int k = 0; k < NV; ++k) { x[k * plane + c] = cur[k]; y[k] = cur[k]; }
  }
}

template <int NV>
__global__ void line_solve_kernel(int nic, int njc, int ldi, const double* fac, const double* b, double* x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nic) return;
  line_substitute<NV>(i, njc, ldi, (size_t)njc * ldi, fac, b, x);
}

// One colour of a line Gauss-Seidel sweep: for the i-lines with i % 3 ==
// colour, t = b - sum over the stencil blocks outside the line's tridiagonal
// band of (-J) x, then x_line = T^-1 t.  Lines of one colour never couple
// (the stencil reaches +-2 in i), so they run in parallel.
template <int NV>
__global__ void line_gs_kernel(int nic, int njc, int ldi, int colour, const double* blocks, const double* fac,
                               const double* b, double* t, double* x) {
  const int i = (blockIdx.x * blockDim.x + threadIdx.x) * 3 + colour;
  if (i >= nic) return;
  constexpr int B = NV * NV;
  const size_t plane = (size_t)njc * ldi;
  for (int j = 0; j < njc; ++j) {
    const size_t c = (size_t)j * ldi + i;
    double acc[NV];
    for (int k = 0; k < NV; ++k) acc[k] = b[k * plane + c];
    for (int s = 3; s < NSLOT; ++s) {
      const int ii = i + kSlotDi[s], jj = j + kSlotDj[s];
      if (ii < 0 || ii >= nic || jj < 0 || jj >= njc) continue;
      const size_t cn = (size_t)jj * ldi + ii;
      double xn[NV];
      for (int m = 0; m < NV; ++m) xn[m] = x[m * plane + cn];
      for (int k = 0; k < NV; ++k) {
        double sum = 0.0;
        for (int m = 0; m < NV; ++m) sum += blocks[(size_t)(s * B + k * NV + m) * plane + c] * xn[m];
        acc[k] += sum;  // A = I/dt - J: off-diagonal blocks are -J
      }
    }
    for (int k = 0; k < NV; ++k) t[k * plane + c] = acc[k];
  }
  line_substitute<NV>(i, njc, ldi, plane, fac, t, x);
}

hipError_t line_factor(int nic, int njc, int ldi, int nv, const double* idt, const double* blocks, double* fac,
                       bool jacobi, hipStream_t s) {
  dim3 g((nic + 63) / 64), b(64);
  if (nv == 5) hipLaunchKernelGGL(line_factor_kernel<5>, g, b, 0, s, nic, njc, ldi, idt, blocks, fac, jacobi ? 1 : 0);
  else hipLaunchKernelGGL(line_factor_kernel<4>, g, b, 0, s, nic, njc, ldi, idt, blocks, fac, jacobi ? 1 : 0);
  return hipGetLastError();
}

hipError_t line_solve(int nic, int njc, int ldi, int nv, const double* fac, const double* bvec, double* x, hipStream_t s) {
  dim3 g((nic + 63) / 64), b(64);
  if (nv == 5) hipLaunchKernelGGL(line_solve_kernel<5>, g, b, 0, s, nic, njc, ldi, fac, bvec, x);
  else hipLaunchKernelGGL(line_solve_kernel<4>, g, b, 0, s, nic, njc, ldi, fac, bvec, x);
  return hipGetLastError();
}

hipError_t line_gs(int nic, int njc, int ldi, int nv, int colour, const double* blocks, const double* fac,
                   const double* bvec, double* t, double* x, hipStream_t s) {
  const int nl = (nic - colour + 2) / 3;
  if (nl <= 0) return hipSuccess;
  dim3 g((nl + 63) / 64), b(64);
  if (nv == 5) hipLaunchKernelGGL(line_gs_kernel<5>, g, b, 0, s, nic, njc, ldi, colour, blocks, fac, bvec, t, x);
  else hipLaunchKernelGGL(line_gs_kernel<4>, g, b, 0, s, nic, njc, ldi, colour, blocks, fac, bvec, t, x);
  return hipGetLastError();
}

// ---- multicolour block ILU(0) on the 13-point stencil ---------------------
// Colour c(i,j) = (i + 3j) mod 5: no two cells of one colour are stencil
// neighbours, so with the colours ordered 0..4 the factorisation and both
// triangular solves run one colour per launch, fully parallel inside it.
// ilu (SoA, same layout as blocks) holds, per row r: slot 0 the inverse of
// the factored diagonal block, slots to lower-colour neighbours L_rn, slots
// to higher-colour neighbours U_rj.
__device__ __forceinline__ int colour5(int i, int j) { return ((i + 3 * j) % 5 + 5) % 5; }

__device__ __forceinline__ int slot_index(int di, int dj) {
  // inverse of kSlotDi/kSlotDj
  for (int s = 0; s < NSLOT; ++s)
    if (kSlotDi[s] == di && kSlotDj[s] == dj) return s;
  return -1;
}

template <int NV>
__global__ void ilu_init_kernel(int nic, int njc, int ldi, const double* idt, const double* blocks, double* ilu) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= nic) return;
  constexpr int B = NV * NV;
  const size_t plane = (size_t)njc * ldi, c = (size_t)j * ldi + i;
  for (int s = 0; s < NSLOT; ++s)
    for (int e = 0; e < B; ++e) {
      double a = -blocks[(size_t)(s * B + e) * plane + c];
      if (s == 0 && (e / NV) == (e % NV)) a += idt[c];
      ilu[(size_t)(s * B + e) * plane + c] = a;
    }
}

template <int NV>
__global__ void ilu_factor_kernel(int nic, int njc, int ldi, int colour, double* ilu) {
  const int j = blockIdx.y;
  const int i = ((colour - 3 * j) % 5 + 5) % 5 + 5 * (blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= nic) return;
  constexpr int B = NV * NV;
  const size_t plane = (size_t)njc * ldi, c = (size_t)j * ldi + i;
  auto at = [&](int s, size_t cell) -> double* { return ilu + (size_t)(s * B) * plane + cell; };  // stride plane
  for (int cn = 0; cn < colour; ++cn) {
    for (int s = 1; s < NSLOT; ++s) {
      const int di = kSlotDi[s], dj = kSlotDj[s];
      const int ni_ = i + di, nj_ = j + dj;
      if (ni_ < 0 || ni_ >= nic || nj_ < 0 || nj_ >= njc || colour5(ni_, nj_) != cn) continue;
      const size_t cnb = (size_t)nj_ * ldi + ni_;
      // L_rn = A_rn * Dinv_n
      double Lb[B];
      const double* Arn = at(s, c);
      const double* Dn = at(0, cnb);
      for (int r = 0; r < NV; ++r)
        for (int q = 0; q < NV; ++q) {
          double sum = 0.0;
          for (int t = 0; t < NV; ++t) sum += Arn[(r * NV + t) * plane] * Dn[(t * NV + q) * plane];
          Lb[r * NV + q] = sum;
        }
      double* Lout = at(s, c);
      for (int e = 0; e < B; ++e) Lout[e * plane] = Lb[e];
      // A_rj -= L_rn U_nj for the higher-colour neighbours j of n inside r's stencil
      for (int t = 1; t < NSLOT; ++t) {
        const int ji = ni_ + kSlotDi[t], jj = nj_ + kSlotDj[t];
        if (ji < 0 || ji >= nic || jj < 0 || jj >= njc || colour5(ji, jj) <= cn) continue;
        const int u = slot_index(di + kSlotDi[t], dj + kSlotDj[t]);
        if (u < 0) continue;
        const double* Unj = at(t, cnb);
        double* Arj = at(u, c);
        for (int r = 0; r < NV; ++r)
          for (int q = 0; q < NV; ++q) {
            double sum = 0.0;
            for (int k = 0; k < NV; ++k) sum += Lb[r * NV + k] * Unj[(k * NV + q) * plane];
            Arj[(r * NV + q) * plane] -= sum;
          }
      }
    }
  }
  double D[B], Dinv[B];
  double* Dd = at(0, c);
  for (int e = 0; e < B; ++e) D[e] = Dd[e * plane];
  inv_small<NV>(D, Dinv);
  for (int e = 0; e < B; ++e) Dd[e * plane] = Dinv[e];
}

// forward (lower = 1): z_r = b_r - sum_{lower n} L_rn z_n
// backward (lower = 0): z_r = Dinv_r (z_r - sum_{upper j} U_rj z_j)
template <int NV>
__global__ void ilu_solve_kernel(int nic, int njc, int ldi, int colour, int lower, const double* ilu, const double* b,
                                 double* z) {
  const int j = blockIdx.y;
  const int i = ((colour - 3 * j) % 5 + 5) % 5 + 5 * (blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= nic) return;
  constexpr int B = NV * NV;
  const size_t plane = (size_t)njc * ldi, c = (size_t)j * ldi + i;
  double acc[NV];
  for (int k = 0; k < NV; ++k) acc[k] = lower ? b[k * plane + c] : z[k * plane + c];
  for (int s = 1; s < NSLOT; ++s) {
    const int ni_ = i + kSlotDi[s], nj_ = j + kSlotDj[s];
    if (ni_ < 0 || ni_ >= nic || nj_ < 0 || nj_ >= njc) continue;
    const int cn = colour5(ni_, nj_);
    if (lower ? (cn >= colour) : (cn <= colour)) continue;
    const size_t cnb = (size_t)nj_ * ldi + ni_;
    double zn[NV];
    for (int m = 0; m < NV; ++m) zn[m] = z[m * plane + cnb];
    for (int k = 0; k < NV; ++k) {
      double sum = 0.0;
      for (int m = 0; m < NV; ++m) sum += ilu[(size_t)(s * B + k * NV + m) * plane + c] * zn[m];
      acc[k] -= sum;
    }
  }
  if (lower) {
    for (int k = 0; k < NV; ++k) z[k * plane + c] = acc[k];
  } else {
    for (int k = 0; k < NV; ++k) {
      double sum = 0.0;
      for (int m = 0; m < NV; ++m) sum += ilu[(size_t)(k * NV + m) * plane + c] * acc[m];
      z[k * plane + c] = sum;
    }
  }
}

hipError_t ilu_factor(int nic, int njc, int ldi, int nv, const double* idt, const double* blocks, double* ilu,
                      hipStream_t s) {
  dim3 gi((nic + 127) / 128, njc), bi(128);
  if (nv == 5) hipLaunchKernelGGL(ilu_init_kernel<5>, gi, bi, 0, s, nic, njc, ldi, idt, blocks, ilu);
  else hipLaunchKernelGGL(ilu_init_kernel<4>, gi, bi, 0, s, nic, njc, ldi, idt, blocks, ilu);
  const int per_row = (nic + 4) / 5;
  dim3 g((per_row + 63) / 64, njc), b(64);
  for (int c = 0; c < 5; ++c) {
    if (nv == 5) hipLaunchKernelGGL(ilu_factor_kernel<5>, g, b, 0, s, nic, njc, ldi, c, ilu);
    else hipLaunchKernelGGL(ilu_factor_kernel<4>, g, b, 0, s, nic, njc, ldi, c, ilu);
  }
  return hipGetLastError();
}

hipError_t ilu_apply(int nic, int njc, int ldi, int nv, const double* ilu, const double* bvec, double* z, hipStream_t s) {
  const int per_row = (nic + 4) / 5;
  dim3 g((per_row + 63) / 64, njc), b(64);
  for (int c = 0; c < 5; ++c) {
    if (nv == 5) hipLaunchKernelGGL(ilu_solve_kernel<5>, g, b, 0, s, nic, njc, ldi, c, 1, ilu, bvec, z);
    else hipLaunchKernelGGL(ilu_solve_kernel<4>, g, b, 0, s, nic, njc, ldi, c, 1, ilu, bvec, z);
  }
  for (int c = 4; c >= 0; --c) {
    if (nv == 5) hipLaunchKernelGGL(ilu_solve_kernel<5>, g, b, 0, s, nic, njc, ldi, c, 0, ilu, bvec, z);
    else hipLaunchKernelGGL(ilu_solve_kernel<4>, g, b, 0, s, nic, njc, ldi, c, 0, ilu, bvec, z);
  }
  return hipGetLastError();
}

// ---- level-scheduled block ILU(0) in natural j-major order -----------------
// Cells ordered o = j*nic + i (i fastest).  The lower neighbours of (i,j) on
// the 13-point stencil are (0,-2), (-1,-1), (0,-1), (1,-1), (-2,0), (-1,0)
// (listed in increasing o); all of them lie on an earlier level L = i + 2j,
// all upper neighbours on a later one.  One persistent work-group walks the
// levels with a barrier between them; the cells of one level run in
// parallel.  Storage as the multicolour variant (slot 0: inverse of the
// factored diagonal, lower slots L_rn, upper slots U_rj), SoA.
constexpr int NLOWER = 6;
__constant__ int kLowDi[NLOWER] = {0, -1, 0, 1, -2, -1};
__constant__ int kLowDj[NLOWER] = {-2, -1, -1, -1, 0, 0};

__device__ __forceinline__ bool is_lower(int di, int dj) { return dj < 0 || (dj == 0 && di < 0); }

template <int NV>
__global__ __launch_bounds__(1024) void lilu_factor_kernel(int nic, int njc, int ldi, double* ilu) {
  constexpr int B = NV * NV;
  const size_t plane = (size_t)njc * ldi;
  const int nlev = nic + 2 * (njc - 1);
  for (int L = 0; L < nlev; ++L) {
    const int jlo = max(0, (L - nic + 2) / 2), jhi = min(njc - 1, L / 2);
    for (int j = jlo + (int)threadIdx.x; j <= jhi; j += blockDim.x) {
      const int i = L - 2 * j;
      if (i < 0 || i >= nic) continue;
      const size_t c = (size_t)j * ldi + i;
      for (int q = 0; q < NLOWER; ++q) {
        const int di = kLowDi[q], dj = kLowDj[q];
        const int ni_ = i + di, nj_ = j + dj;
        if (ni_ < 0 || ni_ >= nic || nj_ < 0) continue;
        const size_t cnb = (size_t)nj_ * ldi + ni_;
        const int s = slot_index(di, dj);
        double Lb[B];
        for (int r = 0; r < NV; ++r)
          for (int t = 0; t < NV; ++t) {
            double sum = 0.0;
            for (int k = 0; k < NV; ++k)
              sum += ilu[(size_t)(s * B + r * NV + k) * plane + c] * ilu[(size_t)(0 * B + k * NV + t) * plane + cnb];
            Lb[r * NV + t] = sum;
          }
        for (int e = 0; e < B; ++e) ilu[(size_t)(s * B + e) * plane + c] = Lb[e];
        // upper neighbours u of n (o(u) > o(n)) that are also in r's stencil
        for (int t = 1; t < NSLOT; ++t) {
          const int ti = kSlotDi[t], tj = kSlotDj[t];
          if (is_lower(ti, tj)) continue;
          const int ui = ni_ + ti, uj = nj_ + tj;
          if (ui < 0 || ui >= nic || uj < 0 || uj >= njc) continue;
          const int u = slot_index(di + ti, dj + tj);
          if (u < 0) continue;
          for (int r = 0; r < NV; ++r)
            for (int q2 = 0; q2 < NV; ++q2) {
              double sum = 0.0;
              for (int k = 0; k < NV; ++k) sum += Lb[r * NV + k] * ilu[(size_t)(t * B + k * NV + q2) * plane + cnb];
              ilu[(size_t)(u * B + r * NV + q2) * plane + c] -= sum;
            }
        }
      }
      double D[B], Dinv[B];
      for (int e = 0; e < B; ++e) D[e] = ilu[(size_t)e * plane + c];
      inv_small<NV>(D, Dinv);
      for (int e = 0; e < B; ++e) ilu[(size_t)e * plane + c] = Dinv[e];
    }
    __threadfence_block();
    __syncthreads();
  }
}

// mode 0: z = (LU)^-1 b ;  mode 1: z = (LU)^-T b  (transposed, for the adjoint)
template <int NV>
__global__ __launch_bounds__(1024) void lilu_solve_kernel(int nic, int njc, int ldi, int mode, const double* ilu,
                                                          const double* b, double* z) {
  constexpr int B = NV * NV;
  const size_t plane = (size_t)njc * ldi;
  const int nlev = nic + 2 * (njc - 1);
  for (int pass = 0; pass < 2; ++pass) {
    // mode 0: pass 0 = L (ascending), pass 1 = U (descending)
    // mode 1: pass 0 = U^T (ascending), pass 1 = L^T (descending)
    const bool asc = pass == 0;
    for (int step = 0; step < nlev; ++step) {
      const int L = asc ? step : nlev - 1 - step;
      const int jlo = max(0, (L - nic + 2) / 2), jhi = min(njc - 1, L / 2);
      for (int j = jlo + (int)threadIdx.x; j <= jhi; j += blockDim.x) {
        const int i = L - 2 * j;
        if (i < 0 || i >= nic) continue;
        const size_t c = (size_t)j * ldi + i;
        double acc[NV];
        for (int k = 0; k < NV; ++k) acc[k] = (pass == 0) ? b[k * plane + c] : z[k * plane + c];
        for (int s = 1; s < NSLOT; ++s) {
          const int di = kSlotDi[s], dj = kSlotDj[s];
          const bool low = is_lower(di, dj);
          if (low != asc) continue;  // ascending passes use earlier-level neighbours
          const int ni_ = i + di, nj_ = j + dj;
          if (ni_ < 0 || ni_ >= nic || nj_ < 0 || nj_ >= njc) continue;
          const size_t cnb = (size_t)nj_ * ldi + ni_;
          double zn[NV];
          for (int m = 0; m < NV; ++m) zn[m] = z[m * plane + cnb];
          if (mode == 0) {
            for (int k = 0; k < NV; ++k) {
              double sum = 0.0;
              for (int m = 0; m < NV; ++m) sum += ilu[(size_t)(s * B + k * NV + m) * plane + c] * zn[m];
              acc[k] -= sum;
            }
          } else {
            // transposed: the entry of row n at the slot pointing back to r
            const int sb = slot_index(-di, -dj);
            for (int k = 0; k < NV; ++k) {
              double sum = 0.0;
              for (int m = 0; m < NV; ++m) sum += ilu[(size_t)(sb * B + m * NV + k) * plane + cnb] * zn[m];
              acc[k] -= sum;
            }
          }
        }
        const bool diag = (mode == 0) ? (pass == 1) : (pass == 0);
        if (diag) {
          for (int k = 0; k < NV; ++k) {
            double sum = 0.0;
            for (int m = 0; m < NV; ++m)
              sum += ((mode == 0) ? ilu[(size_t)(k * NV + m) * plane + c] : ilu[(size_t)(m * NV + k) * plane + c]) * acc[m];
            z[k * plane + c] = sum;
          }
        } else {
          for (int k = 0; k < NV; ++k) z[k * plane + c] = acc[k];
        }
      }
      __threadfence_block();
      __syncthreads();
    }
  }
}

hipError_t lilu_factor(int nic, int njc, int ldi, int nv, const double* idt, const double* blocks, double* ilu,
                       hipStream_t s) {
  dim3 gi((nic + 127) / 128, njc), bi(128);
  if (nv == 5) hipLaunchKernelGGL(ilu_init_kernel<5>, gi, bi, 0, s, nic, njc, ldi, idt, blocks, ilu);
  else hipLaunchKernelGGL(ilu_init_kernel<4>, gi, bi, 0, s, nic, njc, ldi, idt, blocks, ilu);
  if (nv == 5) hipLaunchKernelGGL(lilu_factor_kernel<5>, dim3(1), dim3(1024), 0, s, nic, njc, ldi, ilu);
  else hipLaunchKernelGGL(lilu_factor_kernel<4>, dim3(1), dim3(1024), 0, s, nic, njc, ldi, ilu);
  return hipGetLastError();
}

hipError_t lilu_apply(int nic, int njc, int ldi, int nv, const double* ilu, const double* bvec, double* z, bool transpose,
                      hipStream_t s) {
  if (nv == 5) hipLaunchKernelGGL(lilu_solve_kernel<5>, dim3(1), dim3(1024), 0, s, nic, njc, ldi, transpose ? 1 : 0, ilu, bvec, z);
  else hipLaunchKernelGGL(lilu_solve_kernel<4>, dim3(1), dim3(1024), 0, s, nic, njc, ldi, transpose ? 1 : 0, ilu, bvec, z);
  return hipGetLastError();
}

// ---- adjoint products from the Jacobian blocks ----------------------------
// (J^T psi)_c = sum_s J[c - d_s, c]^T psi_{c - d_s}, with J[r, r + d_s]
// stored in slot s of row r: a 13-neighbour gather, deterministic.
template <int NV>
__global__ void jt_gather_kernel(int nic, int njc, int ldi, const double* blocks, const double* psi, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= nic) return;
  constexpr int B = NV * NV;
  const size_t plane = (size_t)njc * ldi, c = (size_t)j * ldi + i;
  double acc[NV];
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (int s = 0; s < NSLOT; ++s) {
    const int ri = i - kSlotDi[s], rj = j - kSlotDj[s];
    if (ri < 0 || ri >= nic || rj < 0 || rj >= njc) continue;
    const size_t rc = (size_t)rj * ldi + ri;
    double pr[NV];
    for (int m = 0; m < NV; ++m) pr[m] = psi[m * plane + rc];
    for (int k = 0; k < NV; ++k) {
      double sum = 0.0;
      for (int m = 0; m < NV; ++m) sum += blocks[(size_t)(s * B + m * NV + k) * plane + rc] * pr[m];
      acc[k] += sum;
    }
  }
  for (int k = 0; k < NV; ++k) out[k * plane + c] = acc[k];
}

hipError_t jt_gather(int nic, int njc, int ldi, int nv, const double* blocks, const double* psi, double* out,
                     hipStream_t s) {
  dim3 g((nic + 127) / 128, njc), b(128);
  if (nv == 5) hipLaunchKernelGGL(jt_gather_kernel<5>, g, b, 0, s, nic, njc, ldi, blocks, psi, out);
  else hipLaunchKernelGGL(jt_gather_kernel<4>, g, b, 0, s, nic, njc, ldi, blocks, psi, out);
  return hipGetLastError();
}

// w = z * idt - J^T z   (transposed operator I/dt - J^T)
// dbeta_c = psi_4c * dR4_c/dbeta_c  (dR/dbeta is diagonal: beta_c enters only
// the SA production of cell c)
__global__ void dbeta_kernel(int nic, int njc, int ldi, const double* psi, const double* diag, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= nic) return;
  const size_t plane = (size_t)njc * ldi, c = (size_t)j * ldi + i;
  out[c] = psi[4 * plane + c] * diag[4 * plane + c];
}
hipError_t dbeta_apply(int nic, int njc, int ldi, const double* psi, const double* diag, double* out, hipStream_t s) {
  hipLaunchKernelGGL(dbeta_kernel, dim3((nic + 127) / 128, njc), dim3(128), 0, s, nic, njc, ldi, psi, diag, out);
  return hipGetLastError();
}

// Transposed five-colour ILU(0): (LU)^T z = b, U^T ascending colours with the
// transposed diagonal inverse, then L^T descending colours.
template <int NV>
__global__ void ilu_solve_t_kernel(int nic, int njc, int ldi, int colour, int pass, const double* ilu, const double* b,
                                   double* z) {
  const int j = blockIdx.y;
  const int i = ((colour - 3 * j) % 5 + 5) % 5 + 5 * (blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= nic) return;
  constexpr int B = NV * NV;
  const size_t plane = (size_t)njc * ldi, c = (size_t)j * ldi + i;
  double acc[NV];
  for (int k = 0; k < NV; ++k) acc[k] = pass == 0 ? b[k * plane + c] : z[k * plane + c];
  for (int s = 1; s < NSLOT; ++s) {
    const int ni_ = i + kSlotDi[s], nj_ = j + kSlotDj[s];
    if (ni_ < 0 || ni_ >= nic || nj_ < 0 || nj_ >= njc) continue;
    const int cn = colour5(ni_, nj_);
    if (pass == 0 ? (cn >= colour) : (cn <= colour)) continue;
    const int sb = slot_index(-kSlotDi[s], -kSlotDj[s]);  // entry of row n pointing back to this cell
    const size_t cnb = (size_t)nj_ * ldi + ni_;
    double zn[NV];
    for (int m = 0; m < NV; ++m) zn[m] = z[m * plane + cnb];
    for (int k = 0; k < NV; ++k) {
      double sum = 0.0;
      for (int m = 0; m < NV; ++m) sum += ilu[(size_t)(sb * B + m * NV + k) * plane + cnb] * zn[m];
      acc[k] -= sum;
    }
  }
  if (pass == 0) {
    for (int k = 0; k < NV; ++k) {
      double sum = 0.0;
      for (int m = 0; m < NV; ++m) sum += ilu[(size_t)(m * NV + k) * plane + c] * acc[m];
      z[k * plane + c] = sum;
    }
  } else {
    for (int k = 0; k < NV; ++k) z[k * plane + c] = acc[k];
  }
}

hipError_t ilu_apply_t(int nic, int njc, int ldi, int nv, const double* ilu, const double* bvec, double* z, hipStream_t s) {
  const int per_row = (nic + 4) / 5;
  dim3 g((per_row + 63) / 64, njc), b(64);
  for (int c = 0; c < 5; ++c) {
    if (nv == 5) hipLaunchKernelGGL(ilu_solve_t_kernel<5>, g, b, 0, s, nic, njc, ldi, c, 0, ilu, bvec, z);
    else hipLaunchKernelGGL(ilu_solve_t_kernel<4>, g, b, 0, s, nic, njc, ldi, c, 0, ilu, bvec, z);
  }
  for (int c = 4; c >= 0; --c) {
    if (nv == 5) hipLaunchKernelGGL(ilu_solve_t_kernel<5>, g, b, 0, s, nic, njc, ldi, c, 1, ilu, bvec, z);
    else hipLaunchKernelGGL(ilu_solve_t_kernel<4>, g, b, 0, s, nic, njc, ldi, c, 1, ilu, bvec, z);
  }
  return hipGetLastError();
}

// Largest step fraction alpha <= 1 with |d rho| <= f rho and |d p| <= f p per
// cell (d p linearised: (gamma-1)(dE - u d(rho u) - v d(rho v) + |u|^2/2 d rho));
// per-block minima, reduced on the host (min is order independent).
__global__ __launch_bounds__(256) void limit_kernel(int nic, int njc, int ldi, const double* q, const double* dq,
                                                    double f, double* block_min) {
  __shared__ double red[256];
  const size_t plane = (size_t)njc * ldi;
  double a = 1.0;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < (size_t)nic * njc; e += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(e / nic), i = (int)(e % nic);
    const size_t c = (size_t)j * ldi + i;
    const double r = q[c], ru = q[plane + c], rv = q[2 * plane + c], rE = q[3 * plane + c];
    const double u = ru / r, v = rv / r, p = 0.4 * (rE - 0.5 * r * (u * u + v * v));
    const double dr = dq[c], dp = 0.4 * (dq[3 * plane + c] - u * dq[plane + c] - v * dq[2 * plane + c] +
                                         0.5 * (u * u + v * v) * dr);
    if (!(r > 0.0) || !(p > 0.0)) { a = 0.0; continue; }
    if (fabs(dr) > f * r) a = fmin(a, f * r / fabs(dr));
    if (fabs(dp) > f * p) a = fmin(a, f * p / fabs(dp));
    if (!isfinite(dr) || !isfinite(dp)) a = 0.0;
  }
  red[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) block_min[blockIdx.x] = red[0];
}
hipError_t update_limit(int nic, int njc, int ldi, const double* q, const double* dq, double f, double* block_min,
                        int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(limit_kernel, dim3(nblocks), dim3(256), 0, s, nic, njc, ldi, q, dq, f, block_min);
  return hipGetLastError();
}

// inverse local time step per cell: idt = 1/dt (dt from the dt kernel, SoA plane 0)
__global__ void invert_kernel(const double* dt, double* idt, int nic, int njc, int ldi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= nic) return;
  const size_t c = (size_t)j * ldi + i;
  idt[c] = 1.0 / dt[c];
}
hipError_t invert_dt(const double* dt, double* idt, int nic, int njc, int ldi, hipStream_t s) {
  hipLaunchKernelGGL(invert_kernel, dim3((nic + 127) / 128, njc), dim3(128), 0, s, dt, idt, nic, njc, ldi);
  return hipGetLastError();
}

}  // namespace sg
