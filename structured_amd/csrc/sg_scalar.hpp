// Scalar types for the device kernels: plain double, and a forward dual
// number (value + one tangent) so that the same templated residual code
// yields the exact matrix-free tangent J.v in one sweep.
#pragma once
#include <hip/hip_runtime.h>

namespace sg {

struct Dual {
  double v, d;
  __host__ __device__ Dual() : v(0.0), d(0.0) {}
  __host__ __device__ Dual(double x) : v(x), d(0.0) {}
  __host__ __device__ Dual(double x, double t) : v(x), d(t) {}
};

__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  double r = a.v / b.v;
  return Dual(r, (a.d - r * b.d) / b.v);
}
__device__ __forceinline__ Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
__device__ __forceinline__ Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
__device__ __forceinline__ Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
__device__ __forceinline__ Dual operator/(double a, Dual b) {
  double r = a / b.v;
  return Dual(r, -r * b.d / b.v);
}
__device__ __forceinline__ Dual& operator+=(Dual& a, Dual b) { a = a + b; return a; }
__device__ __forceinline__ Dual& operator-=(Dual& a, Dual b) { a = a - b; return a; }

// Reciprocal from the hardware estimate refined by two Newton steps
// (r <- r + r (1 - x r)); within ~1 ulp for the normal, finite operands of
// this solver.  Used in place of IEEE division on the hot path.
__device__ __forceinline__ double frcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}
__device__ __forceinline__ double srcp(double x) { return frcp(x); }
__device__ __forceinline__ Dual srcp(Dual x) {
  double r = frcp(x.v);
  return Dual(r, -r * r * x.d);
}
__device__ __forceinline__ double sdiv(double a, double b) { return a * frcp(b); }
__device__ __forceinline__ Dual sdiv(Dual a, Dual b) {
  double ib = frcp(b.v), r = a.v * ib;
  return Dual(r, (a.d - r * b.d) * ib);
}
__device__ __forceinline__ Dual sdiv(Dual a, double b) {
  double ib = frcp(b);
  return Dual(a.v * ib, a.d * ib);
}
__device__ __forceinline__ Dual sdiv(double a, Dual b) {
  double ib = frcp(b.v), r = a * ib;
  return Dual(r, -r * b.d * ib);
}

// x^(2/3) for x > 0 without pow: z = x^(-2/3) solves z^-3 = x^2; seed from
// the single-precision log2/exp2 units, then two Newton steps
// z <- z (4 - x^2 z^3) / 3 (relative error 2e^2 per step), and
// x^(2/3) = x^2 z^2.
__device__ __forceinline__ double pow23(double x) {
  const float xf = (float)x;
  const float zf = __builtin_amdgcn_exp2f(-(2.0f / 3.0f) * __builtin_amdgcn_logf(xf));
  const double a = x * x;
  double z = (double)zf;
  z = z * (4.0 - a * (z * z * z)) * (1.0 / 3.0);
  z = z * (4.0 - a * (z * z * z)) * (1.0 / 3.0);
  return a * z * z;
}

__device__ __forceinline__ double val(double x) { return x; }
__device__ __forceinline__ double val(Dual x) { return x.v; }
__device__ __forceinline__ double tan_(double) { return 0.0; }
__device__ __forceinline__ double tan_(Dual x) { return x.d; }

// Square root of a positive normal x from the hardware reciprocal square
// root estimate: y = x r, h = r / 2, one coupled Newton (Goldschmidt) step,
// then two residual corrections y + (x - y^2) h (Markstein).  No scaling of
// tiny or huge inputs and no special cases: 0, infinities and negatives are
// not handled (the callers pass squared magnitudes bounded away from zero).
__device__ __forceinline__ double fsqrt(double x) {
  const double r = __builtin_amdgcn_rsq(x);
  double y = x * r, h = 0.5 * r;
  const double e = __builtin_fma(-y, h, 0.5);
  y = __builtin_fma(y, e, y);
  h = __builtin_fma(h, e, h);
  double d = __builtin_fma(-y, y, x);
  y = __builtin_fma(d, h, y);
  d = __builtin_fma(-y, y, x);
  return __builtin_fma(d, h, y);
}

// z^(1/6) for z > 0 without pow: seed from the single-precision log2 / exp2
// units, then two Newton steps on y^6 = z: y <- y (5 + z / y^6) / 6.
__device__ __forceinline__ double pow16(double z) {
  double y = (double)__builtin_amdgcn_exp2f((1.0f / 6.0f) * __builtin_amdgcn_logf((float)z));
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const double y2 = y * y;
    y = y * (5.0 + z * frcp(y2 * y2 * y2)) * (1.0 / 6.0);
  }
  return y;
}

__device__ __forceinline__ double ssqrt(double x) { return fsqrt(x); }
__device__ __forceinline__ Dual ssqrt(Dual x) {
  double r = fsqrt(x.v);
  return Dual(r, x.d * frcp(2.0 * r));
}
__device__ __forceinline__ double sabs(double x) { return fabs(x); }
__device__ __forceinline__ Dual sabs(Dual x) { return x.v < 0.0 ? -x : x; }
// x^(2/3) for the power-law viscosity (src/model/fluid.cpp:38-40).
__device__ __forceinline__ double spow23(double x) { return pow23(x); }
__device__ __forceinline__ Dual spow23(Dual x) {
  double r = pow23(x.v);
  return Dual(r, (2.0 / 3.0) * r * frcp(x.v) * x.d);
}
__device__ __forceinline__ double spow(double x, double e) { return pow(x, e); }
__device__ __forceinline__ Dual spow(Dual x, double e) {
  double r = pow(x.v, e);
  return Dual(r, e * pow(x.v, e - 1.0) * x.d);
}
// z^(1/6) (SA destruction f_w)
__device__ __forceinline__ double spow16(double z) { return pow16(z); }
__device__ __forceinline__ Dual spow16(Dual z) {
  const double y = pow16(z.v);
  return Dual(y, (1.0 / 6.0) * y * frcp(z.v) * z.d);
}
__device__ __forceinline__ double smin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ Dual smin(Dual a, double b) { return a.v < b ? a : Dual(b); }

// Loads of a state component for scalar type S: the dual reads the tangent
// from a second array of the same layout.
template <class S>
struct Field;
template <>
struct Field<double> {
  const double* v;
  __device__ __forceinline__ double operator[](size_t k) const { return v[k]; }
};
template <>
struct Field<Dual> {
  const double* v;
  const double* d;
  __device__ __forceinline__ Dual operator[](size_t k) const { return Dual(v[k], d[k]); }
};

template <class S>
struct OutField;
template <>
struct OutField<double> {
  double* v;
  __device__ __forceinline__ void put(size_t k, double x) const { v[k] = x; }
};
template <>
struct OutField<Dual> {
  double* v;  // may be null: value not stored
  double* d;
  __device__ __forceinline__ void put(size_t k, Dual x) const {
    if (v) v[k] = x.v;
    d[k] = x.d;
  }
};

}  // namespace sg
