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

__device__ __forceinline__ double val(double x) { return x; }
__device__ __forceinline__ double val(Dual x) { return x.v; }
__device__ __forceinline__ double tan_(double) { return 0.0; }
__device__ __forceinline__ double tan_(Dual x) { return x.d; }

__device__ __forceinline__ double ssqrt(double x) { return sqrt(x); }
__device__ __forceinline__ Dual ssqrt(Dual x) {
  double r = sqrt(x.v);
  return Dual(r, x.d / (2.0 * r));
}
__device__ __forceinline__ double sabs(double x) { return fabs(x); }
__device__ __forceinline__ Dual sabs(Dual x) { return x.v < 0.0 ? -x : x; }
// x^(2/3) for the power-law viscosity (src/model/fluid.cpp:38-40).
__device__ __forceinline__ double spow23(double x) { return pow(x, 2.0 / 3.0); }
__device__ __forceinline__ Dual spow23(Dual x) {
  double r = pow(x.v, 2.0 / 3.0);
  return Dual(r, (2.0 / 3.0) * r / x.v * x.d);
}
__device__ __forceinline__ double spow(double x, double e) { return pow(x, e); }
__device__ __forceinline__ Dual spow(Dual x, double e) {
  double r = pow(x.v, e);
  return Dual(r, e * pow(x.v, e - 1.0) * x.d);
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
