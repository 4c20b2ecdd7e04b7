// salsa_dual.h -- first-order dual numbers (value, tangent) for the
// forward-mode derivative kernel.  Dual<double> carries the FP64 state,
// Dual<float> the FP32 collision-kernel coefficients of the mixed path.
//
// Comparisons and min/max look at the value only, so a dual evaluation takes
// exactly the branches of the plain evaluation at the same point and its
// tangent is the derivative of the branch taken.
#pragma once

#include <hip/hip_runtime.h>

namespace salsa {

template <class T>
struct Dual {
  T v, d;
};

// value-level math of the two scalar types
template <class T> struct DMath;
template <> struct DMath<double> {
  static __device__ __forceinline__ double sqrt_(double x) { return ::sqrt(x); }
  static __device__ __forceinline__ double exp_(double x) { return ::exp(x); }
  static __device__ __forceinline__ double log_(double x) { return ::log(x); }
  static __device__ __forceinline__ double cbrt_(double x) { return ::cbrt(x); }
  static __device__ __forceinline__ double pow_(double x, double y) { return ::pow(x, y); }
  static __device__ __forceinline__ double rcp_(double x) { return 1. / x; }
  static __device__ __forceinline__ double div_(double a, double b) { return a / b; }
};
template <> struct DMath<float> {
  // the hardware transcendentals, as the FP32 kernel coefficients use them
  static __device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ __forceinline__ float exp_(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
  static __device__ __forceinline__ float log_(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945f; }
  static __device__ __forceinline__ float cbrt_(float x) {
    return x > 0.f ? __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (1.f / 3.f)) : 0.f;
  }
  static __device__ __forceinline__ float pow_(float x, float y) {
    return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * y);
  }
  static __device__ __forceinline__ float rcp_(float x) { return __builtin_amdgcn_rcpf(x); }
  static __device__ __forceinline__ float div_(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
};

template <class T> __device__ __forceinline__ Dual<T> dconst(T x) { return Dual<T>{x, T(0)}; }
template <class T> __device__ __forceinline__ Dual<T> dvar(T x, T dx) { return Dual<T>{x, dx}; }
template <class U, class T> __device__ __forceinline__ Dual<U> dcast(Dual<T> a) { return Dual<U>{(U)a.v, (U)a.d}; }

// arithmetic
template <class T> __device__ __forceinline__ Dual<T> operator-(Dual<T> a) { return {-a.v, -a.d}; }
template <class T> __device__ __forceinline__ Dual<T> operator+(Dual<T> a, Dual<T> b) { return {a.v + b.v, a.d + b.d}; }
template <class T> __device__ __forceinline__ Dual<T> operator+(Dual<T> a, T b) { return {a.v + b, a.d}; }
template <class T> __device__ __forceinline__ Dual<T> operator+(T a, Dual<T> b) { return {a + b.v, b.d}; }
template <class T> __device__ __forceinline__ Dual<T> operator-(Dual<T> a, Dual<T> b) { return {a.v - b.v, a.d - b.d}; }
template <class T> __device__ __forceinline__ Dual<T> operator-(Dual<T> a, T b) { return {a.v - b, a.d}; }
template <class T> __device__ __forceinline__ Dual<T> operator-(T a, Dual<T> b) { return {a - b.v, -b.d}; }
template <class T> __device__ __forceinline__ Dual<T> operator*(Dual<T> a, Dual<T> b) {
  return {a.v * b.v, a.d * b.v + a.v * b.d};
}
template <class T> __device__ __forceinline__ Dual<T> operator*(Dual<T> a, T b) { return {a.v * b, a.d * b}; }
template <class T> __device__ __forceinline__ Dual<T> operator*(T a, Dual<T> b) { return {a * b.v, a * b.d}; }
template <class T> __device__ __forceinline__ Dual<T> operator/(Dual<T> a, Dual<T> b) {
  const T r = DMath<T>::rcp_(b.v);
  const T q = DMath<T>::div_(a.v, b.v);
  return {q, (a.d - q * b.d) * r};
}
template <class T> __device__ __forceinline__ Dual<T> operator/(Dual<T> a, T b) {
  return {DMath<T>::div_(a.v, b), DMath<T>::div_(a.d, b)};
}
template <class T> __device__ __forceinline__ Dual<T> operator/(T a, Dual<T> b) {
  const T q = DMath<T>::div_(a, b.v);
  return {q, -q * b.d * DMath<T>::rcp_(b.v)};
}
template <class T> __device__ __forceinline__ Dual<T> &operator+=(Dual<T> &a, Dual<T> b) { a = a + b; return a; }
template <class T> __device__ __forceinline__ Dual<T> &operator-=(Dual<T> &a, Dual<T> b) { a = a - b; return a; }

// comparisons on values
template <class T> __device__ __forceinline__ bool operator<(Dual<T> a, Dual<T> b) { return a.v < b.v; }
template <class T> __device__ __forceinline__ bool operator<(Dual<T> a, T b) { return a.v < b; }
template <class T> __device__ __forceinline__ bool operator<(T a, Dual<T> b) { return a < b.v; }
template <class T> __device__ __forceinline__ bool operator>(Dual<T> a, Dual<T> b) { return a.v > b.v; }
template <class T> __device__ __forceinline__ bool operator>(Dual<T> a, T b) { return a.v > b; }
template <class T> __device__ __forceinline__ bool operator>(T a, Dual<T> b) { return a > b.v; }
template <class T> __device__ __forceinline__ bool operator<=(Dual<T> a, T b) { return a.v <= b; }
template <class T> __device__ __forceinline__ bool operator>=(Dual<T> a, Dual<T> b) { return a.v >= b.v; }
template <class T> __device__ __forceinline__ bool operator>=(Dual<T> a, T b) { return a.v >= b; }

// elementary functions
template <class T> __device__ __forceinline__ Dual<T> dsqrt(Dual<T> x) {
  const T s = DMath<T>::sqrt_(x.v);
  return {s, s > T(0) ? x.d * (T(0.5) * DMath<T>::rcp_(s)) : T(0)};
}
template <class T> __device__ __forceinline__ Dual<T> dexp(Dual<T> x) {
  const T e = DMath<T>::exp_(x.v);
  return {e, x.d * e};
}
template <class T> __device__ __forceinline__ Dual<T> dlog(Dual<T> x) {
  return {DMath<T>::log_(x.v), x.d * DMath<T>::rcp_(x.v)};
}
// erf, FP64 only (d/dx erf x = 2/sqrt(pi) exp(-x^2))
__device__ __forceinline__ Dual<double> derf(Dual<double> x) {
  return {::erf(x.v), x.d * (1.1283791670955126 * ::exp(-(x.v * x.v)))};
}
template <class T> __device__ __forceinline__ Dual<T> dcbrt(Dual<T> x) {
  const T c = DMath<T>::cbrt_(x.v);
  return {c, c != T(0) ? x.d * DMath<T>::rcp_(T(3) * c * c) : T(0)};
}
// x**y, constant exponent
template <class T> __device__ __forceinline__ Dual<T> dpow(Dual<T> x, T y) {
  const T p = DMath<T>::pow_(x.v, y);
  return {p, x.v != T(0) ? x.d * y * p * DMath<T>::rcp_(x.v) : T(0)};
}
// x**y, both varying (x > 0)
template <class T> __device__ __forceinline__ Dual<T> dpow(Dual<T> x, Dual<T> y) {
  const T p = DMath<T>::pow_(x.v, y.v);
  return {p, p * (y.d * DMath<T>::log_(x.v) + y.v * x.d * DMath<T>::rcp_(x.v))};
}
// constant base, varying exponent (x > 0)
template <class T> __device__ __forceinline__ Dual<T> dpow(T x, Dual<T> y) {
  const T p = DMath<T>::pow_(x, y.v);
  return {p, p * DMath<T>::log_(x) * y.d};
}
template <class T> __device__ __forceinline__ Dual<T> dmin(Dual<T> a, Dual<T> b) { return b.v < a.v ? b : a; }
template <class T> __device__ __forceinline__ Dual<T> dmin(Dual<T> a, T b) { return b < a.v ? dconst(b) : a; }
template <class T> __device__ __forceinline__ Dual<T> dmax(Dual<T> a, Dual<T> b) { return b.v > a.v ? b : a; }
template <class T> __device__ __forceinline__ Dual<T> dmax(Dual<T> a, T b) { return b > a.v ? dconst(b) : a; }
template <class T> __device__ __forceinline__ Dual<T> dmax(T a, Dual<T> b) { return b.v > a ? b : dconst(a); }
template <class T> __device__ __forceinline__ Dual<T> dabs(Dual<T> a) { return a.v < T(0) ? -a : a; }
template <class T> __device__ __forceinline__ Dual<T> dcube(Dual<T> a) { return a * a * a; }

}  // namespace salsa
