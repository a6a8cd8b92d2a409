// Forward-mode dual numbers for the discrete adjoint: a value and N tangents.
// The interface fluxes and limiter functions are written once as templates;
// instantiated with Dual<N> they return exact derivatives of the same branch
// the value computation takes (comparisons look at values only).
//
// Non-smooth points: fabs at 0 has slope 0, and fmin/fmax at an exact tie take
// the mean of both slopes, which is what a central difference sees there.
#pragma once

#include <hip/hip_runtime.h>

namespace cfd {
namespace ad {

template <int N>
struct Dual {
  double v;
  double d[N];
  __device__ __forceinline__ Dual() {}
  __device__ __forceinline__ Dual(double x) : v(x) {  // NOLINT: constants promote implicitly
#pragma unroll
    for (int k = 0; k < N; ++k) d[k] = 0.0;
  }
};

template <int N>
__device__ __forceinline__ Dual<N> seed(double x, int k) {
  Dual<N> r(x);
  r.d[k] = 1.0;
  return r;
}

template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  const double ib = 1.0 / b.v;
  r.v = a.v / b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
  return r;
}
// Mixed forms with a double on one side (no tangent work for the constant).
template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, double b) {
  Dual<N> r = a;
  r.v = a.v + b;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(double a, const Dual<N>& b) {
  return b + a;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, double b) {
  Dual<N> r = a;
  r.v = a.v - b;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(double a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a - b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = -b.d[k];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, double b) {
  Dual<N> r;
  r.v = a.v * b;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(double a, const Dual<N>& b) {
  return b * a;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, double b) {
  Dual<N> r;
  const double ib = 1.0 / b;
  r.v = a.v / b;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * ib;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(double a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a / b.v;
  const double s = -r.v / b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = s * b.d[k];
  return r;
}

template <int N> __device__ __forceinline__ bool operator<(const Dual<N>& a, const Dual<N>& b) { return a.v < b.v; }
template <int N> __device__ __forceinline__ bool operator>(const Dual<N>& a, const Dual<N>& b) { return a.v > b.v; }
template <int N> __device__ __forceinline__ bool operator<=(const Dual<N>& a, const Dual<N>& b) { return a.v <= b.v; }
template <int N> __device__ __forceinline__ bool operator>=(const Dual<N>& a, const Dual<N>& b) { return a.v >= b.v; }
template <int N> __device__ __forceinline__ bool operator<(const Dual<N>& a, double b) { return a.v < b; }
template <int N> __device__ __forceinline__ bool operator>(const Dual<N>& a, double b) { return a.v > b; }
template <int N> __device__ __forceinline__ bool operator<=(const Dual<N>& a, double b) { return a.v <= b; }
template <int N> __device__ __forceinline__ bool operator>=(const Dual<N>& a, double b) { return a.v >= b; }
template <int N> __device__ __forceinline__ bool operator==(const Dual<N>& a, double b) { return a.v == b; }
template <int N> __device__ __forceinline__ bool operator!=(const Dual<N>& a, double b) { return a.v != b; }

template <int N>
__device__ __forceinline__ Dual<N> sqrt(const Dual<N>& a) {
  Dual<N> r;
  r.v = ::sqrt(a.v);
  const double s = 0.5 / r.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * s;
  return r;
}

template <int N>
__device__ __forceinline__ Dual<N> fabs(const Dual<N>& a) {
  Dual<N> r;
  r.v = ::fabs(a.v);
  const double s = a.v > 0.0 ? 1.0 : (a.v < 0.0 ? -1.0 : 0.0);
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * s;
  return r;
}

template <int N>
__device__ __forceinline__ Dual<N> fmin(const Dual<N>& a, const Dual<N>& b) {
  if (a.v < b.v) return a;
  if (b.v < a.v) return b;
  Dual<N> r;
  r.v = a.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = 0.5 * (a.d[k] + b.d[k]);
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> fmax(const Dual<N>& a, const Dual<N>& b) {
  if (a.v > b.v) return a;
  if (b.v > a.v) return b;
  Dual<N> r;
  r.v = a.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = 0.5 * (a.d[k] + b.d[k]);
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> fmin(double a, const Dual<N>& b) { return fmin(Dual<N>(a), b); }
template <int N> __device__ __forceinline__ Dual<N> fmin(const Dual<N>& a, double b) { return fmin(a, Dual<N>(b)); }
template <int N> __device__ __forceinline__ Dual<N> fmax(double a, const Dual<N>& b) { return fmax(Dual<N>(a), b); }
template <int N> __device__ __forceinline__ Dual<N> fmax(const Dual<N>& a, double b) { return fmax(a, Dual<N>(b)); }

}  // namespace ad

// Value part of a double or a dual number.
__device__ __forceinline__ double value_of(double x) { return x; }
template <int N>
__device__ __forceinline__ double value_of(const ad::Dual<N>& x) {
  return x.v;
}

}  // namespace cfd
