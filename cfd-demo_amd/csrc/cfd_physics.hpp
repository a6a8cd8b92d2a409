// Pointwise 2-D Euler physics, interface fluxes and limiter functions as
// gfx950 device functions. Conserved variables are (rho, rho*u, rho*v, E).
//
// Formulas follow src/cfd_demo/euler.py:21-61, fluxes.py:24-143 and
// limiters.py:16-39. The order of operations follows the reference where it
// decides branches (HLLC wave-speed signs, Roe entropy fix); elsewhere the
// compiler may contract a*b+c into FMAs, and the flux divisions and square
// roots use the hardware reciprocal / reciprocal square root refined by
// Newton steps (frcp, fdiv, fsqrt below), which moves results by a few ulps.
// Decisions that compare against thresholds (pressure_of, the limiter) keep
// correctly rounded arithmetic.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cfd_dual.hpp"

namespace cfd {

constexpr int NP = 15;       // nodes per degree-4 triangle
constexpr int NFP = 5;       // nodes per face
constexpr int NSLOT = 15;    // 3 faces x 5 face nodes
constexpr int NV = 4;        // conserved variables

enum FluxId : int { FLUX_RUSANOV = 0, FLUX_ROE = 1, FLUX_HLLC = 2, FLUX_DYNAMIC = 3 };
enum LimiterId : int { LIM_MINMOD = 0, LIM_MC = 1, LIM_VAN_LEER = 2, LIM_SUPERBEE = 3, LIM_UNLIMITED = 4 };
// BC_EXACT: a prescribed exterior state per face node (rows nb + k of the
// boundary-state table for own face node k) and the limiter's ghost average
// at row nb + NFP.
enum BcKind : int { BC_INTERIOR = 0, BC_SLIP = 1, BC_FIXED = 2, BC_TRANSMISSIVE = 3, BC_EXACT = 4 };

constexpr double ROE_ENTROPY_FIX = 0.1;          // fluxes.py:21
constexpr double TINY = 2.2250738585072014e-308; // finfo(float64).tiny, fluxes.py:24-25

// 1/b from v_rcp_f64 refined by two Newton steps; a/b as a*(1/b) plus one
// residual correction; sqrt(x) from v_rsq_f64 refined by a coupled
// Newton step and a residual correction. Each is within a few ulps of the
// correctly rounded result for the positive, normal arguments the fluxes
// produce (checked against NumPy in the pointwise-flux tests). fsqrt(+0) is
// +0: the seed is the reciprocal square root of x plus the smallest
// subnormal, which is x itself for every x above 2^-1020 (the addend is below
// half an ulp) and keeps the seed finite at x = 0, so x * seed is 0 instead of
// 0 * inf = NaN; the Newton steps use x itself. Negative and NaN arguments
// still give NaN.
__device__ __forceinline__ double frcp(double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  e = fma(-b, r, 1.0);
  return fma(r, e, r);
}
__device__ __forceinline__ double fdiv(double a, double b) {
  const double r = frcp(b);
  const double q = a * r;
  return fma(fma(-b, q, a), r, q);
}
__device__ __forceinline__ double fsqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x + 4.9406564584124654e-324);
  double h = 0.5 * y;
  double g = x * y;
  const double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  return fma(fma(-g, g, x), h, g);
}
// Dual numbers keep exact arithmetic (the adjoint is not on the hot path).
template <int N> __device__ __forceinline__ ad::Dual<N> fdiv(const ad::Dual<N>& a, const ad::Dual<N>& b) { return a / b; }
template <int N> __device__ __forceinline__ ad::Dual<N> fdiv(double a, const ad::Dual<N>& b) { return a / b; }
template <int N> __device__ __forceinline__ ad::Dual<N> fdiv(const ad::Dual<N>& a, double b) { return a / b; }
template <int N> __device__ __forceinline__ ad::Dual<N> fsqrt(const ad::Dual<N>& a) { return sqrt(a); }
template <int N> __device__ __forceinline__ ad::Dual<N> frcp(const ad::Dual<N>& a) { return 1.0 / a; }

template <class T>
struct PrimT {
  T rho, u, v, p;
  T inv;  // 1 / rho
};
using Prim = PrimT<double>;

// conserved_to_primitive (euler.py:21-27).
template <class T>
__device__ __forceinline__ PrimT<T> to_prim(const T q[4], double gm1) {
  PrimT<T> w;
  w.rho = q[0];
  w.inv = frcp(q[0]);
  w.u = q[1] * w.inv;
  w.v = q[2] * w.inv;
  w.p = gm1 * (q[3] - 0.5 * (q[1] * w.u + q[2] * w.v));
  return w;
}

// pressure (euler.py:30-32) with the reference's rounding: contraction off.
__device__ __forceinline__ double pressure_of(const double q[4], double gm1) {
#pragma clang fp contract(off)
  const double u = q[1] / q[0];
  const double v = q[2] / q[0];
  return gm1 * (q[3] - 0.5 * (q[1] * u + q[2] * v));
}

// normal_flux (euler.py:49-61).
template <class T>
__device__ __forceinline__ void normal_flux(const T q[4], const PrimT<T>& w, double nx, double ny, T out[4]) {
  const T un = w.u * nx + w.v * ny;
  out[0] = w.rho * un;
  out[1] = q[1] * un + w.p * nx;
  out[2] = q[2] * un + w.p * ny;
  out[3] = (q[3] + w.p) * un;
}

template <class T>
__device__ __forceinline__ T safe_den(const T& x) {
  return x == 0.0 ? T(TINY) : x;
}

// Minus-side data computed once per face node and reused by every flux.
template <class T>
struct SideT {
  T q[4];
  PrimT<T> w;
  T c;     // sound speed
  T un;    // normal velocity
  T f[4];  // F(q) . n
};
using Side = SideT<double>;

template <class T>
__device__ __forceinline__ void make_side(SideT<T>& s, const T q[4], double nx, double ny, double gamma,
                                          double gm1) {
#pragma unroll
  for (int c = 0; c < 4; ++c) s.q[c] = q[c];
  s.w = to_prim(q, gm1);
  s.c = fsqrt(gamma * s.w.p * s.w.inv);
  s.un = s.w.u * nx + s.w.v * ny;
  normal_flux(q, s.w, nx, ny, s.f);
}

// rusanov (fluxes.py:28-37).
template <class T>
__device__ __forceinline__ void flux_rusanov(const SideT<T>& m, const SideT<T>& p, T out[4]) {
  const T sm = fabs(m.un) + m.c;
  const T sp = fabs(p.un) + p.c;
  const T alpha = fmax(sm, sp);
#pragma unroll
  for (int c = 0; c < 4; ++c) out[c] = 0.5 * (m.f[c] + p.f[c]) - 0.5 * alpha * (p.q[c] - m.q[c]);
}

template <class T>
struct RoeAvgT {
  T rho, u, v, h, c;
};

// _roe_average (fluxes.py:40-51).
template <class T>
__device__ __forceinline__ RoeAvgT<T> roe_average(const SideT<T>& m, const SideT<T>& p, double gm1) {
  RoeAvgT<T> a;
  const T hm = (m.q[3] + m.w.p) * m.w.inv;
  const T hp = (p.q[3] + p.w.p) * p.w.inv;
  const T wm = fsqrt(m.w.rho);
  const T wp = fsqrt(p.w.rho);
  const T inv = fdiv(1.0, wm + wp);
  a.u = (wm * m.w.u + wp * p.w.u) * inv;
  a.v = (wm * m.w.v + wp * p.w.v) * inv;
  a.h = (wm * hm + wp * hp) * inv;
  a.c = fsqrt(gm1 * (a.h - 0.5 * (a.u * a.u + a.v * a.v)));
  a.rho = wm * wp;
  return a;
}

// roe (fluxes.py:54-98).
template <class T>
__device__ __forceinline__ void flux_roe(const SideT<T>& m, const SideT<T>& p, double nx, double ny, double gm1,
                                         T out[4]) {
  const RoeAvgT<T> a = roe_average(m, p, gm1);
  const T un = a.u * nx + a.v * ny;
  const T ut = -a.u * ny + a.v * nx;
  const T d_rho = p.w.rho - m.w.rho;
  const T d_p = p.w.p - m.w.p;
  const T d_un = p.un - m.un;
  const T d_ut = (-p.w.u * ny + p.w.v * nx) - (-m.w.u * ny + m.w.v * nx);
  const T c2 = a.c * a.c;
  const T inv2c2 = fdiv(1.0, 2.0 * c2);
  const T a_minus = (d_p - a.rho * a.c * d_un) * inv2c2;
  const T a_entropy = d_rho - 2.0 * d_p * inv2c2;
  const T a_shear = a.rho * d_ut;
  const T a_plus = (d_p + a.rho * a.c * d_un) * inv2c2;
  const T delta = ROE_ENTROPY_FIX * (fabs(un) + a.c);
  auto fix = [delta](const T& lam) {
    const T mag = fabs(lam);
    return mag < delta ? fdiv(lam * lam + delta * delta, 2.0 * delta) : mag;
  };
  const T lm = fix(un - a.c) * a_minus;
  const T lp = fix(un + a.c) * a_plus;
  const T le = fabs(un) * a_entropy;
  const T ls = fabs(un) * a_shear;
  const T d0 = lm + le + lp;
  const T d1 = lm * (a.u - a.c * nx) + le * a.u + ls * (-ny) + lp * (a.u + a.c * nx);
  const T d2 = lm * (a.v - a.c * ny) + le * a.v + ls * nx + lp * (a.v + a.c * ny);
  const T d3 = lm * (a.h - a.c * un) + le * (0.5 * (a.u * a.u + a.v * a.v)) + ls * ut + lp * (a.h + a.c * un);
  out[0] = 0.5 * (m.f[0] + p.f[0]) - 0.5 * d0;
  out[1] = 0.5 * (m.f[1] + p.f[1]) - 0.5 * d1;
  out[2] = 0.5 * (m.f[2] + p.f[2]) - 0.5 * d2;
  out[3] = 0.5 * (m.f[3] + p.f[3]) - 0.5 * d3;
}

// hllc (fluxes.py:101-143). Only the selected branch's star state is built;
// the branch order and the >=/> tests are the reference's.
template <class T>
__device__ __forceinline__ void flux_hllc(const SideT<T>& m, const SideT<T>& p, double nx, double ny, double gm1,
                                          T out[4]) {
  const RoeAvgT<T> a = roe_average(m, p, gm1);
  const T un = a.u * nx + a.v * ny;
  const T s_m = fmin(m.un - m.c, un - a.c);
  const T s_p = fmax(p.un + p.c, un + a.c);
  const T mass_m = m.w.rho * (s_m - m.un);
  const T mass_p = p.w.rho * (s_p - p.un);
  const T s_star = fdiv(p.w.p - m.w.p + mass_m * m.un - mass_p * p.un, safe_den(mass_m - mass_p));
  if (s_m >= 0.0) {
#pragma unroll
    for (int c = 0; c < 4; ++c) out[c] = m.f[c];
    return;
  }
  const bool left = s_star >= 0.0;
  if (!left && !(s_p > 0.0)) {
#pragma unroll
    for (int c = 0; c < 4; ++c) out[c] = p.f[c];
    return;
  }
  const SideT<T>& k = left ? m : p;
  const T s_k = left ? s_m : s_p;
  const T mass_k = left ? mass_m : mass_p;
  // factor = mass_k / (S_k - S*); factor * p_k / mass_k = p_k / (S_k - S*),
  // so one reciprocal serves both terms of the star energy.
  const T rden = frcp(safe_den(s_k - s_star));
  const T factor = mass_k * rden;
  const T shift = s_star - k.un;
  T star[4];
  star[0] = factor;
  star[1] = factor * (k.w.u + shift * nx);
  star[2] = factor * (k.w.v + shift * ny);
  star[3] = factor * (k.q[3] * k.w.inv + shift * s_star) + shift * (k.w.p * rden);
#pragma unroll
  for (int c = 0; c < 4; ++c) out[c] = k.f[c] + s_k * (star[c] - k.q[c]);
}

template <int FLUX, class T>
__device__ __forceinline__ void interface_flux(const SideT<T>& m, const SideT<T>& p, double nx, double ny, double gm1,
                                               T out[4]) {
  if constexpr (FLUX == FLUX_RUSANOV) {
    flux_rusanov(m, p, out);
  } else if constexpr (FLUX == FLUX_ROE) {
    flux_roe(m, p, nx, ny, gm1, out);
  } else {
    flux_hllc(m, p, nx, ny, gm1, out);
  }
}

// Flux chosen at run time (batched solves; the id is uniform per workgroup).
template <class T>
__device__ __forceinline__ void interface_flux_rt(int id, const SideT<T>& m, const SideT<T>& p, double nx, double ny,
                                                  double gm1, T out[4]) {
  if (id == FLUX_RUSANOV) {
    flux_rusanov(m, p, out);
  } else if (id == FLUX_ROE) {
    flux_roe(m, p, nx, ny, gm1, out);
  } else {
    flux_hllc(m, p, nx, ny, gm1, out);
  }
}

// Limiter functions phi(r) (limiters.py:16-39), selected at run time.
template <class T>
__device__ __forceinline__ T limiter_phi(int id, const T& r) {
  switch (id) {
    case LIM_MINMOD:
      return fmax(0.0, fmin(1.0, r));
    case LIM_MC:
      return fmax(0.0, fmin(fmin(2.0 * r, 0.5 * (1.0 + r)), 2.0));
    case LIM_VAN_LEER:
      return (r + fabs(r)) / (1.0 + fabs(r));
    case LIM_SUPERBEE:
      return fmax(0.0, fmax(fmin(2.0 * r, 1.0), fmin(r, 2.0)));
    default:
      return T(1.0);
  }
}

// NaN-propagating min/max (jnp.min / jnp.max semantics).
__device__ __forceinline__ double nan_min(double a, double b) { return (a < b || a != a) ? a : b; }
__device__ __forceinline__ double nan_max(double a, double b) { return (a > b || a != a) ? a : b; }

}  // namespace cfd
