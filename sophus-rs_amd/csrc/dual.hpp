// Forward-mode dual numbers for host and device code: a value and N partial derivatives.
// The Lie-group templates of lie.hpp instantiated with Dual<N> give exact Jacobians of the
// same code paths (branches decided on the values), the way the reference differentiates
// its cost terms (DualScalar / DualVector, crates/sophus_autodiff/src/dual/).  Include after
// lie.hpp: the math functions below are found by argument-dependent lookup.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "lie.hpp"

namespace sgpu {
namespace dual {

template <int N>
struct Dual {
  double re;
  double d[N];
  SGPU_HD Dual() : re(0.0) {
    for (int k = 0; k < N; ++k) d[k] = 0.0;
  }
  SGPU_HD Dual(double v) : re(v) {  // NOLINT: a constant converts implicitly
    for (int k = 0; k < N; ++k) d[k] = 0.0;
  }
  // the k-th input variable at value v
  static SGPU_HD Dual var(double v, int k) {
    Dual r(v);
    r.d[k] = 1.0;
    return r;
  }
};

template <int N>
SGPU_HD Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.re + b.re);
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int N>
SGPU_HD Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.re - b.re);
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int N>
SGPU_HD Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r(-a.re);
  for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
  return r;
}
template <int N>
SGPU_HD Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.re * b.re);
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.re + a.re * b.d[k];
  return r;
}
template <int N>
SGPU_HD Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  const double inv = 1.0 / b.re;
  Dual<N> r(a.re * inv);
  for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.re * b.d[k]) * inv;
  return r;
}
template <int N>
SGPU_HD Dual<N>& operator+=(Dual<N>& a, const Dual<N>& b) {
  a = a + b;
  return a;
}
template <int N>
SGPU_HD Dual<N>& operator-=(Dual<N>& a, const Dual<N>& b) {
  a = a - b;
  return a;
}
template <int N>
SGPU_HD bool operator<=(const Dual<N>& a, const Dual<N>& b) { return a.re <= b.re; }
template <int N>
SGPU_HD bool operator<(const Dual<N>& a, const Dual<N>& b) { return a.re < b.re; }
template <int N>
SGPU_HD bool operator>(const Dual<N>& a, const Dual<N>& b) { return a.re > b.re; }
template <int N>
SGPU_HD bool operator>=(const Dual<N>& a, const Dual<N>& b) { return a.re >= b.re; }

// f(a) with f(a.re) = v and f'(a.re) = dv
template <int N>
SGPU_HD Dual<N> chain(const Dual<N>& a, double v, double dv) {
  Dual<N> r(v);
  for (int k = 0; k < N; ++k) r.d[k] = dv * a.d[k];
  return r;
}
template <int N>
SGPU_HD Dual<N> sqrt(const Dual<N>& a) {
  const double s = ::sqrt(a.re);
  return chain(a, s, 0.5 / s);
}
template <int N>
SGPU_HD Dual<N> sin(const Dual<N>& a) { return chain(a, ::sin(a.re), ::cos(a.re)); }
template <int N>
SGPU_HD Dual<N> cos(const Dual<N>& a) { return chain(a, ::cos(a.re), -::sin(a.re)); }
template <int N>
SGPU_HD Dual<N> atan2(const Dual<N>& y, const Dual<N>& x) {
  const double den = 1.0 / (x.re * x.re + y.re * y.re);
  Dual<N> r(::atan2(y.re, x.re));
  for (int k = 0; k < N; ++k) r.d[k] = (x.re * y.d[k] - y.re * x.d[k]) * den;
  return r;
}

}  // namespace dual
}  // namespace sgpu
