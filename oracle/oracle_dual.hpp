// oracle_dual.hpp -- forward-mode dual numbers for the FP64 oracle.
//
// TEST INFRASTRUCTURE ONLY.  Dual<N> carries a value and N tangents; every
// operation computes its value exactly as the double operation would, so a
// Dual run reproduces the double run bit for bit in its value part.
#pragma once

#include <cmath>

namespace oracle {

template <int N> struct Dual {
  double v;
  double d[N];
  Dual() : v(0.0) { for (int i = 0; i < N; ++i) d[i] = 0.0; }
  Dual(double x) : v(x) { for (int i = 0; i < N; ++i) d[i] = 0.0; }  // NOLINT: implicit by design
  static Dual seed(double x, int k) { Dual r(x); r.d[k] = 1.0; return r; }
};

template <int N> inline double val(const Dual<N> &x) { return x.v; }

template <int N> inline Dual<N> operator-(const Dual<N> &a) {
  Dual<N> r; r.v = -a.v; for (int i = 0; i < N; ++i) r.d[i] = -a.d[i]; return r;
}
template <int N> inline Dual<N> operator+(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r; r.v = a.v + b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i]; return r;
}
template <int N> inline Dual<N> operator-(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r; r.v = a.v - b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i]; return r;
}
template <int N> inline Dual<N> operator*(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r; r.v = a.v * b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r;
}
template <int N> inline Dual<N> operator/(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r; r.v = a.v / b.v;
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
  return r;
}
template <int N> inline Dual<N> operator+(double a, const Dual<N> &b) { return Dual<N>(a) + b; }
template <int N> inline Dual<N> operator+(const Dual<N> &a, double b) { return a + Dual<N>(b); }
template <int N> inline Dual<N> operator-(double a, const Dual<N> &b) { return Dual<N>(a) - b; }
template <int N> inline Dual<N> operator-(const Dual<N> &a, double b) { return a - Dual<N>(b); }
template <int N> inline Dual<N> operator*(double a, const Dual<N> &b) {
  Dual<N> r; r.v = a * b.v; for (int i = 0; i < N; ++i) r.d[i] = a * b.d[i]; return r;
}
template <int N> inline Dual<N> operator*(const Dual<N> &a, double b) {
  Dual<N> r; r.v = a.v * b; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b; return r;
}
template <int N> inline Dual<N> operator/(const Dual<N> &a, double b) {
  Dual<N> r; r.v = a.v / b; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / b; return r;
}
template <int N> inline Dual<N> operator/(double a, const Dual<N> &b) { return Dual<N>(a) / b; }

template <int N> inline Dual<N> sin(const Dual<N> &a) {
  Dual<N> r; r.v = std::sin(a.v); double c = std::cos(a.v);
  for (int i = 0; i < N; ++i) r.d[i] = c * a.d[i];
  return r;
}
template <int N> inline Dual<N> cos(const Dual<N> &a) {
  Dual<N> r; r.v = std::cos(a.v); double s = -std::sin(a.v);
  for (int i = 0; i < N; ++i) r.d[i] = s * a.d[i];
  return r;
}
template <int N> inline Dual<N> sqrt(const Dual<N> &a) {
  Dual<N> r; r.v = std::sqrt(a.v); double k = 0.5 / r.v;
  for (int i = 0; i < N; ++i) r.d[i] = k * a.d[i];
  return r;
}
template <int N> inline Dual<N> exp(const Dual<N> &a) {
  Dual<N> r; r.v = std::exp(a.v);
  for (int i = 0; i < N; ++i) r.d[i] = r.v * a.d[i];
  return r;
}
template <int N> inline Dual<N> tan(const Dual<N> &a) {
  Dual<N> r; r.v = std::tan(a.v); double k = 1.0 + r.v * r.v;
  for (int i = 0; i < N; ++i) r.d[i] = k * a.d[i];
  return r;
}
template <int N> inline Dual<N> fabs(const Dual<N> &a) { return a.v < 0.0 ? -a : a; }
// x^c with constant exponent.
template <int N> inline Dual<N> pow(const Dual<N> &a, double c) {
  Dual<N> r; r.v = std::pow(a.v, c); double k = c * std::pow(a.v, c - 1.0);
  for (int i = 0; i < N; ++i) r.d[i] = k * a.d[i];
  return r;
}
// x^y with both operands carrying tangents (x > 0).
template <int N> inline Dual<N> pow(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r; r.v = std::pow(a.v, b.v);
  double lx = std::log(a.v);
  for (int i = 0; i < N; ++i) r.d[i] = r.v * (b.d[i] * lx + b.v * a.d[i] / a.v);
  return r;
}
template <int N> inline Dual<N> rem_euclid_t(const Dual<N> &x, double m) {
  Dual<N> r = x; r.v = rem_euclid(x.v, m); return r;
}

}  // namespace oracle
