// Test infrastructure only: a forward-mode dual number for the CPU oracle.
//
// Dual<N> carries a value and N directional derivatives.  The oracle's zone
// routines are templated on their scalar type, so running them on Dual<N>
// with the inputs seeded along N directions gives the exact (to rounding)
// directional derivatives of every output: the reference gradient the GPU
// adjoint is checked against.  Arithmetic follows the usual rules
// (d(ab) = a db + b da, d(a/b) = (da - (a/b) db)/b, d sqrt(a) = da/(2 sqrt a)).
#pragma once
#include <cmath>

namespace npb {

template <int N>
struct Dual {
  double v = 0.0;
  double d[N] = {};
  Dual() = default;
  Dual(double x) : v(x) {}  // constants mix freely with duals
  Dual(int x) : v(double(x)) {}
};

template <int N>
inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v + b.v;
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int N>
inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v - b.v;
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int N>
inline Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r;
  r.v = -a.v;
  for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
  return r;
}
template <int N>
inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v * b.v;
  for (int k = 0; k < N; ++k) r.d[k] = a.v * b.d[k] + a.d[k] * b.v;
  return r;
}
template <int N>
inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v / b.v;
  for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
  return r;
}
template <int N>
inline Dual<N> operator*(double a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a * b.v;
  for (int k = 0; k < N; ++k) r.d[k] = a * b.d[k];
  return r;
}
template <int N>
inline Dual<N> operator*(const Dual<N>& a, double b) { return b * a; }
template <int N>
inline Dual<N> operator/(const Dual<N>& a, double b) {
  Dual<N> r;
  r.v = a.v / b;
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / b;
  return r;
}
template <int N>
inline Dual<N> operator/(double a, const Dual<N>& b) { return Dual<N>(a) / b; }
template <int N>
inline Dual<N> operator+(double a, const Dual<N>& b) { return Dual<N>(a) + b; }
template <int N>
inline Dual<N> operator+(const Dual<N>& a, double b) { return a + Dual<N>(b); }
template <int N>
inline Dual<N> operator-(double a, const Dual<N>& b) { return Dual<N>(a) - b; }
template <int N>
inline Dual<N> operator-(const Dual<N>& a, double b) { return a - Dual<N>(b); }
template <int N>
inline Dual<N>& operator+=(Dual<N>& a, const Dual<N>& b) { return a = a + b; }
template <int N>
inline Dual<N>& operator-=(Dual<N>& a, const Dual<N>& b) { return a = a - b; }
template <int N>
inline Dual<N>& operator*=(Dual<N>& a, const Dual<N>& b) { return a = a * b; }
template <int N>
inline Dual<N>& operator/=(Dual<N>& a, const Dual<N>& b) { return a = a / b; }
template <int N>
inline Dual<N>& operator*=(Dual<N>& a, double b) { return a = a * b; }
template <int N>
inline Dual<N>& operator+=(Dual<N>& a, double b) { return a = a + b; }
template <int N>
inline Dual<N>& operator-=(Dual<N>& a, double b) { return a = a - b; }
template <int N>
inline Dual<N>& operator/=(Dual<N>& a, double b) { return a = a / b; }
template <int N>
inline Dual<N> sqrt(const Dual<N>& a) {
  Dual<N> r;
  r.v = std::sqrt(a.v);
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / (2.0 * r.v);
  return r;
}
// max: the larger operand with its derivatives (the first one on a tie), the
// derivative of max along the branch that attains it.
template <int N>
inline Dual<N> fmax(const Dual<N>& a, const Dual<N>& b) { return a.v >= b.v ? a : b; }
template <int N>
inline Dual<N> fmax(const Dual<N>& a, double b) { return a.v >= b ? a : Dual<N>(b); }
template <int N>
inline Dual<N> fmax(double a, const Dual<N>& b) { return a >= b.v ? Dual<N>(a) : b; }

}  // namespace npb
