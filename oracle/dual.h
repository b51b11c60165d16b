// dual.h -- forward-mode dual numbers for the oracle's tangent-linear build.
//
// TEST INFRASTRUCTURE ONLY.  ostc_oracle.cpp compiled with -DORC_DUAL stores
// every state value as a Dual {value, tangent}; each operation applies the
// chain rule to the tangent, so one run of the oracle gives the exact
// directional derivative (JVP) of everything it computes, to round-off.
// Values follow the same operations in the same order as the double build.
#pragma once
#include <cmath>

namespace orc {

struct Dual {
  double v, d;
  Dual() : v(0), d(0) {}
  Dual(double x) : v(x), d(0) {}  // NOLINT: constants and inputs convert implicitly
  Dual(double x, double t) : v(x), d(t) {}
  Dual& operator+=(const Dual& b) { d += b.d; v += b.v; return *this; }
  Dual& operator-=(const Dual& b) { d -= b.d; v -= b.v; return *this; }
  Dual& operator*=(const Dual& b) { d = d * b.v + v * b.d; v *= b.v; return *this; }
  Dual& operator/=(const Dual& b) { const double q = v / b.v; d = (d - q * b.d) / b.v; v = q; return *this; }
};

inline Dual operator+(const Dual& a) { return a; }
inline Dual operator-(const Dual& a) { return Dual(-a.v, -a.d); }
inline Dual operator+(const Dual& a, const Dual& b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(const Dual& a, const Dual& b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator*(const Dual& a, const Dual& b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(const Dual& a, const Dual& b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
inline Dual operator+(const Dual& a, double b) { return Dual(a.v + b, a.d); }
inline Dual operator+(double a, const Dual& b) { return Dual(a + b.v, b.d); }
inline Dual operator-(const Dual& a, double b) { return Dual(a.v - b, a.d); }
inline Dual operator-(double a, const Dual& b) { return Dual(a - b.v, -b.d); }
inline Dual operator*(const Dual& a, double b) { return Dual(a.v * b, a.d * b); }
inline Dual operator*(double a, const Dual& b) { return Dual(a * b.v, a * b.d); }
inline Dual operator/(const Dual& a, double b) { return Dual(a.v / b, a.d / b); }
inline Dual operator/(double a, const Dual& b) {
  const double q = a / b.v;
  return Dual(q, -q * b.d / b.v);
}
inline bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
inline bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
inline bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
inline bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }

inline Dual sqrt(const Dual& a) {
  const double r = std::sqrt(a.v);
  return Dual(r, a.d / (2 * r));
}
// d|x|/dx = sign(x), taken as +1 at x = 0
inline Dual abs(const Dual& a) { return a.v < 0 ? -a : a; }
inline Dual pow(const Dual& a, double e) {
  const double r = std::pow(a.v, e);
  return Dual(r, e * std::pow(a.v, e - 1) * a.d);
}
inline Dual tanh(const Dual& a) {
  const double t = std::tanh(a.v);
  return Dual(t, (1 - t * t) * a.d);
}

inline double value_of(const Dual& a) { return a.v; }
inline double tangent_of(const Dual& a) { return a.d; }

}  // namespace orc
