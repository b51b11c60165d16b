// count.h -- an operation-counting scalar for the oracle's FLOP-count build.
//
// TEST INFRASTRUCTURE ONLY.  ostc_oracle.cpp compiled with -DORC_COUNT stores
// every state and geometry value as a Counted {value}; each arithmetic
// operation on one increments a counter and rounds exactly as double does,
// so the count build computes the same values as the double build and, in
// addition, how many additions/subtractions, multiplications, divisions and
// square roots it took (each one floating-point operation), and how many
// calls of pow/exp/log/tanh (counted apart, as "transcendental calls").
// Negation, |x| and comparisons are not counted.  The build has no OpenMP,
// so the counters are plain globals.
#pragma once
#include <cmath>

namespace orc {

struct OpCounts {
  unsigned long long add = 0, mul = 0, div = 0, sqrt = 0, trans = 0;
};
inline OpCounts g_ops;

struct Counted {
  double v;
  Counted() : v(0) {}
  Counted(double x) : v(x) {}  // NOLINT: constants and inputs convert implicitly
  Counted& operator+=(const Counted& b) { ++g_ops.add; v += b.v; return *this; }
  Counted& operator-=(const Counted& b) { ++g_ops.add; v -= b.v; return *this; }
  Counted& operator*=(const Counted& b) { ++g_ops.mul; v *= b.v; return *this; }
  Counted& operator/=(const Counted& b) { ++g_ops.div; v /= b.v; return *this; }
};

inline Counted operator+(const Counted& a) { return a; }
inline Counted operator-(const Counted& a) { return Counted(-a.v); }
inline Counted operator+(const Counted& a, const Counted& b) { ++g_ops.add; return Counted(a.v + b.v); }
inline Counted operator-(const Counted& a, const Counted& b) { ++g_ops.add; return Counted(a.v - b.v); }
inline Counted operator*(const Counted& a, const Counted& b) { ++g_ops.mul; return Counted(a.v * b.v); }
inline Counted operator/(const Counted& a, const Counted& b) { ++g_ops.div; return Counted(a.v / b.v); }
inline Counted operator+(const Counted& a, double b) { ++g_ops.add; return Counted(a.v + b); }
inline Counted operator+(double a, const Counted& b) { ++g_ops.add; return Counted(a + b.v); }
inline Counted operator-(const Counted& a, double b) { ++g_ops.add; return Counted(a.v - b); }
inline Counted operator-(double a, const Counted& b) { ++g_ops.add; return Counted(a - b.v); }
inline Counted operator*(const Counted& a, double b) { ++g_ops.mul; return Counted(a.v * b); }
inline Counted operator*(double a, const Counted& b) { ++g_ops.mul; return Counted(a * b.v); }
inline Counted operator/(const Counted& a, double b) { ++g_ops.div; return Counted(a.v / b); }
inline Counted operator/(double a, const Counted& b) { ++g_ops.div; return Counted(a / b.v); }
inline bool operator<(const Counted& a, const Counted& b) { return a.v < b.v; }
inline bool operator>(const Counted& a, const Counted& b) { return a.v > b.v; }
inline bool operator<=(const Counted& a, const Counted& b) { return a.v <= b.v; }
inline bool operator>=(const Counted& a, const Counted& b) { return a.v >= b.v; }
inline bool operator==(const Counted& a, const Counted& b) { return a.v == b.v; }
inline bool operator!=(const Counted& a, const Counted& b) { return a.v != b.v; }

inline Counted sqrt(const Counted& a) { ++g_ops.sqrt; return Counted(std::sqrt(a.v)); }
inline Counted abs(const Counted& a) { return Counted(std::fabs(a.v)); }
inline Counted fabs(const Counted& a) { return Counted(std::fabs(a.v)); }
inline Counted pow(const Counted& a, double e) { ++g_ops.trans; return Counted(std::pow(a.v, e)); }
inline Counted tanh(const Counted& a) { ++g_ops.trans; return Counted(std::tanh(a.v)); }
inline Counted exp(const Counted& a) { ++g_ops.trans; return Counted(std::exp(a.v)); }
inline Counted min(const Counted& a, const Counted& b) { return b.v < a.v ? b : a; }
inline Counted max(const Counted& a, const Counted& b) { return a.v < b.v ? b : a; }

inline double value_of(const Counted& a) { return a.v; }

}  // namespace orc
