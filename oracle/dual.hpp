// dual.hpp -- forward-mode operator-overloading AD scalar for the oracle.
// TEST INFRASTRUCTURE ONLY.  orc::Dual carries (value, tangent); instantiating
// the oracle templates with it yields the directional derivative of the
// reference step.  Branches, comparisons and min/max select on the value with
// the same tie rules as the double code (libm fmin/fmax return the second
// operand on ties, std::min/std::max the first), so the tangent is that of the
// branch the primal computation takes.
#pragma once

#include <cmath>

namespace orc {

struct Dual {
    double v = 0.0, d = 0.0;
    Dual() = default;
    Dual(double x) : v(x), d(0.0) {}  // NOLINT: implicit by design
    Dual(double x, double dx) : v(x), d(dx) {}
};

inline double val(const Dual& x) { return x.v; }
inline bool fin(const Dual& x) { return std::isfinite(x.v); }

inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) {
    const double q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
inline Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
inline Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
inline Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
inline Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
inline Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
inline Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
inline Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
inline Dual operator/(double a, Dual b) {
    const double q = a / b.v;
    return {q, -q * b.d / b.v};
}

inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
inline bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
inline bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
inline bool operator<(Dual a, double b) { return a.v < b; }
inline bool operator>(Dual a, double b) { return a.v > b; }
inline bool operator<=(Dual a, double b) { return a.v <= b; }
inline bool operator>=(Dual a, double b) { return a.v >= b; }
inline bool operator<(double a, Dual b) { return a < b.v; }
inline bool operator>(double a, Dual b) { return a > b.v; }
inline bool operator<=(double a, Dual b) { return a <= b.v; }
inline bool operator>=(double a, Dual b) { return a >= b.v; }

inline Dual sqrt(Dual a) {
    const double s = std::sqrt(a.v);
    return {s, a.d / (2.0 * s)};
}
inline Dual fabs(Dual a) { return a.v < 0.0 ? Dual{-a.v, -a.d} : a; }
inline Dual pow(Dual a, double e) {
    const double p = std::pow(a.v, e);
    return {p, e * std::pow(a.v, e - 1.0) * a.d};
}

inline Dual lmin(Dual a, Dual b) { return (a.v < b.v) ? a : b; }   // libm fmin: second on ties
inline Dual lmax(Dual a, Dual b) { return (a.v > b.v) ? a : b; }   // libm fmax: second on ties
inline Dual smin(Dual a, Dual b) { return (b.v < a.v) ? b : a; }   // std::min: first on ties
inline Dual smax(Dual a, Dual b) { return (a.v < b.v) ? b : a; }   // std::max: first on ties

}  // namespace orc
