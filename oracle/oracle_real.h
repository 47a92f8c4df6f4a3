// oracle_real.h -- the oracle's arithmetic type.
//
// TEST INFRASTRUCTURE ONLY.  The oracle is compiled twice from the same
// source: with real_t = double (the FP64 checker, liboracle.so) and, with
// ORC_TANGENT defined, with real_t = Dual (liboracle_tlm.so), a forward-mode
// number carrying a value and one directional derivative.  The second build
// is the FP64 tangent-linear model of the oracle's step: seeding the
// derivative parts of the state with a direction d and stepping with frozen
// step sizes yields J d exactly (to rounding), which the adjoint's J^T lambda
// must match in <lambda, J d> = <J^T lambda, d>.  Branches (limiters, upwind
// choices, floors) follow the values, so the derivative is the one of the
// branch the forward arithmetic takes, as in the adjoint.
#ifndef PENGUIN_ORACLE_REAL_H
#define PENGUIN_ORACLE_REAL_H

#include <cmath>

#ifdef ORC_TANGENT

struct Dual {
    double v, d;   // value, derivative along the seeded direction
    Dual() : v(0.0), d(0.0) {}
    Dual(double x) : v(x), d(0.0) {}
    Dual(double x, double dx) : v(x), d(dx) {}
    Dual& operator+=(const Dual& b) { v += b.v; d += b.d; return *this; }
    Dual& operator-=(const Dual& b) { v -= b.v; d -= b.d; return *this; }
    Dual& operator*=(const Dual& b) { d = d * b.v + v * b.d; v *= b.v; return *this; }
    Dual& operator/=(const Dual& b) { d = (d * b.v - v * b.d) / (b.v * b.v); v /= b.v; return *this; }
};
inline Dual operator-(const Dual& a) { return Dual(-a.v, -a.d); }
inline Dual operator+(Dual a, const Dual& b) { return a += b; }
inline Dual operator-(Dual a, const Dual& b) { return a -= b; }
inline Dual operator*(Dual a, const Dual& b) { return a *= b; }
inline Dual operator/(Dual a, const Dual& b) { return a /= b; }
inline Dual operator+(Dual a, double b) { return a += Dual(b); }
inline Dual operator-(Dual a, double b) { return a -= Dual(b); }
inline Dual operator*(const Dual& a, double b) { return Dual(a.v * b, a.d * b); }
inline Dual operator/(const Dual& a, double b) { return Dual(a.v / b, a.d / b); }
inline Dual operator+(double a, const Dual& b) { return Dual(a) + b; }
inline Dual operator-(double a, const Dual& b) { return Dual(a) - b; }
inline Dual operator*(double a, const Dual& b) { return Dual(a * b.v, a * b.d); }
inline Dual operator/(double a, const Dual& b) { return Dual(a) / b; }
#define ORC_CMP(op)                                                                   \
    inline bool operator op(const Dual& a, const Dual& b) { return a.v op b.v; }      \
    inline bool operator op(const Dual& a, double b) { return a.v op b; }             \
    inline bool operator op(double a, const Dual& b) { return a op b.v; }
ORC_CMP(<)
ORC_CMP(>)
ORC_CMP(<=)
ORC_CMP(>=)
ORC_CMP(==)
ORC_CMP(!=)
#undef ORC_CMP

inline double val(const Dual& x) { return x.v; }
inline double tan_of(const Dual& x) { return x.d; }
inline Dual msqrt(const Dual& x) { const double s = std::sqrt(x.v); return Dual(s, s > 0.0 ? 0.5 * x.d / s : 0.0); }
inline Dual mfabs(const Dual& x) { return x.v < 0.0 ? -x : x; }
inline Dual mfmax(const Dual& a, const Dual& b) { return a.v >= b.v ? a : b; }   // std::fmax picks a on ties
inline Dual mfmin(const Dual& a, const Dual& b) { return a.v <= b.v ? a : b; }
inline Dual mexp(const Dual& x) { const double e = std::exp(x.v); return Dual(e, e * x.d); }
inline Dual mlog(const Dual& x) { return Dual(std::log(x.v), x.d / x.v); }
inline Dual mlogf(const Dual& x) { return Dual((double)(float)std::log((double)(float)x.v), x.d / x.v); }
inline Dual msin(const Dual& x) { return Dual(std::sin(x.v), std::cos(x.v) * x.d); }
inline Dual mcos(const Dual& x) { return Dual(std::cos(x.v), -std::sin(x.v) * x.d); }
inline Dual mtan(const Dual& x) { const double c = std::cos(x.v); return Dual(std::tan(x.v), x.d / (c * c)); }
inline Dual matan(const Dual& x) { return Dual(std::atan(x.v), x.d / (1.0 + x.v * x.v)); }
inline Dual mpow(const Dual& x, double e) {
    const double p = std::pow(x.v, e);
    return Dual(p, e == 0.0 ? 0.0 : e * std::pow(x.v, e - 1.0) * x.d);
}
inline Dual mpow(const Dual& x, const Dual& e) {
    const double p = std::pow(x.v, e.v);
    const double dx = e.v == 0.0 ? 0.0 : e.v * std::pow(x.v, e.v - 1.0) * x.d;
    const double de = e.d != 0.0 ? p * std::log(x.v) * e.d : 0.0;
    return Dual(p, dx + de);
}
inline Dual mfmod(const Dual& x, const Dual& m) {
    const double q = std::trunc(x.v / m.v);
    return Dual(std::fmod(x.v, m.v), x.d - q * m.d);
}
inline Dual mfloor(const Dual& x) { return Dual(std::floor(x.v), 0.0); }
inline Dual mcopysign(const Dual& a, const Dual& b) {
    const bool flip = std::signbit(a.v) != std::signbit(b.v);
    return flip ? -a : a;
}
inline bool msignbit(const Dual& x) { return std::signbit(x.v); }
typedef Dual real_t;

#else

typedef double real_t;

#endif

inline double val(double x) { return x; }
inline double tan_of(double) { return 0.0; }
inline double msqrt(double x) { return std::sqrt(x); }
inline double mfabs(double x) { return std::fabs(x); }
inline double mfmax(double a, double b) { return std::fmax(a, b); }
inline double mfmin(double a, double b) { return std::fmin(a, b); }
inline double mexp(double x) { return std::exp(x); }
inline double mlog(double x) { return std::log(x); }
// the single-precision logarithm of a float argument, correctly rounded
// (the double logarithm of the float, rounded to float)
inline double mlogf(double x) { return (double)(float)std::log((double)(float)x); }
inline double msin(double x) { return std::sin(x); }
inline double mcos(double x) { return std::cos(x); }
inline double mtan(double x) { return std::tan(x); }
inline double matan(double x) { return std::atan(x); }
inline double mpow(double x, double e) { return std::pow(x, e); }
inline double mfmod(double x, double m) { return std::fmod(x, m); }
inline double mfloor(double x) { return std::floor(x); }
inline double mcopysign(double a, double b) { return std::copysign(a, b); }
inline bool msignbit(double x) { return std::signbit(x); }

#endif
