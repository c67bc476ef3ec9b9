// mp_dual.h -- first-order dual numbers for the adjoint's local Jacobians.
//
// The point-wise process code (mp_process.h) is generic in its rate type R
// and its wide (state) type W = wide_t<R>.  The forward kernels use
// R = float or double with W = double.  The adjoint instantiates it with
// R = W = Dual: a value and one tangent.  Seeding the tangent of one input
// and running a process once gives that input's column of the process's
// Jacobian; every comparison acts on the value, so the derivative follows
// the branch the forward takes at the point.
#pragma once

#include "mp_common.h"

namespace shield {

struct Dual {
  double v, d;
  Dual() = default;
  SH_FN constexpr Dual(double x) : v(x), d(0.0) {}
  SH_FN constexpr Dual(double x, double dx) : v(x), d(dx) {}
};

SH_FN Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
SH_FN Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
SH_FN Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
SH_FN Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
SH_FN Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
SH_FN Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
SH_FN Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
SH_FN Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
SH_FN Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
SH_FN Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
SH_FN Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
SH_FN Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
SH_FN Dual operator/(double a, Dual b) {
  const double q = a / b.v;
  return Dual(q, -q * b.d / b.v);
}
SH_FN Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
SH_FN Dual& operator-=(Dual& a, Dual b) { return a = a - b; }
SH_FN Dual& operator*=(Dual& a, Dual b) { return a = a * b; }

#define SH_DUAL_CMP(op)                                                                    \
  SH_FN bool operator op(Dual a, Dual b) { return a.v op b.v; }       \
  SH_FN bool operator op(Dual a, double b) { return a.v op b; }       \
  SH_FN bool operator op(double a, Dual b) { return a op b.v; }
SH_DUAL_CMP(<)
SH_DUAL_CMP(>)
SH_DUAL_CMP(<=)
SH_DUAL_CMP(>=)
SH_DUAL_CMP(==)
#undef SH_DUAL_CMP

SH_FN Dual mexp(Dual x) {
  const double e = exp(x.v);
  return Dual(e, e * x.d);
}
SH_FN Dual mlog(Dual x) { return Dual(log(x.v), x.d / x.v); }
SH_FN Dual msqrt(Dual x) {
  const double r = sqrt(x.v);
  return Dual(r, 0.5 * x.d / r);
}
SH_FN Dual mabs(Dual x) { return x.v < 0.0 ? -x : x; }
// min / max select an operand by value (the reference's a < b ? a : b), so
// the tangent is the selected operand's
SH_FN Dual mmin(Dual a, Dual b) { return a.v < b.v ? a : b; }
SH_FN Dual mmax(Dual a, Dual b) { return a.v > b.v ? a : b; }
SH_FN Dual mmin(Dual a, double b) { return mmin(a, Dual(b)); }
SH_FN Dual mmin(double a, Dual b) { return mmin(Dual(a), b); }
SH_FN Dual mmax(Dual a, double b) { return mmax(a, Dual(b)); }
SH_FN Dual mmax(double a, Dual b) { return mmax(Dual(a), b); }
// positive difference (microphysics_funcs.py:45-49)
SH_FN Dual dim(Dual x, Dual y) {
  const Dual d = x - y;
  return d.v > 0.0 ? d : Dual(0.0);
}
SH_FN Dual dim(Dual x, double y) { return dim(x, Dual(y)); }
SH_FN Dual dim(double x, Dual y) { return dim(Dual(x), y); }

template <>
SH_FN Dual pw_el<Dual>(Dual x, double p) { return mexp(p * mlog(x)); }
template <>
SH_FN Dual pw<Dual>(Dual x, double p) {
  const double y = pow(x.v, p);
  return Dual(y, p * y / x.v * x.d);
}

}  // namespace shield
