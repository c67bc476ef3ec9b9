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
  __device__ __forceinline__ constexpr Dual(double x) : v(x), d(0.0) {}
  __device__ __forceinline__ constexpr Dual(double x, double dx) : v(x), d(dx) {}
};

__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
__device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
__device__ __forceinline__ Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
__device__ __forceinline__ Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
__device__ __forceinline__ Dual operator/(double a, Dual b) {
  const double q = a / b.v;
  return Dual(q, -q * b.d / b.v);
}
__device__ __forceinline__ Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
__device__ __forceinline__ Dual& operator-=(Dual& a, Dual b) { return a = a - b; }
__device__ __forceinline__ Dual& operator*=(Dual& a, Dual b) { return a = a * b; }

#define SH_DUAL_CMP(op)                                                                    \
  __device__ __forceinline__ bool operator op(Dual a, Dual b) { return a.v op b.v; }       \
  __device__ __forceinline__ bool operator op(Dual a, double b) { return a.v op b; }       \
  __device__ __forceinline__ bool operator op(double a, Dual b) { return a op b.v; }
SH_DUAL_CMP(<)
SH_DUAL_CMP(>)
SH_DUAL_CMP(<=)
SH_DUAL_CMP(>=)
SH_DUAL_CMP(==)
#undef SH_DUAL_CMP

__device__ __forceinline__ Dual mexp(Dual x) {
  const double e = exp(x.v);
  return Dual(e, e * x.d);
}
__device__ __forceinline__ Dual mlog(Dual x) { return Dual(log(x.v), x.d / x.v); }
__device__ __forceinline__ Dual msqrt(Dual x) {
  const double r = sqrt(x.v);
  return Dual(r, 0.5 * x.d / r);
}
__device__ __forceinline__ Dual mabs(Dual x) { return x.v < 0.0 ? -x : x; }
// min / max select an operand by value (the reference's a < b ? a : b), so
// the tangent is the selected operand's
__device__ __forceinline__ Dual mmin(Dual a, Dual b) { return a.v < b.v ? a : b; }
__device__ __forceinline__ Dual mmax(Dual a, Dual b) { return a.v > b.v ? a : b; }
__device__ __forceinline__ Dual mmin(Dual a, double b) { return mmin(a, Dual(b)); }
__device__ __forceinline__ Dual mmin(double a, Dual b) { return mmin(Dual(a), b); }
__device__ __forceinline__ Dual mmax(Dual a, double b) { return mmax(a, Dual(b)); }
__device__ __forceinline__ Dual mmax(double a, Dual b) { return mmax(Dual(a), b); }
// positive difference (microphysics_funcs.py:45-49)
__device__ __forceinline__ Dual dim(Dual x, Dual y) {
  const Dual d = x - y;
  return d.v > 0.0 ? d : Dual(0.0);
}
__device__ __forceinline__ Dual dim(Dual x, double y) { return dim(x, Dual(y)); }
__device__ __forceinline__ Dual dim(double x, Dual y) { return dim(Dual(x), y); }

template <>
__device__ __forceinline__ Dual pw_el<Dual>(Dual x, double p) { return mexp(p * mlog(x)); }
template <>
__device__ __forceinline__ Dual pw<Dual>(Dual x, double p) {
  const double y = pow(x.v, p);
  return Dual(y, p * y / x.v * x.d);
}

}  // namespace shield
