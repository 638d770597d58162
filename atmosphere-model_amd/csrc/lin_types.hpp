// lin_types.hpp -- the scalar type of the linearised dynamics kernels.
//
// Jac<K> carries a value and K partial derivatives.  The per-point bodies
// of the linearised step (lin_ops.hpp) are templates on their scalar type:
//   double   the forward model, the reference's operations in order;
//   Jac<1>   the tangent-linear model: the derivative slot holds the
//            tangent of the value, seeded from the tangent of each input;
//   Jac<K>   the adjoint: each of the body's K reads seeds its own slot, so
//            the written value carries its partial derivative with respect
//            to every value read; the adjoint kernel scatters the output's
//            adjoint times each partial back to the read positions.
// Comparisons, min/max and sign selections act on the value, so limiter
// branches and upwind switches are frozen at the trajectory: the
// linearisation is that of the branch the forward model takes.
#pragma once

#include "atm_common.hpp"

// Every branch decision of the linearised program passes through
// LIN_DECIDE: the value comparisons below, the Fortran MAX, MIN, SIGN and
// ABS selections, and the remap's overlap and clamp tests.  It is the
// decision itself unless a build defines it first: the host build of the
// program records the decisions of one run and replays them in another
// (its frozen-branch mode), so that finite differences differentiate the
// branches the trajectory took.
#ifndef LIN_DECIDE
#define LIN_DECIDE(b) (b)
#endif

template <int K>
struct Jac {
    double v;
    double d[K];
    __host__ __device__ Jac() : v(0.) {
        for (int i = 0; i < K; i++) d[i] = 0.;
    }
    __host__ __device__ Jac(double x) : v(x) {  // a constant
        for (int i = 0; i < K; i++) d[i] = 0.;
    }
    // an input: value x with derivative 1 in slot s
    __host__ __device__ static Jac seed(double x, int s) {
        Jac r(x);
        r.d[s] = 1.;
        return r;
    }
};

#define JAC_UNARY_MAP(expr_v, deriv)                      \
    Jac<K> r;                                            \
    r.v = (expr_v);                                      \
    const double dd_ = (deriv);                          \
    for (int i = 0; i < K; i++) r.d[i] = dd_ * a.d[i];   \
    return r;

template <int K>
__host__ __device__ inline Jac<K> operator+(const Jac<K> &a, const Jac<K> &b) {
    Jac<K> r;
    r.v = a.v + b.v;
    for (int i = 0; i < K; i++) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <int K>
__host__ __device__ inline Jac<K> operator-(const Jac<K> &a, const Jac<K> &b) {
    Jac<K> r;
    r.v = a.v - b.v;
    for (int i = 0; i < K; i++) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <int K>
__host__ __device__ inline Jac<K> operator+(const Jac<K> &a) {
    return a;
}
template <int K>
__host__ __device__ inline Jac<K> operator-(const Jac<K> &a) {
    Jac<K> r;
    r.v = -a.v;
    for (int i = 0; i < K; i++) r.d[i] = -a.d[i];
    return r;
}
template <int K>
__host__ __device__ inline Jac<K> operator*(const Jac<K> &a, const Jac<K> &b) {
    Jac<K> r;
    r.v = a.v * b.v;
    for (int i = 0; i < K; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <int K>
__host__ __device__ inline Jac<K> operator/(const Jac<K> &a, const Jac<K> &b) {
    Jac<K> r;
    r.v = a.v / b.v;
    const double rb = 1. / b.v;
    for (int i = 0; i < K; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) * rb;
    return r;
}
// mixed with plain doubles (constants)
template <int K>
__host__ __device__ inline Jac<K> operator+(const Jac<K> &a, double b) {
    Jac<K> r = a;
    r.v = a.v + b;
    return r;
}
template <int K>
__host__ __device__ inline Jac<K> operator+(double a, const Jac<K> &b) {
    Jac<K> r = b;
    r.v = a + b.v;
    return r;
}
template <int K>
__host__ __device__ inline Jac<K> operator-(const Jac<K> &a, double b) {
    Jac<K> r = a;
    r.v = a.v - b;
    return r;
}
template <int K>
__host__ __device__ inline Jac<K> operator-(double a, const Jac<K> &b) {
    Jac<K> r;
    r.v = a - b.v;
    for (int i = 0; i < K; i++) r.d[i] = -b.d[i];
    return r;
}
template <int K>
__host__ __device__ inline Jac<K> operator*(const Jac<K> &a, double b) {
    Jac<K> r;
    r.v = a.v * b;
    for (int i = 0; i < K; i++) r.d[i] = a.d[i] * b;
    return r;
}
template <int K>
__host__ __device__ inline Jac<K> operator*(double a, const Jac<K> &b) {
    Jac<K> r;
    r.v = a * b.v;
    for (int i = 0; i < K; i++) r.d[i] = a * b.d[i];
    return r;
}
template <int K>
__host__ __device__ inline Jac<K> operator/(const Jac<K> &a, double b) {
    Jac<K> r;
    r.v = a.v / b;
    for (int i = 0; i < K; i++) r.d[i] = a.d[i] / b;
    return r;
}
template <int K>
__host__ __device__ inline Jac<K> operator/(double a, const Jac<K> &b) {
    Jac<K> r;
    r.v = a / b.v;
    const double f = -r.v / b.v;
    for (int i = 0; i < K; i++) r.d[i] = f * b.d[i];
    return r;
}

// comparisons on the value (frozen switches)
template <int K> __host__ __device__ inline bool operator>(const Jac<K> &a, const Jac<K> &b) { return LIN_DECIDE(a.v > b.v); }
template <int K> __host__ __device__ inline bool operator<(const Jac<K> &a, const Jac<K> &b) { return LIN_DECIDE(a.v < b.v); }
template <int K> __host__ __device__ inline bool operator>=(const Jac<K> &a, const Jac<K> &b) { return LIN_DECIDE(a.v >= b.v); }
template <int K> __host__ __device__ inline bool operator<=(const Jac<K> &a, const Jac<K> &b) { return LIN_DECIDE(a.v <= b.v); }
template <int K> __host__ __device__ inline bool operator>(const Jac<K> &a, double b) { return LIN_DECIDE(a.v > b); }
template <int K> __host__ __device__ inline bool operator<(const Jac<K> &a, double b) { return LIN_DECIDE(a.v < b); }

template <int K>
__host__ __device__ inline Jac<K> fabs(const Jac<K> &a) {
    const bool neg = LIN_DECIDE(a.v < 0.);
    JAC_UNARY_MAP(neg ? -a.v : a.v, neg ? -1. : 1.)
}
template <int K>
__host__ __device__ inline Jac<K> log(const Jac<K> &a) {
    JAC_UNARY_MAP(::log(a.v), 1. / a.v)
}
// Fortran MAX / MIN of two values: the derivative of the operand selected
template <int K>
__host__ __device__ inline Jac<K> dmax2(const Jac<K> &a, const Jac<K> &b) { return LIN_DECIDE(a.v >= b.v) ? a : b; }
template <int K>
__host__ __device__ inline Jac<K> dmin2(const Jac<K> &a, const Jac<K> &b) { return LIN_DECIDE(a.v <= b.v) ? a : b; }
// Fortran SIGN(a, b) = |a| with the sign of b; piecewise constant in b
template <int K>
__host__ __device__ inline Jac<K> fsign(const Jac<K> &a, const Jac<K> &b) {
    const bool an = LIN_DECIDE(a.v < 0.), bn = LIN_DECIDE(__builtin_signbit(b.v) != 0);
    const double s = (an ? -1. : 1.) * (bn ? -1. : 1.);
    const double m = an ? -a.v : a.v, sv = bn ? -m : m;
    JAC_UNARY_MAP(sv == 0. ? __builtin_copysign(0., bn ? -1. : 1.) : sv, s)
}
// x / 3 (the forward value is div3's correctly rounded quotient)
template <int K>
__host__ __device__ inline Jac<K> div3(const Jac<K> &a) {
    JAC_UNARY_MAP(div3(a.v), 1. / 3.)
}
// x**kappa: d/dx = kappa x**kappa / x
template <int K>
__host__ __device__ inline Jac<K> pow_kappa(const Jac<K> &a) {
    const double pk = pow_kappa(a.v);
    JAC_UNARY_MAP(pk, ATM_KAPPA * pk / a.v)
}
#undef JAC_UNARY_MAP

// value of a scalar of either kind
__host__ __device__ inline double val(double x) { return x; }
template <int K>
__host__ __device__ inline double val(const Jac<K> &x) { return x.v; }
