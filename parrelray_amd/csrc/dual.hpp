// dual.hpp -- forward-mode dual numbers with K tangent directions, used by
// the adjoint kernel for the once-per-ray Jacobians (ray set-up and the
// terminal colour).  The per-step Jacobian-transpose products use the
// hand-written reverse mode Kerr::rhs_vjp instead.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#define PRR_HD_DUAL __host__ __device__

namespace prr {

// Keep the scalar math functions visible next to the Dual overloads below
// (a declaration in namespace prr would otherwise hide the global ones).
using ::acos;
using ::atan2;
using ::cos;
using ::fabs;
using ::fmax;
using ::fmin;
using ::sin;
using ::sincos;
using ::sqrt;

template <int K> struct Dual {
    double v;
    double d[K];
    PRR_HD_DUAL Dual() : v(0.0) {
        for (int i = 0; i < K; ++i) d[i] = 0.0;
    }
    PRR_HD_DUAL Dual(double x) : v(x) { // NOLINT: constants convert implicitly
        for (int i = 0; i < K; ++i) d[i] = 0.0;
    }
    PRR_HD_DUAL static Dual var(double x, int dir) {
        Dual r(x);
        r.d[dir] = 1.0;
        return r;
    }
};

#define DUAL_T template <int K> __host__ __device__ inline

DUAL_T Dual<K> operator+(const Dual<K> &a, const Dual<K> &b) {
    Dual<K> r;
    r.v = a.v + b.v;
    for (int i = 0; i < K; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
DUAL_T Dual<K> operator-(const Dual<K> &a, const Dual<K> &b) {
    Dual<K> r;
    r.v = a.v - b.v;
    for (int i = 0; i < K; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
DUAL_T Dual<K> operator-(const Dual<K> &a) {
    Dual<K> r;
    r.v = -a.v;
    for (int i = 0; i < K; ++i) r.d[i] = -a.d[i];
    return r;
}
DUAL_T Dual<K> operator*(const Dual<K> &a, const Dual<K> &b) {
    Dual<K> r;
    r.v = a.v * b.v;
    for (int i = 0; i < K; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
DUAL_T Dual<K> operator/(const Dual<K> &a, const Dual<K> &b) {
    Dual<K> r;
    const double ib = 1.0 / b.v;
    r.v = a.v * ib;
    for (int i = 0; i < K; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
    return r;
}
DUAL_T Dual<K> operator+(const Dual<K> &a, double b) { return a + Dual<K>(b); }
DUAL_T Dual<K> operator+(double a, const Dual<K> &b) { return Dual<K>(a) + b; }
DUAL_T Dual<K> operator-(const Dual<K> &a, double b) { return a - Dual<K>(b); }
DUAL_T Dual<K> operator-(double a, const Dual<K> &b) { return Dual<K>(a) - b; }
DUAL_T Dual<K> operator*(const Dual<K> &a, double b) {
    Dual<K> r;
    r.v = a.v * b;
    for (int i = 0; i < K; ++i) r.d[i] = a.d[i] * b;
    return r;
}
DUAL_T Dual<K> operator*(double a, const Dual<K> &b) { return b * a; }
DUAL_T Dual<K> operator/(const Dual<K> &a, double b) { return a * (1.0 / b); }
DUAL_T Dual<K> operator/(double a, const Dual<K> &b) { return Dual<K>(a) / b; }
DUAL_T Dual<K> &operator+=(Dual<K> &a, const Dual<K> &b) { return a = a + b; }
DUAL_T Dual<K> &operator-=(Dual<K> &a, const Dual<K> &b) { return a = a - b; }
DUAL_T Dual<K> &operator*=(Dual<K> &a, const Dual<K> &b) { return a = a * b; }
DUAL_T bool operator<(const Dual<K> &a, const Dual<K> &b) { return a.v < b.v; }
DUAL_T bool operator>(const Dual<K> &a, const Dual<K> &b) { return a.v > b.v; }
DUAL_T bool operator<(const Dual<K> &a, double b) { return a.v < b; }
DUAL_T bool operator>(const Dual<K> &a, double b) { return a.v > b; }

// f(a) with f(a.v) = fv and f'(a.v) = fp
DUAL_T Dual<K> chain(const Dual<K> &a, double fv, double fp) {
    Dual<K> r;
    r.v = fv;
    for (int i = 0; i < K; ++i) r.d[i] = fp * a.d[i];
    return r;
}
DUAL_T Dual<K> sqrt(const Dual<K> &a) {
    const double s = ::sqrt(a.v);
    return chain(a, s, 0.5 / s);
}
DUAL_T Dual<K> sin(const Dual<K> &a) { return chain(a, ::sin(a.v), ::cos(a.v)); }
DUAL_T Dual<K> cos(const Dual<K> &a) { return chain(a, ::cos(a.v), -::sin(a.v)); }
DUAL_T void sincos(const Dual<K> &a, Dual<K> *s, Dual<K> *c) {
    double sv, cv;
    ::sincos(a.v, &sv, &cv);
    *s = chain(a, sv, cv);
    *c = chain(a, cv, -sv);
}
DUAL_T Dual<K> acos(const Dual<K> &a) { return chain(a, ::acos(a.v), -1.0 / ::sqrt(1.0 - a.v * a.v)); }
DUAL_T Dual<K> atan2(const Dual<K> &y, const Dual<K> &x) {
    Dual<K> r;
    r.v = ::atan2(y.v, x.v);
    const double q = 1.0 / (x.v * x.v + y.v * y.v);
    for (int i = 0; i < K; ++i) r.d[i] = (x.v * y.d[i] - y.v * x.d[i]) * q;
    return r;
}
DUAL_T Dual<K> fabs(const Dual<K> &a) { return a.v < 0.0 ? -a : a; }
DUAL_T Dual<K> fmax(const Dual<K> &a, const Dual<K> &b) { return a.v >= b.v ? a : b; }
DUAL_T Dual<K> fmin(const Dual<K> &a, const Dual<K> &b) { return a.v <= b.v ? a : b; }
DUAL_T Dual<K> fmax(const Dual<K> &a, double b) { return a.v >= b ? a : Dual<K>(b); }
DUAL_T Dual<K> fmin(const Dual<K> &a, double b) { return a.v <= b ? a : Dual<K>(b); }

template <int K> __host__ __device__ inline double value(const Dual<K> &a) { return a.v; }
__host__ __device__ inline double value(double a) { return a; }

#undef DUAL_T

} // namespace prr
