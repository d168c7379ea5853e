// geodesic.hpp -- metric plug-ins, ray set-up and colouring, shared by the
// gfx950 kernels and by host-side evaluation (prr_debug_rhs).
//
// The reference selects one metric header at compile time (main/metric.h:12)
// whose getGeod / getMetric / User_* hooks are inlined into trace_RKF.  Here
// each metric is a struct with a templated right-hand side `rhs<T>` (T =
// double for the plain form, float for the FP32 bulk stages) and the kernel
// is instantiated once per metric, so all four ship in one library and are
// chosen at run time.
//
// State vectors:
//   Flat, SchwKS, SchwSph: y = (x^1, x^2, x^3, t, u^1, u^2, u^3, u^t), the
//     reference's Ray r/u pair (main/base.h:172-176); f(y) = (u, getGeod(x, u)).
//   Kerr (Boyer-Lindquist): y = (1/r, d(1/r)/ds, cos th, d(cos th)/ds, phi,
//     lambda) in Mino time s, with E, L and Carter's Q held per ray.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "internal.hpp"

namespace prr {

#define PRR_HD __host__ __device__ inline
// Out of line: for code that is rare on the integrator's path, so that its
// registers do not count against the loop's.
#define PRR_NOINLINE __attribute__((noinline))

// ----------------------------------------------------------------- helpers
struct D3 {
    double x, y, z;
};
PRR_HD D3 d3(double x, double y, double z) { return D3{x, y, z}; }
PRR_HD D3 operator+(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
PRR_HD D3 operator-(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
PRR_HD D3 operator*(double s, D3 a) { return d3(s * a.x, s * a.y, s * a.z); }
PRR_HD double dot3(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PRR_HD D3 normalize3(D3 a) { return (1.0 / sqrt(dot3(a, a))) * a; }

PRR_HD void sincos_t(double x, double *s, double *c) { sincos(x, s, c); }

// Light FP64 sine and cosine for the per-step Cartesian maps: Cody-Waite
// reduction by pi/2 (three 33-bit parts, exact products for |x| < 2^20 pi/2)
// and Taylor polynomials on |r| <= pi/4 (truncation < 1e-19).  Register
// light, unlike a library sincos with a large-argument reduction path.
PRR_HD void sincos_cw(double x, double *s, double *c) {
    const double k = rint(x * 0.63661977236758134308);
    double r = fma(-k, 1.5707963267341256, x);
    r = fma(-k, 6.077100506303966e-11, r);
    r = fma(-k, 2.0222662487959506e-21, r);
    const double r2 = r * r;
    const double sp =
        r + r * r2 *
                (-1.0 / 6.0 +
                 r2 * (1.0 / 120.0 +
                       r2 * (-1.0 / 5040.0 +
                             r2 * (1.0 / 362880.0 +
                                   r2 * (-1.0 / 39916800.0 +
                                         r2 * (1.0 / 6227020800.0 +
                                               r2 * (-1.0 / 1307674368000.0 + r2 * (1.0 / 355687428096000.0))))))));
    const double cp =
        1.0 + r2 * (-0.5 +
                    r2 * (1.0 / 24.0 +
                          r2 * (-1.0 / 720.0 +
                                r2 * (1.0 / 40320.0 +
                                      r2 * (-1.0 / 3628800.0 +
                                            r2 * (1.0 / 479001600.0 +
                                                  r2 * (-1.0 / 87178291200.0 + r2 * (1.0 / 20922789888000.0))))))));
    const int q = int(int64_t(k) & 3);
    if (q == 0) {
        *s = sp;
        *c = cp;
    } else if (q == 1) {
        *s = cp;
        *c = -sp;
    } else if (q == 2) {
        *s = -sp;
        *c = -cp;
    } else {
        *s = -cp;
        *c = sp;
    }
}

// FP32 sine and cosine for screening tests: the hardware approximations on
// the device (|error| ~ 1e-6 on [-pi, pi]); x is reduced to [-pi, pi] first.
PRR_HD void sincos_screen(double x, float *s, float *c) {
    const float r = float(x - 6.28318530717958647692 * rint(x * 0.15915494309189533577));
#if defined(__HIP_DEVICE_COMPILE__)
    *s = __sinf(r);
    *c = __cosf(r);
#else
    *s = sinf(r);
    *c = cosf(r);
#endif
}
PRR_HD void sincos_t(float x, float *s, float *c) { sincosf(x, s, c); }
PRR_HD double sqrt_t(double x) { return sqrt(x); }
PRR_HD float sqrt_t(float x) { return sqrtf(x); }

// Reciprocal: FP64 to ~10 ulp on the device; the hardware reciprocal (1 ulp)
// for the FP32 bulk stages on the device; exact division on the host.
PRR_HD double rcp_t(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    // hardware estimate and one Newton step, three instructions instead of the
    // ~10 of the IEEE division sequence.  Measured on gfx950 over 2^20 normal
    // x across 60 binades: the estimate is within 4.6e-8 of 1 / x, one step
    // within 2.2e-15 (a second step would give the rounded quotient).  x = 0
    // gives NaN.
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
#else
    return 1.0 / x;
#endif
}
PRR_HD float rcp_t(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
template <class T> PRR_HD T rcp_t(const T &x) { return T(1.0) / x; }
// FP32 square root for screens and bounds: the hardware approximation
// (v_sqrt_f32, ~1 ulp) on the device instead of the correctly rounded sequence.
PRR_HD float sqrt_f32(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}
// FP32 reciprocal square root: the hardware approximation on the device.
PRR_HD float rsqrt_f32(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / sqrtf(x);
#endif
}

// max(|a|, |b|) for values that are never signalling NaNs.  On the device one
// v_max_f64 with absolute-value source modifiers: fmax must quiet signalling
// NaNs first, which costs a v_max_f64 per operand whose origin the compiler
// cannot see (a loop-carried state).  NaN operands give the other operand,
// as fmax does.
PRR_HD double fmax_abs(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r;
    __asm__("v_max_f64 %0, |%1|, |%2|" : "=v"(r) : "v"(a), "v"(b));
    return r;
#else
    return fmax(fabs(a), fabs(b));
#endif
}

// Per-ray constants: mass and spin, and for Kerr the conserved energy E,
// axial angular momentum L and Carter constant Q.
template <class T> struct Consts {
    T M, a, E, L, Q;
};

// Right-hand-side constants of the 8-component metrics.
template <class T> struct MassCoef {
    T M;
    T q2; // Schwarzschild coordinates: the squared charge of Reissner-Nordstrom (0: Schwarzschild)
};

// ----------------------------------------------------------------- metrics
// metrics/flat.cuh: straight lines, getGeod = 0 (:88-90).
struct Flat {
    // the integrator carries the adjoint's checkpoint and replay paths
    static constexpr bool adj = true, ckpt = true, replay = true;
    static constexpr int N = 8;
    static constexpr int kind = PRR_FLAT;
    static constexpr bool spherical = false;
    static constexpr int NERR = 8;
    static constexpr bool mino = false;
    template <class T> PRR_HD static MassCoef<T> coef(const Consts<double> &c) { return MassCoef<T>{T(c.M)}; }
    template <class T> PRR_HD static void rhs(const T *y, const MassCoef<T> &, T *f) {
        f[0] = y[4];
        f[1] = y[5];
        f[2] = y[6];
        f[3] = y[7];
        f[4] = f[5] = f[6] = f[7] = T(0);
    }
};

// metrics/schwarzschild.cuh:90-119 (Kerr-Schild Cartesian), factored: with
// q = x.u / r, w = r_s / r,
//   du^i = x^i r_s/(2 r^3) [3q^2 - 2|u|^2 - (u^t)^2 + w (q + u^t)^2]
//   du^t = r_s/(2 r^2) [2|u|^2 - 4q^2 - 2 q u^t - w (q + u^t)^2].
struct SchwKS {
    // the integrator carries the adjoint's checkpoint and replay paths
    static constexpr bool adj = true, ckpt = true, replay = true;
    static constexpr int N = 8;
    static constexpr int kind = PRR_SCHW_KS_CART;
    static constexpr bool spherical = false;
    static constexpr int NERR = 8;
    static constexpr bool mino = false;
    template <class T> PRR_HD static MassCoef<T> coef(const Consts<double> &c) { return MassCoef<T>{T(c.M)}; }
    template <class T> PRR_HD static void rhs(const T *y, const MassCoef<T> &c, T *f) {
        const T x = y[0], yy = y[1], z = y[2];
        const T ux = y[4], uy = y[5], uz = y[6], ut = y[7];
        const T rs = T(2) * c.M;
        const T r = sqrt_t(x * x + yy * yy + z * z);
        const T ir = T(1) / r;
        const T q = (x * ux + yy * uy + z * uz) * ir;
        const T uu = ux * ux + uy * uy + uz * uz;
        const T w = rs * ir;
        const T qt = q + ut;
        const T k = T(0.5) * w * ir * ir; // r_s / (2 r^3)
        const T Phi = k * (T(3) * q * q - T(2) * uu - ut * ut + w * qt * qt);
        f[0] = ux;
        f[1] = uy;
        f[2] = uz;
        f[3] = ut;
        f[4] = x * Phi;
        f[5] = yy * Phi;
        f[6] = z * Phi;
        f[7] = k * r * (T(2) * uu - T(4) * q * q - T(2) * q * ut - w * qt * qt);
    }
};

// "metrics/schwarzschild - sph.cuh":94-115, including its tan/sin clamp
// (|tan th| < 1e-4 -> tan = sin = 1e-4, :98-101).
struct SchwSph {
    // the integrator carries the adjoint's checkpoint and replay paths
    static constexpr bool adj = true, ckpt = true, replay = true;
    static constexpr int N = 8;
    static constexpr int kind = PRR_SCHW_SPH;
    static constexpr bool spherical = true;
    static constexpr int NERR = 8;
    static constexpr bool mino = false;
    // the metric's second parameter (Consts::a, the spin slot) is the charge q
    template <class T> PRR_HD static MassCoef<T> coef(const Consts<double> &c) {
        return MassCoef<T>{T(c.M), T(c.a * c.a)};
    }
    template <class T> PRR_HD static void rhs(const T *y, const MassCoef<T> &c, T *f) {
        const T r = y[0];
        const T ur = y[4], uth = y[5], uph = y[6], ut = y[7];
        T s, co;
        sincos_t(y[1], &s, &co);
        T tn = s / co;
        if (fabs(tn) < T(1e-4)) {
            tn = T(1e-4);
            s = T(1e-4);
        }
        // f = 1 - r_s/r + q^2/r^2 (Reissner-Nordstrom; q = 0: the reference's
        // Schwarzschild), A = f'/2 = (M - q^2/r)/r^2
        const T rs = T(2) * c.M;
        const T ir = T(1) / r;
        const T fr = T(1) - rs * ir + c.q2 * ir * ir;
        const T A = (c.M - c.q2 * ir) * ir * ir;
        f[0] = ur;
        f[1] = uth;
        f[2] = uph;
        f[3] = ut;
        f[4] = A * (ur * ur / fr - fr * ut * ut) + r * fr * (uth * uth + s * s * uph * uph);
        f[5] = T(-2) * ir * ur * uth + s * co * uph * uph;
        f[6] = T(-2) * ir * ur * uph - T(2) * uth * uph / tn;
        f[7] = T(-2) * A / fr * ur * ut;
    }
};

// Ellis wormhole, ds^2 = -dt^2 + dl^2 + (b^2 + l^2) (dth^2 + sin^2 th dph^2)
// (Ellis 1973; Morris and Thorne 1988), in (l, th, ph, t) with l on both
// sides of the throat l = 0 of radius b (the metric's second parameter,
// Consts::a).  The reference keeps a second sky for this case ("l < 0 if
// wormhole", main/base.h:168) without a metric that reaches it.  With
// R^2 = b^2 + l^2: Gamma^l_thth = -l, Gamma^l_phph = -l sin^2 th,
// Gamma^th_lth = Gamma^ph_lph = l / R^2, and the angular symbols of flat
// space; tan and sin are clamped to 1e-4 in magnitude near the axis, with
// their signs (the path crosses the axis: axis_post_step).
struct Ellis {
    // the integrator carries the adjoint's checkpoint and replay paths
    static constexpr bool adj = true, ckpt = true, replay = true;
    static constexpr int N = 8;
    static constexpr int kind = PRR_ELLIS;
    static constexpr bool spherical = true;
    static constexpr int NERR = 8;
    static constexpr bool mino = false;
    template <class T> PRR_HD static MassCoef<T> coef(const Consts<double> &c) {
        return MassCoef<T>{T(0), T(c.a * c.a)}; // q2: b^2
    }
    template <class T> PRR_HD static void rhs(const T *y, const MassCoef<T> &c, T *f) {
        const T l = y[0];
        const T ul = y[4], uth = y[5], uph = y[6], ut = y[7];
        T s, co;
        sincos_t(y[1], &s, &co);
        T tn = s / co;
        if (fabs(tn) < T(1e-4)) { // within 1e-4 of the axis, keeping the signs
            tn = copysign(T(1e-4), tn);
            s = copysign(T(1e-4), s);
        }
        const T k = l / (c.q2 + l * l); // l / R^2
        f[0] = ul;
        f[1] = uth;
        f[2] = uph;
        f[3] = ut;
        f[4] = l * (uth * uth + s * s * uph * uph);
        f[5] = T(-2) * k * ul * uth + s * co * uph * uph;
        f[6] = T(-2) * k * ul * uph - T(2) * uth * uph / tn;
        f[7] = T(0);
    }
};

// Kerr, Boyer-Lindquist (metric of metrics/kerr.cuh:77-100; the reference's
// getGeod is a stub).  Integrated in Mino time s (d lambda = Sigma ds) with
// u = 1/r and mu = cos th, where the geodesic equations separate into two
// polynomial oscillators and two quadratures (Carter 1968):
//   u'' = P'(u)/2 = u (c1 + u (c2 + u c3)),        P(u) = R(r) / r^4,
//   mu'' = Th'(mu)/2 = mu (d1 + d2 mu^2),           Th(mu) = (1 - mu^2) Theta,
//   phi' = u (2 M a E - a^2 L u) / D(u) + L / (1 - mu^2),  D = 1 - 2 M u + a^2 u^2,
//   lambda' = Sigma = 1/u^2 + a^2 mu^2,
// with K = (L - a E)^2 + Q, A = a (a E - L):
//   c1 = 2 E A - K, c2 = 3 M K, c3 = 2 (A^2 - a^2 K),
//   d1 = -(Q + L^2 - a^2 E^2), d2 = -2 a^2 E^2.
// State y = (u, u', mu, mu', phi, lambda); E = -p_t, L = p_phi and Carter's
// Q are fixed per ray.  No term cancels catastrophically, so the stages are
// well conditioned in FP32, and u' -> -E as r -> infinity (no pole in s).
// t is not integrated (it enters nothing observable).
template <class T> struct KerrCoef {
    T c1, c2, c3, d1, d2, pa, pb, m2, a2, L;
};
// FP64-anchored stage data of the FP32 bulk form: f(y0 + d) - f(y0) is a
// polynomial / rational function of d whose coefficients depend on y0 only.
struct KerrDelta {
    float g1u, g2u, c3, g1m, g2m, d2; // u'' and mu'' Taylor coefficients
    float n1, n2, D0, dD1, a2;        // phi' first term: d (n1 + d n2) / (D0 D)
    float mu0, s20, Ls, u0, iu02;     // phi' second term and lambda'
};

struct Kerr {
    // the integrator carries the adjoint's checkpoint and replay paths
    static constexpr bool adj = true, ckpt = true, replay = true;
    static constexpr int N = 6;
    static constexpr int NERR = 5; // lambda is bookkeeping: outside the error norm
    static constexpr int kind = PRR_KERR_BL;
    static constexpr bool spherical = true;
    static constexpr bool mino = true;

    // Coefficients in any scalar type (double, or the duals of the adjoint).
    // Contraction is off so that every kernel that computes them (the lane at
    // its ray's load, the adjoint's step transpose) gets the same bits.
    template <class T>
    PRR_HD static KerrCoef<T> coef_of(double M, const T &a, const T &E, const T &L, const T &Q) {
#pragma clang fp contract(off)
        const T K = (L - a * E) * (L - a * E) + Q, A = a * (a * E - L);
        KerrCoef<T> k;
        k.c1 = T(2.0) * E * A - K;
        k.c2 = T(3.0 * M) * K;
        k.c3 = T(2.0) * (A * A - a * a * K);
        k.d1 = T(0.0) - (Q + L * L - a * a * E * E);
        k.d2 = T(-2.0) * a * a * E * E;
        k.pa = T(2.0 * M) * a * E;
        k.pb = a * a * L;
        k.m2 = T(2.0 * M);
        k.a2 = a * a;
        k.L = L;
        return k;
    }
    // Coefficients of a ray: computed in FP64, then rounded to T.
    template <class T> PRR_HD static KerrCoef<T> coef(const Consts<double> &c) {
        return cast<T>(coef_of<double>(c.M, c.a, c.E, c.L, c.Q));
    }
    template <class T> PRR_HD static KerrCoef<T> cast(const KerrCoef<double> &d) {
        return KerrCoef<T>{T(d.c1), T(d.c2), T(d.c3), T(d.d1), T(d.d2), T(d.pa), T(d.pb), T(d.m2), T(d.a2), T(d.L)};
    }

    // One reciprocal for the three denominators D, 1 - mu^2 and u^2.
    template <class T> PRR_HD static void rhs(const T *y, const KerrCoef<T> &k, T *f) {
        const T u = y[0], mu = y[2];
        const T D = T(1) - u * (k.m2 - k.a2 * u);
        const T s2 = T(1) - mu * mu;
        const T u2 = u * u;
        const T q = rcp_t(D * s2 * u2);
        f[0] = y[1];
        f[1] = u * (k.c1 + u * (k.c2 + u * k.c3));
        f[2] = y[3];
        f[3] = mu * (k.d1 + k.d2 * mu * mu);
        f[4] = u * (k.pa - k.pb * u) * (q * s2 * u2) + k.L * (q * D * u2);
        f[5] = q * D * s2 + k.a2 * mu * mu;
    }

    // FP64 part of an FP32-bulk attempt: F = f(y0) and the difference
    // coefficients at y0.
    PRR_HD static void delta_prep(const double *y, const KerrCoef<double> &k, double *F, KerrDelta &P) {
        const double u0 = y[0], mu0 = y[2];
        const double D0 = 1.0 - u0 * (k.m2 - k.a2 * u0), s20 = 1.0 - mu0 * mu0, u02 = u0 * u0;
        const double q = rcp_t(D0 * s20 * u02);
        const double iD0 = q * s20 * u02, is20 = q * D0 * u02, iu02 = q * D0 * s20;
        const double N0 = u0 * (k.pa - k.pb * u0);
        F[0] = y[1];
        F[1] = u0 * (k.c1 + u0 * (k.c2 + u0 * k.c3));
        F[2] = y[3];
        F[3] = mu0 * (k.d1 + k.d2 * mu0 * mu0);
        F[4] = N0 * iD0 + k.L * is20;
        F[5] = iu02 + k.a2 * mu0 * mu0;
        P.g1u = float(k.c1 + u0 * (2.0 * k.c2 + 3.0 * k.c3 * u0));
        P.g2u = float(k.c2 + 3.0 * k.c3 * u0);
        P.c3 = float(k.c3);
        P.g1m = float(k.d1 + 3.0 * k.d2 * mu0 * mu0);
        P.g2m = float(3.0 * k.d2 * mu0);
        P.d2 = float(k.d2);
        const double dD1 = 2.0 * k.a2 * u0 - k.m2;
        P.n1 = float((D0 * (k.pa - 2.0 * k.pb * u0) - N0 * dD1) * iD0);
        P.n2 = float((-k.pb * D0 - k.a2 * N0) * iD0);
        P.D0 = float(D0);
        P.dD1 = float(dD1);
        P.a2 = float(k.a2);
        P.mu0 = float(mu0);
        P.s20 = float(s20);
        P.Ls = float(k.L * is20);
        P.u0 = float(u0);
        P.iu02 = float(iu02);
    }

    // g = f(y0 + d) - f(y0) in FP32.
    PRR_HD static void delta_rhs(const float *d, const KerrDelta &P, float *g) {
        const float du = d[0], dm = d[2];
        const float D = P.D0 + du * (P.dD1 + P.a2 * du);
        const float tm = dm * (2.0f * P.mu0 + dm); // mu^2 - mu0^2
        const float s2 = P.s20 - tm;
        const float u = P.u0 + du;
        const float u2 = u * u;
        const float q = rcp_t(D * s2 * u2);
        g[0] = d[1];
        g[1] = du * (P.g1u + du * (P.g2u + du * P.c3));
        g[2] = d[3];
        g[3] = dm * (P.g1m + dm * (P.g2m + dm * P.d2));
        g[4] = du * (P.n1 + du * P.n2) * (q * s2 * u2) + P.Ls * tm * (q * D * u2);
        g[5] = -du * (2.0f * P.u0 + du) * P.iu02 * (q * D * s2) + P.a2 * tm;
    }

    // Every accepted step re-imposes the first integrals u'^2 = P(u) and
    // mu'^2 = Th(mu) in FP64 (E, L, Q are exact per-ray constants), keeping
    // the signs, except near turning points where the square roots are
    // ill-conditioned.  Truncation and FP32 rounding otherwise move the state
    // off the constraint surface, and a pass near the polar axis amplifies
    // that drift (phi jumps by ~pi L / sqrt(K sin^2 th_min)).
    PRR_HD static void project(double *y, const Consts<double> &c) {
        const double a = c.a, E = c.E, L = c.L, Q = c.Q, M = c.M;
        const double K = (L - a * E) * (L - a * E) + Q, A = a * (a * E - L);
        const double u = y[0], mu = y[2];
        const double u2 = u * u, m2 = mu * mu;
        const double t = E + A * u2;
        const double P = t * t - (1.0 - 2.0 * M * u + a * a * u2) * K * u2;
        // y' <- y' + (P - y'^2) / (2 y'): one Newton step towards copysign(sqrt(P), y')
        // from a state that is already within the step tolerance of it
        // (relative error d -> d^2 / 2), at a third of the cost of sqrt
        if (P > 1e-4 * E * E && y[1] * y[1] > 1e-4 * E * E) y[1] += (P - y[1] * y[1]) * rcp_t(2.0 * y[1]);
        const double Th = Q * (1.0 - m2) - m2 * L * L + a * a * E * E * m2 * (1.0 - m2);
        const double Kth = fabs(Q) + L * L + a * a * E * E;
        if (Th > 1e-4 * Kth && y[3] * y[3] > 1e-4 * Kth) y[3] += (Th - y[3] * y[3]) * rcp_t(2.0 * y[3]);
    }

    // Reverse mode of rhs: given fb = dL/df, accumulates dL/dy into yb[0..4]
    // (phi and lambda do not enter f) and dL/d(coefficients) into kb.
    PRR_HD static void rhs_vjp(const double *y, const KerrCoef<double> &k, const double *fb, double *yb,
                               KerrCoef<double> &kb) {
        const double u = y[0], mu = y[2];
        const double D = 1.0 - u * (k.m2 - k.a2 * u), s2 = 1.0 - mu * mu, u2 = u * u;
        // one reciprocal for the three denominators, as in rhs
        const double q = rcp_t(D * s2 * u2);
        const double iD = q * s2 * u2, is2 = q * D * u2, iu2 = q * D * s2;
        const double N = u * (k.pa - k.pb * u);
        // f0 = w, f2 = nu
        yb[1] += fb[0];
        yb[3] += fb[2];
        // f1 = c1 u + c2 u^2 + c3 u^3
        yb[0] += fb[1] * (k.c1 + u * (2.0 * k.c2 + 3.0 * k.c3 * u));
        kb.c1 += fb[1] * u;
        kb.c2 += fb[1] * u2;
        kb.c3 += fb[1] * u2 * u;
        // f3 = d1 mu + d2 mu^3
        yb[2] += fb[3] * (k.d1 + 3.0 * k.d2 * mu * mu);
        kb.d1 += fb[3] * mu;
        kb.d2 += fb[3] * mu * mu * mu;
        // f4 = N / D + L / s2, N = pa u - pb u^2, D = 1 - m2 u + a2 u^2
        {
            const double Nb = fb[4] * iD, Db = -fb[4] * N * iD * iD, s2b = -fb[4] * k.L * is2 * is2;
            kb.L += fb[4] * is2;
            yb[0] += Nb * (k.pa - 2.0 * k.pb * u) + Db * (2.0 * k.a2 * u - k.m2);
            kb.pa += Nb * u;
            kb.pb -= Nb * u2;
            kb.m2 -= Db * u;
            kb.a2 += Db * u2;
            yb[2] += s2b * (-2.0 * mu);
        }
        // f5 = 1/u^2 + a2 mu^2
        yb[0] += fb[5] * (-2.0 * iu2 / u);
        yb[2] += fb[5] * 2.0 * k.a2 * mu;
        kb.a2 += fb[5] * mu * mu;
    }

    // Chain rule of coef(): dL/d(coefficients) -> dL/d(E, L, Q, a) (added to cb).
    PRR_HD static void coef_vjp(const Consts<double> &c, const KerrCoef<double> &kb, double cb[4]) {
        const double a = c.a, E = c.E, L = c.L, Q = c.Q, M = c.M;
        const double K = (L - a * E) * (L - a * E) + Q, A = a * (a * E - L);
        double Kb = 0.0, Ab = 0.0, Eb = 0.0, Lb = kb.L, Qb = 0.0, ab = 0.0;
        // c1 = 2 E A - K
        Eb += kb.c1 * 2.0 * A;
        Ab += kb.c1 * 2.0 * E;
        Kb -= kb.c1;
        // c2 = 3 M K
        Kb += kb.c2 * 3.0 * M;
        // c3 = 2 (A^2 - a^2 K)
        Ab += kb.c3 * 4.0 * A;
        ab += kb.c3 * (-4.0 * a * K);
        Kb += kb.c3 * (-2.0 * a * a);
        // d1 = -(Q + L^2 - a^2 E^2)
        Qb -= kb.d1;
        Lb -= kb.d1 * 2.0 * L;
        ab += kb.d1 * 2.0 * a * E * E;
        Eb += kb.d1 * 2.0 * a * a * E;
        // d2 = -2 a^2 E^2
        ab += kb.d2 * (-4.0 * a * E * E);
        Eb += kb.d2 * (-4.0 * a * a * E);
        // pa = 2 M a E, pb = a^2 L, a2 = a^2 (m2 = 2M is fixed)
        ab += kb.pa * 2.0 * M * E + kb.pb * 2.0 * a * L + kb.a2 * 2.0 * a;
        Eb += kb.pa * 2.0 * M * a;
        Lb += kb.pb * a * a;
        // K = (L - a E)^2 + Q, A = a (a E - L)
        Lb += Kb * 2.0 * (L - a * E) - Ab * a;
        Eb += Kb * (-2.0 * a * (L - a * E)) + Ab * a * a;
        ab += Kb * (-2.0 * E * (L - a * E)) + Ab * (2.0 * a * E - L);
        Qb += Kb;
        cb[0] += Eb;
        cb[1] += Lb;
        cb[2] += Qb;
        cb[3] += ab;
    }
};

// ----------------------------------------------------------------- geometry
// Radius of a state (for the horizon, escape and FP64-region tests).
template <class Mt> PRR_HD double radius_of(const double *y) {
    if (Mt::mino) return rcp_t(y[0]);
    if (Mt::kind == PRR_ELLIS) return fabs(y[0]); // distance from the throat on either side
    if (Mt::spherical) return y[0];
    return sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
}

// Cartesian position (spherical-type coordinates map as flat space).
template <class Mt> PRR_HD D3 cart_of(const double *y) {
    if (!Mt::spherical) return d3(y[0], y[1], y[2]);
    double r, st, ct, sp, cp;
    if (Mt::mino) { // (u, u', mu, mu', phi, lambda)
        r = rcp_t(y[0]);
        ct = y[2];
        st = sqrt(fmax(1.0 - ct * ct, 0.0));
        sincos_cw(y[4], &sp, &cp);
    } else {
        r = y[0];
        sincos_cw(y[1], &st, &ct);
        sincos_cw(y[2], &sp, &cp);
    }
    return d3(r * st * cp, r * st * sp, r * ct);
}

// FP32 Cartesian position for the sphere screen (accuracy ~1e-6 relative).
// Kerr: the cylindrical radius and height (rho, z) of the screen's FP32
// position, from the same r, sin th and cos th values as cart_of_f32.
PRR_HD void kerr_rz_f32(const double *y, float P[2]) {
    const float r = rcp_t(float(y[0])), ct = float(y[2]);
    const float st = sqrt_f32(fmaxf(1.0f - ct * ct, 0.0f));
    P[0] = r * st;
    P[1] = r * ct;
}

template <class Mt> PRR_HD void cart_of_f32(const double *y, float P[3]) {
    if (!Mt::spherical) {
        P[0] = float(y[0]);
        P[1] = float(y[1]);
        P[2] = float(y[2]);
        return;
    }
    float r, st, ct, sp, cp;
    if (Mt::mino) {
        r = rcp_t(float(y[0]));
        ct = float(y[2]);
        st = sqrt_f32(fmaxf(1.0f - ct * ct, 0.0f));
        sincos_screen(y[4], &sp, &cp);
    } else {
        r = float(y[0]);
        sincos_screen(y[1], &st, &ct);
        sincos_screen(y[2], &sp, &cp);
    }
    P[0] = r * st * cp;
    P[1] = r * st * sp;
    P[2] = r * ct;
}

// pos_cartTOspher / vel_cartTOspher (cudamaths/transform.cuh:15-31) with the
// sign of the phi rate corrected (the reference's is (vx y - x vy)/(x^2+y^2)).
PRR_HD void cart_to_sph(D3 p, D3 d, double xs[3], double us[3]) {
    const double x2y2 = p.x * p.x + p.y * p.y, r2 = x2y2 + p.z * p.z, r = sqrt(r2);
    xs[0] = r;
    xs[1] = acos(p.z / r);
    xs[2] = atan2(p.y, p.x);
    us[0] = dot3(p, d) / r;
    us[1] = (p.z * (p.x * d.x + p.y * d.y) - x2y2 * d.z) / (r2 * sqrt(x2y2));
    us[2] = (p.x * d.y - p.y * d.x) / x2y2;
}

// createRays_* (kernels/ssaa.cuh:36-48, 92-99) and rayDir_* (main/Camera.h:30-42)
// for ray `local` of frame fr.
// SplitMix64's output function (Steele, Lea and Flood, OOPSLA 2014): a
// bijective 64-bit mix.  Hashing a counter gives the counter-based uniform
// stream of the RANDOM and JITTERED distributions: no per-pixel generator
// state (the reference keeps a curand state per pixel, kernels/ssaa.cuh:26-29).
PRR_HD uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// Uniform in [0, 1) (53 bits) for coordinate dim (0: u, 1: v) of ray `local` of a frame.
PRR_HD double sample_u01(uint32_t seed, int64_t local, int dim) {
    return double(mix64(mix64(seed) + 2ull * uint64_t(local) + uint64_t(dim)) >> 11) * 0x1.0p-53;
}

// Sample position of ray `local` of a frame: ray (j W + i) s + k is sample k
// of pixel (i, j) (kernels/ssaa.cuh:45).  NONE: the pixel centre
// (createRays_none, :92-99); REGULAR: grid point ((p + 1) / (n + 1),
// (q + 1) / (n + 1)) for k = q n + p (createRays_regular, :36-48); RANDOM:
// uniform over the pixel (createRays_random, :55-66, whose offset is not
// scaled to the pixel -- D23); JITTERED: uniform inside cell (p, q)
// (createRays_jittered, :73-86).
PRR_HD void ray_uv(const FrameDev &fr, int64_t local, double &u, double &v) {
    const int64_t pix = local / fr.s;
    const int k = int(local - pix * fr.s);
    const int j = int(pix / fr.W), i = int(pix - int64_t(j) * fr.W);
    if (fr.ssaa == PRR_SSAA_RANDOM) {
        u = (i + sample_u01(fr.seed, local, 0)) / fr.W;
        v = (j + sample_u01(fr.seed, local, 1)) / fr.H;
    } else if (fr.ssaa == PRR_SSAA_JITTERED) {
        const int q = k / fr.n, p = k - q * fr.n;
        u = (i + (p + sample_u01(fr.seed, local, 0)) / fr.n) / fr.W;
        v = (j + (q + sample_u01(fr.seed, local, 1)) / fr.n) / fr.H;
    } else if (fr.n == 0) {
        u = (i + 0.5) / fr.W;
        v = (j + 0.5) / fr.H;
    } else {
        const int q = k / fr.n, p = k - q * fr.n;
        u = (i + double(p + 1) / (fr.n + 1)) / fr.W;
        v = (j + double(q + 1) / (fr.n + 1)) / fr.H;
    }
}

// Direction of the camera ray through screen position (u, v) in [0, 1]^2
// (Camera::rayDir_PinHole / rayDir_Spherical, main/Camera.h:30-42).
PRR_HD D3 camera_dir(const FrameDev &fr, double u, double v) {
    if (fr.projection == PRR_PROJ_SPHERICAL) {
        const double nu = (0.5 - v) * fr.vfov, mu = (0.5 - u) * fr.aspect * fr.vfov;
        double sb, cb, sa, ca;
        sincos(fr.beta + nu, &sb, &cb);
        sincos(fr.alpha + mu, &sa, &ca);
        return normalize3(d3(cb * ca, cb * sa, sb));
    }
    const double cr = (2.0 * u - 1.0) * fr.tan_half * fr.aspect;
    const double cu = (1.0 - 2.0 * v) * fr.tan_half;
    return normalize3(d3(fr.F[0], fr.F[1], fr.F[2]) + cr * d3(fr.R[0], fr.R[1], fr.R[2]) +
                      cu * d3(fr.U[0], fr.U[1], fr.U[2]));
}

PRR_HD D3 camera_ray(const FrameDev &fr, int64_t local) {
    double u, v;
    ray_uv(fr, local, u, v);
    return camera_dir(fr, u, v);
}

// Larger root of g_tt (u^t)^2 + B u^t + C = 0 with u0 < 0 -> 1
// (initRays, kernels/solver.cuh:48-53; only u^t is set, D4).
PRR_HD double null_ut(double A, double B, double C) {
    const double rt = sqrt(B * B - 4.0 * A * C);
    double u0 = fmax((-B + rt) / (2.0 * A), (-B - rt) / (2.0 * A));
    if (u0 < 0.0) u0 = 1.0;
    return u0;
}

// Kerr BL metric components at (r, th): g_rr, g_thth, g_phph, g_tph, g_tt.
struct KerrG {
    double rr, thth, phph, tph, tt, Sig, Del;
};
PRR_HD KerrG kerr_g(double M, double a, double r, double th) {
    double s, c;
    sincos(th, &s, &c);
    const double s2 = s * s, a2 = a * a;
    KerrG g;
    g.Sig = r * r + a2 * c * c;
    g.Del = r * r - 2.0 * M * r + a2;
    g.rr = g.Sig / g.Del;
    g.thth = g.Sig;
    g.phph = (r * r + a2 + 2.0 * M * r * a2 * s2 / g.Sig) * s2;
    g.tph = -2.0 * M * r * a * s2 / g.Sig;
    g.tt = -(1.0 - 2.0 * M * r / g.Sig);
    return g;
}

// Kerr's integrator state and constants at (r, theta, phi) from the metric
// g there, sin / cos theta, 1 / r and u = dx/dlambda (cs.M and cs.a are set
// by the caller).
PRR_HD void kerr_state_at(const KerrG &g, double st, double ct, double ir, double ph, const double u[4], double a,
                          double *y, Consts<double> &cs) {
    cs.E = -(g.tt * u[3] + g.tph * u[2]);
    cs.L = g.phph * u[2] + g.tph * u[3];
    const double pth = g.Sig * u[1];
    cs.Q = pth * pth + ct * ct * (cs.L * cs.L / (st * st) - a * a * cs.E * cs.E);
    y[0] = ir;
    y[1] = -g.Sig * u[0] * ir * ir;
    y[2] = ct;
    y[3] = -st * pth;
    y[4] = ph;
    y[5] = 0.0;
}

// (x, u) in the metric's coordinates -> integrator state y and per-ray
// constants.  For the 8-component metrics y = (x, u).  For Kerr (x = (r, th,
// phi, t), u = dx/dlambda): E = -p_t, L = p_phi, Q = p_th^2 + cos^2 th
// (L^2 / sin^2 th - a^2 E^2), u' = du/ds = -Sigma u^r / r^2 and
// mu' = -sin th Sigma u^th.
template <class Mt> PRR_HD void state_from_xu(const double x[4], const double u[4], double M, double a, double *y,
                                              Consts<double> &cs) {
    cs.M = M;
    cs.a = a;
    cs.E = cs.L = cs.Q = 0.0;
    if (!Mt::mino) {
        for (int i = 0; i < 4; ++i) {
            y[i] = x[i];
            y[4 + i] = u[i];
        }
        return;
    }
    const KerrG g = kerr_g(M, a, x[0], x[1]);
    double st, ct;
    sincos(x[1], &st, &ct);
    kerr_state_at(g, st, ct, 1.0 / x[0], x[2], u, a, y, cs);
}

// Integrator state -> (x, u); for Kerr, t is returned as 0 and
// u = dx/dlambda = (dx/ds) / Sigma.
template <class Mt> PRR_HD void xu_from_state(const double *y, const Consts<double> &cs, double x[4], double u[4]) {
    if (!Mt::mino) {
        for (int i = 0; i < 4; ++i) {
            x[i] = y[i];
            u[i] = y[4 + i];
        }
        return;
    }
    double f[6];
    Kerr::rhs<double>(y, Kerr::coef<double>(cs), f);
    const double r = 1.0 / y[0], mu = y[2];
    const double th = acos(fmin(fmax(mu, -1.0), 1.0));
    const double st = sqrt(fmax(1.0 - mu * mu, 0.0));
    const double a = cs.a, M = cs.M;
    const double Sig = r * r + a * a * mu * mu, Del = r * r - 2.0 * M * r + a * a;
    const double W = (r * r + a * a) * cs.E - a * cs.L;
    x[0] = r;
    x[1] = th;
    x[2] = y[4];
    x[3] = 0.0;
    u[0] = -y[1] * r * r / Sig;
    u[1] = -y[3] / (st * Sig);
    u[2] = f[4] / Sig;
    u[3] = ((r * r + a * a) * W / Del - a * (a * cs.E * st * st - cs.L)) / Sig;
}

// initRays (kernels/solver.cuh:26-56): camera position and unit direction ->
// null state.  Cartesian metrics take u = (d, u^t); spherical-type metrics
// convert (pos, d) with vel_cartTOspher first (the SPHERICAL branch, :34-36).
template <class Mt> PRR_HD void init_state(D3 pos, D3 d, double M, double a, double *y, Consts<double> &cs) {
    double x[4], u[4];
    if (Mt::spherical) {
        double xs[3], us[3];
        cart_to_sph(pos, d, xs, us);
        x[0] = xs[0];
        x[1] = xs[1];
        x[2] = xs[2];
        u[0] = us[0];
        u[1] = us[1];
        u[2] = us[2];
        if (Mt::kind == PRR_ELLIS) { // a camera at areal radius r > b, on the l > 0 side: l = sqrt(r^2 - b^2)
            const double r = xs[0];
            x[0] = sqrt(fmax(r * r - a * a, 0.0));
            u[0] = us[0] * r / x[0];
        }
    } else {
        x[0] = pos.x;
        x[1] = pos.y;
        x[2] = pos.z;
        u[0] = d.x;
        u[1] = d.y;
        u[2] = d.z;
    }
    x[3] = 0.0;
    double A, B, C;
    if (Mt::kind == PRR_SCHW_KS_CART) {
        const double rs = 2.0 * M;
        const double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        const double xu = x[0] * u[0] + x[1] * u[1] + x[2] * u[2];
        A = -1.0 + rs / r;
        B = 2.0 * rs * xu / (r * r);
        C = u[0] * u[0] + u[1] * u[1] + u[2] * u[2] + rs * xu * xu / (r * r * r);
    } else if (Mt::kind == PRR_SCHW_SPH) { // a: the charge q
        const double f = 1.0 - 2.0 * M / x[0] + a * a / (x[0] * x[0]), s = sin(x[1]);
        A = -f;
        B = 0.0;
        C = u[0] * u[0] / f + x[0] * x[0] * (u[1] * u[1] + s * s * u[2] * u[2]);
    } else if (Mt::kind == PRR_ELLIS) { // a: the throat radius b
        const double R2 = a * a + x[0] * x[0], s = sin(x[1]);
        A = -1.0;
        B = 0.0;
        C = u[0] * u[0] + R2 * (u[1] * u[1] + s * s * u[2] * u[2]);
    } else if (Mt::kind == PRR_KERR_BL) {
        const KerrG g = kerr_g(M, a, x[0], x[1]);
        A = g.tt;
        B = 2.0 * g.tph * u[2];
        C = g.rr * u[0] * u[0] + g.thth * u[1] * u[1] + g.phph * u[2] * u[2];
    } else {
        A = -1.0;
        B = 0.0;
        C = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    }
    u[3] = null_ut(A, B, C);
    state_from_xu<Mt>(x, u, M, a, y, cs);
}

// initRays for a camera ray of frame fr.  For Kerr the camera position's
// coordinates, sin / cos theta and the metric there are per-frame constants
// of the frame table (frame_dev), so a ray only converts its direction into
// (dr, dtheta, dphi) and solves the null condition; the other metrics take
// the general path.
template <class Mt> PRR_HD void init_state_camera(const FrameDev &fr, D3 d, double *y, Consts<double> &cs) {
    const D3 p = d3(fr.pos[0], fr.pos[1], fr.pos[2]);
    if constexpr (Mt::kind == PRR_KERR_BL) {
        double u[4];
        u[0] = dot3(p, d) * fr.cam_ir;
        u[1] = (p.z * (p.x * d.x + p.y * d.y) - fr.cam_x2y2 * d.z) * fr.cam_iu1;
        u[2] = (p.x * d.y - p.y * d.x) * fr.cam_ix2y2;
        const double A = fr.g_tt, B = 2.0 * fr.g_tph * u[2];
        const double C = fr.g_rr * u[0] * u[0] + fr.g_sig * u[1] * u[1] + fr.g_phph * u[2] * u[2];
        u[3] = null_ut(A, B, C);
        KerrG g;
        g.tt = fr.g_tt;
        g.tph = fr.g_tph;
        g.rr = fr.g_rr;
        g.thth = g.Sig = fr.g_sig;
        g.phph = fr.g_phph;
        g.Del = 0.0;
        cs.M = fr.M;
        cs.a = fr.a;
        kerr_state_at(g, fr.cam_st, fr.cam_ct, fr.cam_ir, fr.cam_ph, u, fr.a, y, cs);
    } else {
        init_state<Mt>(p, d, fr.M, fr.a, y, cs);
    }
}

// User_ENDLOOP of the spherical metric ("metrics/schwarzschild - sph.cuh":69-78):
// pole flip, then phi in [0, 2pi) and theta in [0, pi].
PRR_HD void sph_post_step(double *y) {
    const double pi = 3.14159265358979323846, two_pi = 6.28318530717958647692;
    if (fabs(y[1]) <= 0.04 || fabs(y[1] - pi) <= 0.04) {
        y[2] += pi;
        y[5] = -y[5];
    }
    y[2] += two_pi;
    y[2] = fmod(y[2], two_pi);
    y[1] = fabs(y[1]);
    y[1] = fmod(y[1], two_pi);
    if (y[1] > pi && y[1] <= two_pi) y[1] = two_pi - y[1];
}

// Chart map of the wormhole's (theta, phi) across the polar axis: a state
// at theta < 0 (or > pi) is the same point and direction as the reflected
// theta with phi + pi and -theta', so the path stays continuous (the
// reference's flip within 0.04 rad of the axis, sph_post_step, moves the
// point); then phi in [0, 2pi).
PRR_HD void axis_post_step(double *y) {
    const double pi = 3.14159265358979323846, two_pi = 6.28318530717958647692;
    if (y[1] < 0.0 || y[1] > pi) {
        y[1] = y[1] < 0.0 ? -y[1] : two_pi - y[1];
        y[2] += pi;
        y[5] = -y[5];
    }
    y[2] -= two_pi * floor(y[2] / two_pi);
    if (y[2] >= two_pi) y[2] = 0.0;
}

// The chart's step-end map for the metrics in (r or l, theta, phi, t).
template <class Mt> PRR_HD void chart_post_step(double *y) {
    if (Mt::kind == PRR_SCHW_SPH) sph_post_step(y);
    if (Mt::kind == PRR_ELLIS) axis_post_step(y);
}

// Direction used for the celestial lookup at escape: the spatial velocity for
// Cartesian metrics (getColorDefault2, main/base.h:119-128), the position for
// Schwarzschild coordinates (getColorDefault, :68-84), the Cartesian
// coordinate velocity for Kerr: with e_r, e_th, e_ph the flat-space frame of
// (th, phi), dx/ds ~ -u' e_r / u^2 + r th' e_th + r sin th phi' e_ph, scaled
// by u^2 > 0: -u' e_r - u mu' / sin th e_th + u sin th phi' e_ph.
template <class Mt> PRR_HD D3 escape_dir(const double *y, const Consts<double> &cs) {
    if (Mt::kind == PRR_SCHW_SPH || Mt::kind == PRR_ELLIS) {
        double st, ct, sp, cp;
        sincos(y[1], &st, &ct);
        sincos(y[2], &sp, &cp);
        return d3(st * cp, st * sp, ct);
    }
    if (Mt::mino) {
        double f[6];
        Kerr::rhs<double>(y, Kerr::coef<double>(cs), f);
        double sp, cp;
        const double u = y[0], ct = y[2], st = sqrt(fmax(1.0 - ct * ct, 0.0));
        sincos(y[4], &sp, &cp);
        const D3 er = d3(st * cp, st * sp, ct), et = d3(ct * cp, ct * sp, -st), ep = d3(-sp, cp, 0.0);
        return normalize3((-y[1]) * er + (-u * y[3] / st) * et + (u * st * f[4]) * ep);
    }
    return normalize3(d3(y[4], y[5], y[6]));
}

// ----------------------------------------------------------------- colour
// Float arithmetic below is kept free of fused multiply-adds so that it
// rounds exactly like the reference's float expressions.
PRR_HD float4 texel(const TexDev &t, int i, int j) { return t.texels[size_t(j) * t.w + i]; }

// sqavg (cudamaths/f3.cuh:101-103): per-channel RMS of four texels.
PRR_HD float sqavg1(float a, float b, float c, float d) {
#pragma clang fp contract(off)
    return sqrtf(0.25f * (a * a + b * b + c * c + d * d));
}

// Bilinear lookup, texel centres at integer coordinates as tex2d::operator[](f2)
// (cudamaths/tex2d.cuh:112-119), u wrapped and v clamped.
PRR_HD float3 bilinear(const TexDev &t, double u, double v) {
#pragma clang fp contract(off)
    if (!(u == u) || !(v == v) || t.w <= 0 || t.h <= 0) return make_float3(0.f, 0.f, 0.f); // NaN guard
    u = u - floor(u);
    v = fmin(fmax(v, 0.0), 1.0);
    const double X = u * t.w, Y = v * t.h;
    int i0 = int(floor(X)), j0 = int(floor(Y));
    const float fx = float(X - i0);
    float fy = float(Y - j0);
    if (i0 >= t.w) i0 -= t.w;
    const int i1 = i0 + 1 >= t.w ? 0 : i0 + 1;
    if (j0 >= t.h - 1) {
        j0 = t.h - 1;
        fy = 0.0f;
    }
    const int j1 = j0 + 1 < t.h ? j0 + 1 : t.h - 1;
    const float4 c00 = texel(t, i0, j0), c10 = texel(t, i1, j0), c01 = texel(t, i0, j1), c11 = texel(t, i1, j1);
    const float w00 = (1.0f - fx) * (1.0f - fy), w10 = fx * (1.0f - fy), w01 = (1.0f - fx) * fy, w11 = fx * fy;
    return make_float3(c00.x * w00 + c10.x * w10 + c01.x * w01 + c11.x * w11,
                       c00.y * w00 + c10.y * w10 + c01.y * w01 + c11.y * w11,
                       c00.z * w00 + c10.z * w10 + c01.z * w01 + c11.z * w11);
}

// OutSphere::getColorDefault{,2} (main/base.h:68-162).
PRR_HD float3 sky_lookup(const TexDev &t, double th, double ph) {
#pragma clang fp contract(off)
    const double pi = 3.14159265358979323846;
    if (t.w == 0) return make_float3(0.f, 0.f, 0.f);
    if (!(ph <= 2.0 * pi && ph >= 0.0)) return make_float3(0.5f, 0.5f, 0.0f);
    if (!(th >= 0.0 && th <= pi)) return make_float3(0.0f, 0.5f, 0.8f);
    if (t.filter == PRR_FILTER_BILINEAR) return bilinear(t, ph / (2.0 * pi), th / pi);
    const double uvx = ph / (2.0 * pi) * t.w, uvy = th / pi * t.h;
    int px = int(floor(uvx)), py = int(floor(uvy));
    if (px == t.w) px -= 1;
    if (py == t.h) py -= 1;
    if (px < 0 || px >= t.w || py < 0 || py >= t.h) return make_float3(0.0f, 0.5f, 0.0f);
    if (px == t.w - 1 || py == t.h - 1) {
        const float4 c = texel(t, px, py);
        return make_float3(c.x, c.y, c.z);
    }
    const float4 a = texel(t, px, py), b = texel(t, px, py + 1), c = texel(t, px + 1, py), d = texel(t, px + 1, py + 1);
    return make_float3(sqavg1(a.x, b.x, c.x, d.x), sqavg1(a.y, b.y, c.y, d.y), sqavg1(a.z, b.z, c.z, d.z));
}

// Sphere::getColor (scene/sphere.h:59-64) at the unit normal n.
PRR_HD float3 sphere_color(const TexDev *t, D3 n) {
#pragma clang fp contract(off)
    const double pi = 3.14159265358979323846;
    const double th = acos(fmin(fmax(n.z, -1.0), 1.0));
    const double ph = atan2(n.y, n.x) + pi;
    float3 c = (t && t->w) ? bilinear(*t, ph / (2.0 * pi), th / pi) : make_float3(1.f, 1.f, 1.f);
    const float shade = fminf(fmaxf(0.5f + 0.5f * float(n.z), 0.0f), 1.0f);
    return make_float3(c.x * shade, c.y * shade, c.z * shade);
}

// Segment P0 -> P1 against spheres (exact chord intersection, smallest t).
PRR_HD int chord_hit(const FrameDev &fr, D3 P0, D3 P1, double &t_hit) {
    int best = -1;
    t_hit = 2.0;
    const D3 d = P1 - P0;
    const double A = dot3(d, d);
    for (int k = 0; k < fr.n_spheres; ++k) {
        const D3 m = P0 - d3(fr.sph[k][0], fr.sph[k][1], fr.sph[k][2]);
        const double C = dot3(m, m) - fr.sph[k][3] * fr.sph[k][3];
        double t;
        if (C <= 0.0) {
            t = 0.0;
        } else {
            if (A <= 0.0) continue;
            const double B = dot3(m, d), disc = B * B - A * C;
            if (disc < 0.0) continue;
            t = (-B - sqrt(disc)) / A;
            if (t < 0.0 || t > 1.0) continue;
        }
        if (t < t_hit) {
            t_hit = t;
            best = k;
        }
    }
    return best;
}

// A metric whose integrator has no adjoint paths (checkpoints, replay
// records, event capture): the render instantiation, with their registers
// and branches compiled out.
// Instantiations of the integrator: renders carry no adjoint code (NoAdj);
// the Kerr adjoint's checkpointing passes and its replay each carry only
// their own path (AdjCkpt, AdjReplay).
template <class M> struct NoAdj : M {
    static constexpr bool adj = false, ckpt = false, replay = false;
};
template <class M> struct AdjCkpt : M {
    static constexpr bool adj = true, ckpt = true, replay = false;
};
template <class M> struct AdjReplay : M {
    static constexpr bool adj = true, ckpt = false, replay = true;
};

} // namespace prr
