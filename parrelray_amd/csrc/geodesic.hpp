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
//   Kerr (Boyer-Lindquist): y = (r, mu = cos theta, phi, p_r, p_mu) with the
//     conserved E = -p_t and L = p_phi held per ray; f(y) = Hamilton's
//     equations of H = g^{ab} p_a p_b / 2.  t is not integrated (it enters
//     nothing observable).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "internal.hpp"

namespace prr {

#define PRR_HD __host__ __device__ inline

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
PRR_HD void sincos_t(float x, float *s, float *c) { sincosf(x, s, c); }
PRR_HD double sqrt_t(double x) { return sqrt(x); }
PRR_HD float sqrt_t(float x) { return sqrtf(x); }

template <class T> struct Consts {
    T M, a, E, L;
};

// ----------------------------------------------------------------- metrics
// metrics/flat.cuh: straight lines, getGeod = 0 (:88-90).
struct Flat {
    static constexpr int N = 8;
    static constexpr int kind = PRR_FLAT;
    static constexpr bool spherical = false;
    static constexpr bool hamiltonian = false;
    template <class T> PRR_HD static void rhs(const T *y, const Consts<T> &, T *f) {
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
    static constexpr int N = 8;
    static constexpr int kind = PRR_SCHW_KS_CART;
    static constexpr bool spherical = false;
    static constexpr bool hamiltonian = false;
    template <class T> PRR_HD static void rhs(const T *y, const Consts<T> &c, T *f) {
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
    static constexpr int N = 8;
    static constexpr int kind = PRR_SCHW_SPH;
    static constexpr bool spherical = true;
    static constexpr bool hamiltonian = false;
    template <class T> PRR_HD static void rhs(const T *y, const Consts<T> &c, T *f) {
        const T r = y[0];
        const T ur = y[4], uth = y[5], uph = y[6], ut = y[7];
        T s, co;
        sincos_t(y[1], &s, &co);
        T tn = s / co;
        if (fabs(tn) < T(1e-4)) {
            tn = T(1e-4);
            s = T(1e-4);
        }
        const T rs = T(2) * c.M;
        const T ir = T(1) / r;
        const T fr = T(1) - rs * ir; // 1 - r_s/r
        const T A = T(0.5) * rs * ir * ir;
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

// Kerr, Boyer-Lindquist (metric of metrics/kerr.cuh:77-100; the reference's
// getGeod is a stub), integrated as Hamilton's equations in the canonical
// pair (mu = cos th, p_mu = -p_th / sin th), so that the right-hand side is
// rational: no trigonometry in the hot loop.  With s2 = 1 - mu^2,
// Sigma = r^2 + a^2 mu^2, Delta = r^2 - 2Mr + a^2, W = (r^2 + a^2) E - a L,
// Lm = L - a E s2 and N = 2 Sigma H
//   = Delta p_r^2 + s2 p_mu^2 - W^2/Delta + Lm^2 / s2:
//   r'  = Delta p_r / Sigma,  mu' = s2 p_mu / Sigma,
//   ph' = (a W / Delta + Lm / s2) / Sigma,
//   p_r'  = -(dN/dr - 2 r N / Sigma) / (2 Sigma),
//     dN/dr = Delta' (p_r^2 + (W/Delta)^2) - 4 r E W / Delta,
//   p_mu' = -mu (L^2 / s2^2 - a^2 E^2 - p_mu^2 - a^2 N / Sigma) / Sigma.
// The chart is singular only where a ray with L = 0 crosses the axis
// exactly (s2 = 0); such a ray ends as NAN.
struct Kerr {
    static constexpr int N = 5;
    static constexpr int kind = PRR_KERR_BL;
    static constexpr bool spherical = true;
    static constexpr bool hamiltonian = true;
    template <class T> PRR_HD static void rhs(const T *y, const Consts<T> &c, T *f) {
        const T r = y[0], mu = y[1], pr = y[3], pm = y[4];
        const T a = c.a, M = c.M, E = c.E, L = c.L;
        const T r2 = r * r, a2 = a * a;
        const T s2 = T(1) - mu * mu;
        const T Sig = r2 + a2 * mu * mu;
        const T Del = r2 - T(2) * M * r + a2;
        const T iS = T(1) / Sig, iD = T(1) / Del, is2 = T(1) / s2;
        const T W = (r2 + a2) * E - a * L;
        const T Lm = L - a * E * s2;
        const T WiD = W * iD;
        const T Lmis2 = Lm * is2;
        const T Nn = Del * pr * pr + s2 * pm * pm - W * WiD + Lm * Lmis2;
        const T NiS = Nn * iS;
        const T dNr = T(2) * (r - M) * (pr * pr + WiD * WiD) - T(4) * r * E * WiD;
        const T Lis2 = L * is2;
        f[0] = Del * pr * iS;
        f[1] = s2 * pm * iS;
        f[2] = (a * WiD + Lmis2) * iS;
        f[3] = T(-0.5) * iS * (dNr - T(2) * r * NiS);
        f[4] = -mu * iS * (Lis2 * Lis2 - a2 * E * E - pm * pm - a2 * NiS);
    }

    // Reverse mode of rhs: given fb = dL/df, accumulate dL/dy into yb[5]
    // (yb[2] stays untouched: phi does not enter f) and dL/d(E, L, a) into
    // cb[0..2].  Written by hand from the forward expression graph above,
    // in reverse order.
    PRR_HD static void rhs_vjp(const double *y, const Consts<double> &c, const double *fb, double *yb, double *cb) {
        const double r = y[0], mu = y[1], pr = y[3], pm = y[4];
        const double a = c.a, M = c.M, E = c.E, L = c.L;
        // forward values
        const double r2 = r * r, a2 = a * a;
        const double s2 = 1.0 - mu * mu;
        const double Sig = r2 + a2 * mu * mu;
        const double Del = r2 - 2.0 * M * r + a2;
        const double iS = 1.0 / Sig, iD = 1.0 / Del, is2 = 1.0 / s2;
        const double W = (r2 + a2) * E - a * L;
        const double Lm = L - a * E * s2;
        const double WiD = W * iD;
        const double Lmis2 = Lm * is2;
        const double Nn = Del * pr * pr + s2 * pm * pm - W * WiD + Lm * Lmis2;
        const double NiS = Nn * iS;
        const double Aq = pr * pr + WiD * WiD;
        const double dNr = 2.0 * (r - M) * Aq - 4.0 * r * E * WiD;
        const double Lis2 = L * is2;
        const double g2 = a * WiD + Lmis2;
        const double g3 = dNr - 2.0 * r * NiS;
        const double g4 = Lis2 * Lis2 - a2 * E * E - pm * pm - a2 * NiS;
        // adjoints
        double rb = 0, mub = 0, prb = 0, pmb = 0, Eb = 0, Lb = 0, ab = 0;
        double iSb = 0, Delb = 0, s2b = 0, WiDb = 0, Lmis2b = 0, NiSb = 0, dNrb = 0, Lis2b = 0, a2b = 0;
        // f0 = Del pr iS
        Delb += fb[0] * pr * iS;
        prb += fb[0] * Del * iS;
        iSb += fb[0] * Del * pr;
        // f1 = s2 pm iS
        s2b += fb[1] * pm * iS;
        pmb += fb[1] * s2 * iS;
        iSb += fb[1] * s2 * pm;
        // f2 = g2 iS, g2 = a WiD + Lmis2
        iSb += fb[2] * g2;
        {
            const double g2b = fb[2] * iS;
            ab += g2b * WiD;
            WiDb += g2b * a;
            Lmis2b += g2b;
        }
        // f3 = -0.5 iS g3, g3 = dNr - 2 r NiS
        iSb += -0.5 * fb[3] * g3;
        {
            const double g3b = -0.5 * fb[3] * iS;
            dNrb += g3b;
            rb += g3b * (-2.0 * NiS);
            NiSb += g3b * (-2.0 * r);
        }
        // f4 = -mu iS g4, g4 = Lis2^2 - a2 E^2 - pm^2 - a2 NiS
        mub += -fb[4] * iS * g4;
        iSb += -fb[4] * mu * g4;
        {
            const double g4b = -fb[4] * mu * iS;
            Lis2b += g4b * 2.0 * Lis2;
            a2b += g4b * (-E * E - NiS);
            Eb += g4b * (-2.0 * a2 * E);
            pmb += g4b * (-2.0 * pm);
            NiSb += g4b * (-a2);
        }
        // Lis2 = L is2
        Lb += Lis2b * is2;
        double is2b = Lis2b * L;
        // dNr = 2 (r - M) Aq - 4 r E WiD, Aq = pr^2 + WiD^2
        rb += dNrb * (2.0 * Aq - 4.0 * E * WiD);
        prb += dNrb * 4.0 * (r - M) * pr;
        WiDb += dNrb * (4.0 * (r - M) * WiD - 4.0 * r * E);
        Eb += dNrb * (-4.0 * r * WiD);
        // NiS = Nn iS
        const double Nnb = NiSb * iS;
        iSb += NiSb * Nn;
        // Nn = Del pr^2 + s2 pm^2 - W WiD + Lm Lmis2
        Delb += Nnb * pr * pr;
        prb += Nnb * 2.0 * Del * pr;
        s2b += Nnb * pm * pm;
        pmb += Nnb * 2.0 * s2 * pm;
        double Wb = -Nnb * WiD;
        WiDb += -Nnb * W;
        double Lmb = Nnb * Lmis2;
        Lmis2b += Nnb * Lm;
        // Lmis2 = Lm is2
        Lmb += Lmis2b * is2;
        is2b += Lmis2b * Lm;
        // WiD = W iD
        Wb += WiDb * iD;
        const double iDb = WiDb * W;
        // Lm = L - a E s2
        Lb += Lmb;
        ab += Lmb * (-E * s2);
        Eb += Lmb * (-a * s2);
        s2b += Lmb * (-a * E);
        // W = (r2 + a2) E - a L
        double r2b = Wb * E;
        a2b += Wb * E;
        Eb += Wb * (r2 + a2);
        ab += Wb * (-L);
        Lb += Wb * (-a);
        // reciprocals
        const double Sigb = -iSb * iS * iS;
        Delb += -iDb * iD * iD;
        s2b += -is2b * is2 * is2;
        // Del = r2 - 2 M r + a2
        r2b += Delb;
        rb += Delb * (-2.0 * M);
        a2b += Delb;
        // Sig = r2 + a2 mu^2
        r2b += Sigb;
        a2b += Sigb * mu * mu;
        mub += Sigb * 2.0 * a2 * mu;
        // s2 = 1 - mu^2
        mub += s2b * (-2.0 * mu);
        // squares
        ab += a2b * 2.0 * a;
        rb += r2b * 2.0 * r;
        yb[0] += rb;
        yb[1] += mub;
        yb[3] += prb;
        yb[4] += pmb;
        cb[0] += Eb;
        cb[1] += Lb;
        cb[2] += ab;
    }
};

// ----------------------------------------------------------------- geometry
// Radius of a state (for the horizon, escape and FP64-region tests).
template <class Mt> PRR_HD double radius_of(const double *y) {
    if (Mt::spherical) return y[0];
    return sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
}

// Cartesian position (spherical-type coordinates map as flat space).
template <class Mt> PRR_HD D3 cart_of(const double *y) {
    if (!Mt::spherical) return d3(y[0], y[1], y[2]);
    double st, ct, sp, cp;
    if (Mt::hamiltonian) { // y[1] = cos th
        ct = y[1];
        st = sqrt(fmax(1.0 - ct * ct, 0.0));
    } else {
        sincos(y[1], &st, &ct);
    }
    sincos(y[2], &sp, &cp);
    return d3(y[0] * st * cp, y[0] * st * sp, y[0] * ct);
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
PRR_HD void ray_uv(const FrameDev &fr, int64_t local, double &u, double &v) {
    const int64_t pix = local / fr.s;
    const int k = int(local - pix * fr.s);
    const int j = int(pix / fr.W), i = int(pix - int64_t(j) * fr.W);
    if (fr.n == 0) {
        u = (i + 0.5) / fr.W;
        v = (j + 0.5) / fr.H;
    } else {
        const int q = k / fr.n, p = k - q * fr.n;
        u = (i + double(p + 1) / (fr.n + 1)) / fr.W;
        v = (j + double(q + 1) / (fr.n + 1)) / fr.H;
    }
}

PRR_HD D3 camera_ray(const FrameDev &fr, int64_t local) {
    double u, v;
    ray_uv(fr, local, u, v);
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

// (x, u) in the metric's coordinates -> integrator state y and per-ray
// constants.  For the 8-component metrics y = (x, u).
template <class Mt> PRR_HD void state_from_xu(const double x[4], const double u[4], double M, double a, double *y,
                                              Consts<double> &cs) {
    cs.M = M;
    cs.a = a;
    cs.E = cs.L = 0.0;
    if (!Mt::hamiltonian) {
        for (int i = 0; i < 4; ++i) {
            y[i] = x[i];
            y[4 + i] = u[i];
        }
        return;
    }
    const KerrG g = kerr_g(M, a, x[0], x[1]);
    cs.E = -(g.tt * u[3] + g.tph * u[2]);
    cs.L = g.phph * u[2] + g.tph * u[3];
    double st, ct;
    sincos(x[1], &st, &ct);
    y[0] = x[0];
    y[1] = ct;
    y[2] = x[2];
    y[3] = g.rr * u[0];
    y[4] = -g.thth * u[1] / st; // p_mu = -p_th / sin th
}

// Integrator state -> (x, u); for Kerr, t is returned as 0.
template <class Mt> PRR_HD void xu_from_state(const double *y, const Consts<double> &cs, double x[4], double u[4]) {
    if (!Mt::hamiltonian) {
        for (int i = 0; i < 4; ++i) {
            x[i] = y[i];
            u[i] = y[4 + i];
        }
        return;
    }
    double f[5];
    Kerr::rhs<double>(y, cs, f);
    const double th = acos(fmin(fmax(y[1], -1.0), 1.0));
    const KerrG g = kerr_g(cs.M, cs.a, y[0], th);
    const double det = g.tt * g.phph - g.tph * g.tph;
    x[0] = y[0];
    x[1] = th;
    x[2] = y[2];
    x[3] = 0.0;
    u[0] = f[0];
    u[1] = -sqrt(fmax(1.0 - y[1] * y[1], 0.0)) * y[4] / g.Sig; // th' = -sin th p_mu / Sigma
    u[2] = f[2];
    u[3] = -(g.phph * cs.E + g.tph * cs.L) / det;
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
    } else if (Mt::kind == PRR_SCHW_SPH) {
        const double f = 1.0 - 2.0 * M / x[0], s = sin(x[1]);
        A = -f;
        B = 0.0;
        C = u[0] * u[0] / f + x[0] * x[0] * (u[1] * u[1] + s * s * u[2] * u[2]);
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

// Direction used for the celestial lookup at escape: the spatial velocity for
// Cartesian metrics (getColorDefault2, main/base.h:119-128), the position for
// Schwarzschild coordinates (getColorDefault, :68-84), the Cartesian
// coordinate velocity for Kerr.
template <class Mt> PRR_HD D3 escape_dir(const double *y, const Consts<double> &cs) {
    if (Mt::kind == PRR_SCHW_SPH) {
        double st, ct, sp, cp;
        sincos(y[1], &st, &ct);
        sincos(y[2], &sp, &cp);
        return d3(st * cp, st * sp, ct);
    }
    if (Mt::hamiltonian) {
        double f[5];
        Kerr::rhs<double>(y, cs, f);
        double sp, cp;
        const double ct = y[1], st = sqrt(fmax(1.0 - ct * ct, 0.0));
        sincos(y[2], &sp, &cp);
        const double r = y[0];
        const double Sig = r * r + cs.a * cs.a * ct * ct;
        const double thdot = -st * y[4] / Sig;
        const D3 er = d3(st * cp, st * sp, ct), et = d3(ct * cp, ct * sp, -st), ep = d3(-sp, cp, 0.0);
        return normalize3(f[0] * er + (r * thdot) * et + (r * st * f[2]) * ep);
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

} // namespace prr
