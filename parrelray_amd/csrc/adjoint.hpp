// adjoint.hpp -- pieces of the Kerr adjoint (gradient of an image loss with
// respect to spin a, camera radius r_cam, inclination i and the sphere's
// orbital angle phi_s), shared by the adjoint kernels and host evaluation.
//
// Per ray the render is the map  p -> (y_0, c) = init(p) -> y_{N-1} =
// Phi_{N-2} o ... o Phi_0 (y_0) -> y* = Phi(y_{N-1}; tau) -> colour C(y*, c,
// phi_s), with y the Mino-time state (u, u', mu, mu', phi), c = (E, L, Q, a)
// the per-ray constants, Phi_n one accepted RKF45 step with its size frozen,
// and tau the located sub-step of the terminal event G(y*, phi_s) = 0
// (escape: G = 1/r_esc - u; sphere: G = |P(y*) - c_s|^2 - R^2).  The adjoint
//   lambda* = w . dC/dy*,  kappa = (lambda* . f*) / (grad G . f*),
//   lambda_{N-1} = (lambda* - kappa grad G) . dPhi(.; tau)/dy,
//   lambda_n = lambda_{n+1} . dPhi_n/dy_n,
//   dL/dp = lambda_0 . dy_0/dp + lambda_c . dc/dp (+ dC/dphi_s - kappa dG/dphi_s)
// (f* = f(y*) = dPhi/dtau; the kappa term is the implicit derivative of tau)
// uses a hand-written reverse sweep through the six RKF stages
// (kerr_step_vjp, built on Kerr::rhs_vjp) and forward-mode duals for the
// once-per-ray Jacobians dy_0/dp and dC/dy*.
#pragma once

#include "dual.hpp"
#include "lane.hpp"

namespace prr {

template <class T> struct V3 {
    T x, y, z;
};
template <class T> PRR_HD V3<T> v3t(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> PRR_HD V3<T> vadd(const V3<T> &a, const V3<T> &b) { return v3t<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> PRR_HD V3<T> vsub(const V3<T> &a, const V3<T> &b) { return v3t<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T, class S> PRR_HD V3<T> vscale(const S &s, const V3<T> &a) { return v3t<T>(a.x * s, a.y * s, a.z * s); }
template <class T> PRR_HD T vdot(const V3<T> &a, const V3<T> &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> PRR_HD V3<T> vnormalize(const V3<T> &a) {
    const T k = T(1.0) / sqrt(vdot(a, a));
    return vscale(k, a);
}
template <class T> PRR_HD V3<T> vcross(const V3<T> &a, const V3<T> &b) {
    return v3t<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Camera of an orbiting frame (prr_frame_from_orbit): position, then
// Camera::computeVectors (main/Camera.cpp:19-28) and rayDir_PinHole
// (main/Camera.h:30-33) for screen position (u, v).
template <class T>
PRR_HD void orbit_camera_ray(const T &r_cam, const T &incl, double cam_phi, double tan_half, double aspect, double u,
                             double v, V3<T> &pos, V3<T> &d) {
    T si, ci;
    sincos(incl, &si, &ci);
    double sp, cp;
    ::sincos(cam_phi, &sp, &cp);
    pos = v3t<T>(r_cam * si * cp, r_cam * si * sp, r_cam * ci);
    // alpha = cam_phi + pi, beta = pi/2 - incl: cos b = sin i, sin b = cos i
    const double ca = -cp, sa = -sp;
    const V3<T> F = vnormalize(v3t<T>(si * ca, si * sa, T(0.0) - ci));
    const V3<T> U = vnormalize(v3t<T>(ci * ca, ci * sa, si));
    const V3<T> R = vnormalize(vcross(F, U));
    const double cr = (2.0 * u - 1.0) * tan_half * aspect, cu = (1.0 - 2.0 * v) * tan_half;
    d = vnormalize(vadd(F, vadd(vscale(cr, R), vscale(cu, U))));
}

// init_state<Kerr> in scalar type T: z = (u, u', mu, mu', phi, E, L, Q).
template <class T> PRR_HD void kerr_init_t(const V3<T> &p, const V3<T> &d, double M, const T &a, T z[8]) {
    const T x2y2 = p.x * p.x + p.y * p.y, r2 = x2y2 + p.z * p.z, r = sqrt(r2);
    const T mu = p.z / r;
    const T ph = atan2(p.y, p.x);
    const T ur = vdot(p, d) / r;
    const T uth = (p.z * (p.x * d.x + p.y * d.y) - x2y2 * d.z) / (r2 * sqrt(x2y2));
    const T uph = (p.x * d.y - p.y * d.x) / x2y2;
    const T s2 = T(1.0) - mu * mu, a2 = a * a;
    const T Sig = r2 + a2 * mu * mu;
    const T Del = r2 - T(2.0 * M) * r + a2;
    const T g_rr = Sig / Del, g_thth = Sig;
    const T g_phph = (r2 + a2 + T(2.0 * M) * r * a2 * s2 / Sig) * s2;
    const T g_tph = T(-2.0 * M) * r * a * s2 / Sig;
    const T g_tt = T(0.0) - (T(1.0) - T(2.0 * M) * r / Sig);
    // larger root of g_tt ut^2 + 2 g_tph uph ut + (g_rr ur^2 + g_thth uth^2 + g_phph uph^2) = 0
    const T B = T(2.0) * g_tph * uph;
    const T C = g_rr * ur * ur + g_thth * uth * uth + g_phph * uph * uph;
    const T rt = sqrt(B * B - T(4.0) * g_tt * C);
    const T u1 = (T(0.0) - B + rt) / (T(2.0) * g_tt), u2 = (T(0.0) - B - rt) / (T(2.0) * g_tt);
    const T ut = value(u1) >= value(u2) ? u1 : u2;
    const T E = T(0.0) - (g_tt * ut + g_tph * uph);
    const T L = g_phph * uph + g_tph * ut;
    const T pth = g_thth * uth;
    z[0] = T(1.0) / r;
    z[1] = T(0.0) - Sig * ur / r2;
    z[2] = mu;
    z[3] = T(0.0) - sqrt(s2) * pth;
    z[4] = ph;
    z[5] = E;
    z[6] = L;
    z[7] = pth * pth + mu * mu * (L * L / s2 - a2 * E * E);
}

// escape_dir<Kerr> in scalar type T from the state y[5] and (E, L, Q, a).
template <class T>
PRR_HD V3<T> kerr_escape_dir_t(const T y[5], const T &E, const T &L, const T &Q, const T &a, double M) {
    T yy[6] = {y[0], y[1], y[2], y[3], y[4], T(0.0)}, f[6];
    Kerr::rhs<T>(yy, Kerr::coef_of<T>(M, a, E, L, Q), f);
    const T u = y[0], ct = y[2], st = sqrt(T(1.0) - ct * ct);
    T sp, cp;
    sincos(y[4], &sp, &cp);
    const V3<T> er = v3t<T>(st * cp, st * sp, ct), et = v3t<T>(ct * cp, ct * sp, T(0.0) - st);
    const V3<T> ep = v3t<T>(T(0.0) - sp, cp, T(0.0));
    return vnormalize(vadd(vscale(T(0.0) - y[1], er), vadd(vscale(T(0.0) - u * y[3] / st, et), vscale(u * st * f[4], ep))));
}

// Bilinear lookup (same cell selection as `bilinear`) in scalar type T:
// the texel values are constants, the weights carry the derivatives.
template <class T> PRR_HD void bilinear_t(const TexDev &t, T u, T v, T out[3]) {
    if (t.w <= 0 || !(value(u) == value(u)) || !(value(v) == value(v))) {
        out[0] = out[1] = out[2] = T(0.0);
        return;
    }
    u = u - T(floor(value(u)));
    v = fmin(fmax(v, 0.0), 1.0);
    const T X = u * double(t.w), Y = v * double(t.h);
    int i0 = int(floor(value(X))), j0 = int(floor(value(Y)));
    const T fx = X - double(i0);
    T fy = Y - double(j0);
    if (i0 >= t.w) i0 -= t.w;
    const int i1 = i0 + 1 >= t.w ? 0 : i0 + 1;
    if (j0 >= t.h - 1) {
        j0 = t.h - 1;
        fy = T(0.0);
    }
    const int j1 = j0 + 1 < t.h ? j0 + 1 : t.h - 1;
    const float4 c00 = texel(t, i0, j0), c10 = texel(t, i1, j0), c01 = texel(t, i0, j1), c11 = texel(t, i1, j1);
    const T w00 = (T(1.0) - fx) * (T(1.0) - fy), w10 = fx * (T(1.0) - fy), w01 = (T(1.0) - fx) * fy, w11 = fx * fy;
    out[0] = w00 * double(c00.x) + w10 * double(c10.x) + w01 * double(c01.x) + w11 * double(c11.x);
    out[1] = w00 * double(c00.y) + w10 * double(c10.y) + w01 * double(c01.y) + w11 * double(c11.y);
    out[2] = w00 * double(c00.z) + w10 * double(c10.z) + w01 * double(c01.z) + w11 * double(c11.z);
}

// Sky colour of an escaping ray (bilinear filter; SQAVG is piecewise constant
// and has zero derivative).
template <class T> PRR_HD void sky_color_t(const TexDev &t, const V3<T> &n, T out[3]) {
    const double pi = 3.14159265358979323846;
    const T th = acos(fmin(fmax(n.z, -1.0), 1.0));
    T ph = atan2(n.y, n.x);
    if (value(ph) < 0.0) ph = ph + 2.0 * pi;
    if (t.filter != PRR_FILTER_BILINEAR) {
        const float3 c = sky_lookup(t, value(th), value(ph));
        out[0] = T(double(c.x));
        out[1] = T(double(c.y));
        out[2] = T(double(c.z));
        return;
    }
    bilinear_t<T>(t, ph / (2.0 * pi), th / pi, out);
}

// Sphere::getColor (scene/sphere.h:59-64) at the outward unit normal n, in
// scalar type T (bilinear texture filter).
template <class T> PRR_HD void sphere_color_t(const TexDev *t, const V3<T> &n, T out[3]) {
    const double pi = 3.14159265358979323846;
    const T th = acos(fmin(fmax(n.z, -1.0), 1.0));
    const T ph = atan2(n.y, n.x) + pi;
    T col[3];
    if (t && t->w) bilinear_t<T>(*t, ph / (2.0 * pi), th / pi, col);
    else col[0] = col[1] = col[2] = T(1.0);
    T shade = T(0.5) + 0.5 * n.z;
    shade = fmin(fmax(shade, 0.0), 1.0);
    for (int k = 0; k < 3; ++k) out[k] = col[k] * shade;
}

// cart_of<Kerr> in scalar type T.
template <class T> PRR_HD V3<T> kerr_cart_t(const T y[5]) {
    const T r = T(1.0) / y[0], ct = y[2], st = sqrt(T(1.0) - ct * ct);
    T sp, cp;
    sincos(y[4], &sp, &cp);
    return v3t<T>(r * st * cp, r * st * sp, r * ct);
}

// One accepted RKF45 step y -> y + sum b4_s k_s (FP64, step size h frozen).
PRR_HD void kerr_step(const double y[6], double h, const Consts<double> &cs, double out[6]) {
    const KerrCoef<double> kc = Kerr::coef<double>(cs);
    double k[6][6];
    for (int s = 0; s < 6; ++s) {
        double ys[6];
        for (int i = 0; i < 6; ++i) {
            double acc = y[i];
            for (int j = 0; j < s; ++j) acc += Fehl::a(s, j) * k[j][i];
            ys[i] = acc;
        }
        double f[6];
        Kerr::rhs<double>(ys, kc, f);
        for (int i = 0; i < 6; ++i) k[s][i] = h * f[i];
    }
    for (int i = 0; i < 6; ++i) {
        double acc = 0.0;
        for (int s = 0; s < 6; ++s) acc += Fehl::b4(s) * k[s][i];
        out[i] = y[i] + acc;
    }
}

// lam[9] = adjoint of (y_{n+1}[0..4], E, L, Q, a) on entry, of (y_n, E, L,
// Q, a) on exit, through one accepted step of size h (reverse sweep over the
// stages; lambda = y[5] feeds nothing back and is not carried).
PRR_HD void kerr_step_vjp(const double y[6], double h, const Consts<double> &cs, double lam[9]) {
    const KerrCoef<double> kc = Kerr::coef<double>(cs);
    // forward: the stage states; the right-hand side reads components 0..3
    // only, so phi's and lambda's stage values are not formed, and the
    // adjoint of phi passes through unchanged (d f / d phi = 0)
    double Y[6][4];
    {
        double k[6][4];
        for (int s = 0; s < 6; ++s) {
            double ys[6];
            for (int i = 0; i < 4; ++i) {
                double acc = y[i];
                for (int j = 0; j < s; ++j) acc += Fehl::a(s, j) * k[j][i];
                ys[i] = acc;
                Y[s][i] = acc;
            }
            ys[4] = ys[5] = 0.0;
            double f[6];
            Kerr::rhs<double>(ys, kc, f);
            for (int i = 0; i < 4; ++i) k[s][i] = h * f[i];
        }
    }
    double lk[6][4];
    for (int s = 0; s < 6; ++s)
        for (int i = 0; i < 4; ++i) lk[s][i] = Fehl::b4(s) * lam[i];
    double ly[4] = {lam[0], lam[1], lam[2], lam[3]};
    KerrCoef<double> kb{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int s = 5; s >= 0; --s) {
        double ys[6] = {Y[s][0], Y[s][1], Y[s][2], Y[s][3], 0.0, 0.0};
        double fb[6], yb[6] = {0, 0, 0, 0, 0, 0};
        for (int i = 0; i < 4; ++i) fb[i] = h * lk[s][i];
        fb[4] = h * (Fehl::b4(s) * lam[4]);
        fb[5] = 0.0;
        Kerr::rhs_vjp(ys, kc, fb, yb, kb);
        for (int i = 0; i < 4; ++i) {
            ly[i] += yb[i];
            for (int j = 0; j < s; ++j) lk[j][i] += Fehl::a(s, j) * yb[i];
        }
    }
    for (int i = 0; i < 4; ++i) lam[i] = ly[i];
    Kerr::coef_vjp(cs, kb, lam + 5);
}

} // namespace prr
