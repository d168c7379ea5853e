// lane.hpp -- the per-ray integrator: set-up, one Euler step / RKF45
// attempt, termination events and output, for one ray held by one lane.
//
// The persistent kernels (kernels.hip) call these functions once per loop
// trip; they are __host__ __device__ so that the same code also runs on the
// host for selected rays (prr_debug_trace_rays).
//
// Replaces trace_Euler / trace_RKF (kernels/solver.cuh:68-220) and the
// per-step User_UPDATE / User_ENDLOOP hooks (metrics/schwarzschild.cuh:55-79).
// For the adaptive solver the two events whose position depends on the step
// sequence are located on the integrated path instead of at the first step
// beyond them:
//   * escape: the state where r crosses r_escape (Newton on the crossing
//     fraction theta of the last step, each iterate a fresh FP64 RK45 step of
//     theta h from the step's start);
//   * sphere hit: the first point of the path inside a sphere (bracketed
//     Illinois iteration on |P(theta) - c|^2 - R^2 over the same FP64
//     sub-steps), for steps whose chord passes within a margin of a sphere.
// The fixed-step Euler solver keeps the reference's chord and first-step
// semantics.
#pragma once

#include "geodesic.hpp"

namespace prr {

// Fehlberg 4(5) tableau (kernels/solver.cuh:125-131; Fehlberg 1969, NASA TR R-315).
struct Fehl {
    static PRR_HD constexpr double a(int s, int j) {
        constexpr double A[6][5] = {{0, 0, 0, 0, 0},
                                    {1.0 / 4.0, 0, 0, 0, 0},
                                    {3.0 / 32.0, 9.0 / 32.0, 0, 0, 0},
                                    {1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0, 0, 0},
                                    {439.0 / 216.0, -8.0, 3680.0 / 513.0, -845.0 / 4104.0, 0},
                                    {-8.0 / 27.0, 2.0, -3544.0 / 2565.0, 1859.0 / 4104.0, -11.0 / 40.0}};
        return A[s][j];
    }
    // c_s = sum_j a_sj
    static PRR_HD constexpr double c(int s) {
        constexpr double C[6] = {0.0, 1.0 / 4.0, 3.0 / 8.0, 12.0 / 13.0, 1.0, 1.0 / 2.0};
        return C[s];
    }
    static PRR_HD constexpr double b4(int s) {
        constexpr double B[6] = {25.0 / 216.0, 0.0, 1408.0 / 2565.0, 2197.0 / 4104.0, -1.0 / 5.0, 0.0};
        return B[s];
    }
    // b5 - b4: the embedded error estimate
    static PRR_HD constexpr double e(int s) {
        constexpr double E[6] = {16.0 / 135.0 - 25.0 / 216.0,       0.0,
                                 6656.0 / 12825.0 - 1408.0 / 2565.0, 28561.0 / 56430.0 - 2197.0 / 4104.0,
                                 -9.0 / 50.0 + 1.0 / 5.0,            2.0 / 55.0};
        return E[s];
    }
};

// Fixed-size state vector passed by value to out-of-line functions (so that
// the caller's lane state never has its address taken and stays in VGPRs).
template <int N> struct Vec {
    double v[N];
};

// Located event: state, fraction of the step, sphere index.
template <int N> struct Located {
    Vec<N> y;
    double theta;
    int k;
};

// 1 - cos^2 th below which Kerr's FP32-bulk attempts run their stages in FP64.
constexpr double kAxisFP64 = 0.05;
// Largest turn of a Kerr step, in units of the ray's angular distance from the polar axis.
constexpr double kAxisStep = 0.25;

// Lane state of one ray.
template <class Mt> struct Lane {
    double y[Mt::N];
    Consts<double> cs;
    double h;   // next step size in the integration variable (Mino time for Kerr)
    double lam; // affine parameter
    int32_t steps, attempts, att64;
    int32_t frame;
    int32_t hit; // sphere index of a SPHERE termination
    int64_t rid;
    float P0[3]; // FP32 Cartesian position at the start of the step (sphere screen)
};

PRR_HD int find_frame(const JobDev &job, int64_t id) {
    int lo = 0, hi = job.nframes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (job.frames[mid].ray_begin <= id) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <class Mt> PRR_HD bool finite_state(const double *y) {
    bool ok = true;
#pragma unroll
    for (int i = 0; i < Mt::N; ++i) ok = ok && __builtin_isfinite(y[i]);
    return ok;
}

// Sigma = dlambda/ds of the Mino-time form; 1 for the affine metrics.
// Contraction is off here and in rkf_step_size so that the adjoint's replay
// recomputes the step sizes bit for bit wherever these are inlined.
template <class Mt> PRR_HD double lambda_rate(const double *y, const Consts<double> &cs) {
#pragma clang fp contract(off)
    if (!Mt::mino) return 1.0;
    return 1.0 / (y[0] * y[0]) + cs.a * cs.a * y[2] * y[2];
}

// Step size of an RKF45 attempt from the proposal h: the affine bounds
// [hmin, hmax] map to [hmin, hmax] / Sigma in Mino time.  at_min: no
// rejection (the reference's h == hmin, kernels/solver.cuh:185).
template <class Mt>
PRR_HD double rkf_step_size(const TracerDev &tr, const double *y, const Consts<double> &cs, double h, bool &at_min) {
#pragma clang fp contract(off)
    if (!Mt::mino) {
        at_min = h == tr.hmin;
        return h;
    }
    const double iS = 1.0 / lambda_rate<Mt>(y, cs);
    const double lo = tr.hmin * iS;
    // near the polar axis a step may turn the ray by at most kAxisStep times
    // its angular distance from the axis (sin th): the L / sin^2 th spike of
    // phi' is then resolved by several steps, which the embedded error
    // estimate alone does not guarantee (angular speed sqrt(Q + L^2 + a^2 E^2 mu^2))
    const double s2 = 1.0 - y[2] * y[2];
    const double K = cs.Q + cs.L * cs.L + cs.a * cs.a * cs.E * cs.E * y[2] * y[2];
    const double hax = kAxisStep * sqrt(fmax(s2, 0.0) / K);
    h = fmin(fmax(h, lo), fmin(tr.hmax * iS, fmax(hax, lo)));
    at_min = h <= lo;
    return h;
}

// Six RK stages in precision T from the FP64 state y (plain or FP32 form of
// the 8-component metrics, and Kerr's FP64 form); the 4th-order increment d4
// and the error vector e in FP64.
template <class Mt, class T>
PRR_HD void rkf_stages(const double *y, double h, const Consts<double> &cd, double *d4, double *e) {
    constexpr int N = Mt::N;
    const auto c = Mt::template coef<T>(cd);
    const T hT = T(h);
    T yT[N];
#pragma unroll
    for (int i = 0; i < N; ++i) yT[i] = T(y[i]);
    T k[6][N];
    Mt::template rhs<T>(yT, c, k[0]);
#pragma unroll
    for (int i = 0; i < N; ++i) k[0][i] *= hT;
#pragma unroll
    for (int s = 1; s < 6; ++s) {
        T ys[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            T acc = yT[i];
#pragma unroll
            for (int j = 0; j < s; ++j) acc += T(Fehl::a(s, j)) * k[j][i];
            ys[i] = acc;
        }
        Mt::template rhs<T>(ys, c, k[s]);
#pragma unroll
        for (int i = 0; i < N; ++i) k[s][i] *= hT;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double a4 = 0.0, ae = 0.0;
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            if (Fehl::b4(s) != 0.0) a4 += Fehl::b4(s) * double(k[s][i]);
            if (Fehl::e(s) != 0.0) ae += Fehl::e(s) * double(k[s][i]);
        }
        d4[i] = a4;
        e[i] = ae;
    }
}

// Kerr's FP32-bulk attempt: stage 1 and the difference coefficients in FP64,
// stages 2-6 evaluate g_s = f(y + d_s) - f(y) in FP32 with
// d_s = h (c_s F + sum_j a_sj g_j).  Since sum b4 = 1 and sum (b5 - b4) = 0,
//   d4 = h F + h sum b4_s g_s,   e = h sum (b5 - b4)_s g_s,
// so FP32 rounding enters only through the small differences g_s and the
// error estimate never sees the rounding of f itself.
PRR_HD void kerr_stages_fp32(const double *y, double h, const Consts<double> &cd, double *d4, double *e) {
    constexpr int N = Kerr::N;
    double F[N];
    KerrDelta P;
    Kerr::delta_prep(y, Kerr::coef<double>(cd), F, P);
    const float hf = float(h);
    float Ff[N];
#pragma unroll
    for (int i = 0; i < N; ++i) Ff[i] = float(F[i]);
    float g[6][N];
#pragma unroll
    for (int i = 0; i < N; ++i) g[0][i] = 0.0f;
#pragma unroll
    for (int s = 1; s < 6; ++s) {
        float d[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            float acc = float(Fehl::c(s)) * Ff[i];
#pragma unroll
            for (int j = 1; j < s; ++j) acc += float(Fehl::a(s, j)) * g[j][i];
            d[i] = hf * acc;
        }
        Kerr::delta_rhs(d, P, g[s]);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        float a4 = 0.0f, ae = 0.0f;
#pragma unroll
        for (int s = 1; s < 6; ++s) {
            if (Fehl::b4(s) != 0.0) a4 += float(Fehl::b4(s)) * g[s][i];
            if (Fehl::e(s) != 0.0) ae += float(Fehl::e(s)) * g[s][i];
        }
        d4[i] = h * F[i] + double(hf * a4);
        e[i] = double(hf * ae);
    }
}

template <class Mt, class T> PRR_HD void euler_rhs(const double *y, const Consts<double> &cd, double *f) {
    constexpr int N = Mt::N;
    const auto c = Mt::template coef<T>(cd);
    T yT[N], fT[N];
#pragma unroll
    for (int i = 0; i < N; ++i) yT[i] = T(y[i]);
    Mt::template rhs<T>(yT, c, fT);
#pragma unroll
    for (int i = 0; i < N; ++i) f[i] = double(fT[i]);
}

// One FP64 RK45 step (4th-order solution) of size h from y: the sub-steps of
// the event location.  Out of line: cold code.
template <class Mt>
PRR_HD __noinline__ void rk_substep(const double *y, double h, const Consts<double> &cs, double *out) {
    double d4[Mt::N], e[Mt::N];
    rkf_stages<Mt, double>(y, h, cs, d4, e);
    for (int i = 0; i < Mt::N; ++i) out[i] = y[i] + d4[i];
}

// Escape crossing function g (g < 0 inside r_escape) and dg/dt along the
// integration variable.
template <class Mt> PRR_HD void escape_g(const double *y, double r_esc, double &g, double &dg) {
    if (Mt::mino) {
        g = 1.0 / r_esc - y[0];
        dg = -y[1];
    } else if (Mt::spherical) {
        g = y[0] - r_esc;
        dg = y[4];
    } else {
        const double r = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
        g = r - r_esc;
        dg = (y[0] * y[4] + y[1] * y[5] + y[2] * y[6]) / r;
    }
}

// State at the r = r_escape crossing inside the step yo -> yn of size h
// (theta = the crossing fraction).
template <class Mt>
PRR_HD __noinline__ Located<Mt::N> escape_locate(const Vec<Mt::N> yo, const Vec<Mt::N> yn, double h,
                                                 const Consts<double> cs, double r_esc) {
    Located<Mt::N> out;
    out.k = -1;
    double go, gn, dg;
    escape_g<Mt>(yo.v, r_esc, go, dg);
    escape_g<Mt>(yn.v, r_esc, gn, dg);
    if (!(go < 0.0)) { // started outside: nothing to locate
        out.y = yo;
        out.theta = 0.0;
        return out;
    }
    double th = go / (go - gn);
    th = fmin(fmax(th, 0.0), 1.0);
    const double tol = 1e-13 * (Mt::mino ? 1.0 / r_esc : r_esc);
    for (int it = 0; it < 6; ++it) {
        rk_substep<Mt>(yo.v, th * h, cs, out.y.v);
        double g;
        escape_g<Mt>(out.y.v, r_esc, g, dg);
        if (fabs(g) <= tol || !(dg * h != 0.0)) break;
        th = fmin(fmax(th - g / (dg * h), 0.0), 1.0);
        if (it == 5) rk_substep<Mt>(yo.v, th * h, cs, out.y.v);
    }
    out.theta = th;
    return out;
}

// Sphere k: G = |P - c|^2 - R^2 at a state.
template <class Mt> PRR_HD double sphere_G(const FrameDev &fr, int k, const double *y) {
    const D3 m = cart_of<Mt>(y) - d3(fr.sph[k][0], fr.sph[k][1], fr.sph[k][2]);
    return dot3(m, m) - fr.sph[k][3] * fr.sph[k][3];
}

// FP32 screen: can the path of this step come near sphere k?  The chord's
// closest distance to the centre against R plus half the chord length (a
// bound on the path's bulge for any step that turns by less than ~1 rad)
// plus a relative margin for FP32 rounding.
PRR_HD bool sphere_screen(const FrameDev &fr, int k, const float P0[3], const float P1[3]) {
    const float cx = float(fr.sph[k][0]), cy = float(fr.sph[k][1]), cz = float(fr.sph[k][2]);
    const float R = float(fr.sph[k][3]);
    const float mx = P0[0] - cx, my = P0[1] - cy, mz = P0[2] - cz;
    const float dx = P1[0] - P0[0], dy = P1[1] - P0[1], dz = P1[2] - P0[2];
    const float A = dx * dx + dy * dy + dz * dz;
    float t = A > 0.0f ? -(mx * dx + my * dy + mz * dz) / A : 0.0f;
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    const float qx = mx + t * dx, qy = my + t * dy, qz = mz + t * dz;
    const float dist = sqrtf(qx * qx + qy * qy + qz * qz);
    const float scale = fabsf(cx) + fabsf(cy) + fabsf(cz) + R + fabsf(P0[0]) + fabsf(P0[1]) + fabsf(P0[2]);
    return dist <= R + 0.5f * sqrtf(A) + 1e-4f * scale;
}

// First point of the step yo -> yn (size h) inside a screened sphere: the
// sphere index k (or -1), the state there and the fraction theta of the step.
template <class Mt>
PRR_HD __noinline__ Located<Mt::N> sphere_locate(const FrameDev &fr, unsigned mask, const Vec<Mt::N> yo,
                                                 const Vec<Mt::N> yn, double h, const Consts<double> cs) {
    constexpr int N = Mt::N;
    Located<N> out;
    out.k = -1;
    out.theta = 2.0;
    double yt[N];
    const D3 P0 = cart_of<Mt>(yo.v), P1 = cart_of<Mt>(yn.v);
    const D3 dP = P1 - P0;
    for (int k = 0; k < fr.n_spheres; ++k) {
        if (!((mask >> k) & 1u)) continue;
        const D3 c = d3(fr.sph[k][0], fr.sph[k][1], fr.sph[k][2]);
        const double R2 = fr.sph[k][3] * fr.sph[k][3];
        const D3 m = P0 - c;
        const double G0 = dot3(m, m) - R2;
        if (G0 <= 0.0) { // the step starts inside (reference: chord t = 0)
            if (0.0 < out.theta) {
                out.theta = 0.0;
                out.k = k;
                out.y = yo;
            }
            continue;
        }
        // the path's entry: a bracket [0, hi] with G(0) > 0 >= G(hi)
        double hi = -1.0, Ghi = 0.0;
        const double G1 = sphere_G<Mt>(fr, k, yn.v);
        if (G1 <= 0.0) {
            hi = 1.0;
            Ghi = G1;
            for (int i = 0; i < N; ++i) yt[i] = yn.v[i];
        } else {
            // interior dip: sample the path at the chord's closest point, then
            // refine the minimum by successive parabolic interpolation
            const double A = dot3(dP, dP);
            double t1 = A > 0.0 ? -dot3(m, dP) / A : 0.5;
            t1 = fmin(fmax(t1, 0.02), 0.98);
            double ta = 0.0, Ga = G0, tb = 1.0, Gb = G1;
            rk_substep<Mt>(yo.v, t1 * h, cs, yt);
            double G1s = sphere_G<Mt>(fr, k, yt);
            for (int it = 0; it < 4 && G1s > 0.0; ++it) {
                // vertex of the parabola through (ta, Ga), (t1, G1s), (tb, Gb)
                const double p = (t1 - ta) * (G1s - Gb), q = (t1 - tb) * (G1s - Ga);
                const double den = 2.0 * (p - q);
                if (!(den != 0.0)) break;
                double tv = t1 - ((t1 - ta) * p - (t1 - tb) * q) / den;
                if (!(tv > ta && tv < tb) || fabs(tv - t1) < 1e-9) break;
                double yv[N];
                rk_substep<Mt>(yo.v, tv * h, cs, yv);
                const double Gv = sphere_G<Mt>(fr, k, yv);
                if (Gv <= 0.0) {
                    t1 = tv;
                    G1s = Gv;
                    for (int i = 0; i < N; ++i) yt[i] = yv[i];
                    break;
                }
                if (Gv < G1s) { // tv becomes the middle point
                    if (tv < t1) {
                        tb = t1;
                        Gb = G1s;
                    } else {
                        ta = t1;
                        Ga = G1s;
                    }
                    t1 = tv;
                    G1s = Gv;
                } else if (tv < t1) {
                    ta = tv;
                    Ga = Gv;
                } else {
                    tb = tv;
                    Gb = Gv;
                }
            }
            if (G1s > 0.0) continue; // the path stays outside: a miss
            hi = t1;
            Ghi = G1s;
        }
        if (hi >= out.theta) continue;
        // Illinois regula falsi on [0, hi]
        double lo = 0.0, Glo = G0;
        int side = 0;
        double yb[N];
        for (int i = 0; i < N; ++i) yb[i] = yt[i];
        const double gtol = 1e-14 * (R2 + dot3(c, c));
        for (int it = 0; it < 60 && fabs(Ghi) > gtol && hi - lo > 1e-15; ++it) {
            double t = (lo * Ghi - hi * Glo) / (Ghi - Glo);
            if (!(t > lo && t < hi)) t = 0.5 * (lo + hi);
            rk_substep<Mt>(yo.v, t * h, cs, yt);
            const double Gt = sphere_G<Mt>(fr, k, yt);
            if (Gt > 0.0) {
                lo = t;
                Glo = Gt;
                if (side == -1) Ghi *= 0.5;
                side = -1;
            } else {
                hi = t;
                Ghi = Gt;
                for (int i = 0; i < N; ++i) yb[i] = yt[i];
                if (side == 1) Glo *= 0.5;
                side = 1;
            }
        }
        if (hi < out.theta) {
            out.theta = hi;
            out.k = k;
            for (int i = 0; i < N; ++i) out.y.v[i] = yb[i];
        }
    }
    return out;
}

// lane_start and lane_finish run once per ray; they are kept out of line so
// that their transcendental constants and temporaries do not occupy
// registers across the integration loop.  Out-of-line functions take and
// return lane data by value.
template <class Mt> PRR_HD __noinline__ Lane<Mt> lane_start(const JobDev &job, int64_t id) {
    Lane<Mt> L;
    L.rid = id;
    L.frame = job.explicit_x0u0 ? 0 : find_frame(job, id);
    const FrameDev &fr = job.frames[L.frame];
    if (job.explicit_x0u0) {
        const double *s = job.explicit_x0u0 + 8 * id;
        state_from_xu<Mt>(s, s + 4, fr.M, fr.a, L.y, L.cs);
    } else {
        const D3 d = camera_ray(fr, id - fr.ray_begin);
        init_state<Mt>(d3(fr.pos[0], fr.pos[1], fr.pos[2]), d, fr.M, fr.a, L.y, L.cs);
    }
    L.h = job.tr.h / lambda_rate<Mt>(L.y, L.cs);
    L.lam = 0.0;
    L.steps = L.attempts = L.att64 = 0;
    L.hit = -1;
    L.P0[0] = L.P0[1] = L.P0[2] = 0.0f;
    if (fr.n_spheres > 0) cart_of_f32<Mt>(L.y, L.P0);
    return L;
}

// Writes the outputs of a finished ray (rayToRGB, metrics/schwarzschild.cuh:151-183).
template <class Mt> PRR_HD __noinline__ void lane_finish(const JobDev &job, const Lane<Mt> L, int cls) {
    const int64_t id = L.rid;
    const FrameDev &fr = job.frames[L.frame];
    D3 endpoint;
    float3 color;
    if (cls == PRR_SPHERE) {
        const D3 m = cart_of<Mt>(L.y) - d3(fr.sph[L.hit][0], fr.sph[L.hit][1], fr.sph[L.hit][2]);
        endpoint = normalize3(m);
        const int slot = fr.sph_tex[L.hit];
        color = sphere_color((slot > 0 && slot < PRR_MAX_TEXTURES) ? &job.tex[slot] : nullptr, endpoint);
    } else if (cls == PRR_ESCAPE) {
        const D3 n = escape_dir<Mt>(L.y, L.cs);
        endpoint = n;
        double th, ph;
        if (Mt::kind == PRR_SCHW_SPH) {
            const double two_pi = 6.28318530717958647692, pi = 3.14159265358979323846;
            th = L.y[1];
            ph = L.y[2];
            if (ph == two_pi) ph = 0.0; // main/base.h:86-89
            if (th == pi) th = 0.0;
        } else {
            th = acos(fmin(fmax(n.z, -1.0), 1.0));
            ph = atan2(n.y, n.x);
            if (ph < 0.0) ph += 6.28318530717958647692;
        }
        color = sky_lookup(job.tex[0], th, ph);
    } else {
        endpoint = cart_of<Mt>(L.y);
        color = make_float3(0.f, 0.f, 0.f);
    }
    const OutDev &o = job.out;
    if (o.color) {
        o.color[3 * id] = color.x;
        o.color[3 * id + 1] = color.y;
        o.color[3 * id + 2] = color.z;
    }
    if (o.cls) o.cls[id] = int8_t(cls);
    if (o.endpoint) {
        o.endpoint[3 * id] = endpoint.x;
        o.endpoint[3 * id + 1] = endpoint.y;
        o.endpoint[3 * id + 2] = endpoint.z;
    }
    if (o.steps) o.steps[id] = L.steps;
    if (o.attempts) o.attempts[id] = L.attempts;
    if (o.state8) {
        double x[4], u[4];
        xu_from_state<Mt>(L.y, L.cs, x, u);
        for (int c = 0; c < 4; ++c) {
            o.state8[8 * id + c] = x[c];
            o.state8[8 * id + 4 + c] = u[c];
        }
    }
}

// Event record of a terminal RKF45 step (the adjoint's last step): the state
// at the step's start, the step size and the located fraction.
PRR_HD __noinline__ void record_event(EventDev *ev, const Vec<8> yo, double h, double theta, int cls, int hit) {
    for (int i = 0; i < 8; ++i) ev->y_prev[i] = yo.v[i];
    ev->h = h;
    ev->theta = theta;
    ev->cls = cls;
    ev->hit = hit;
}

template <int N> PRR_HD Vec<8> widen8(const double *y) {
    Vec<8> o;
    for (int i = 0; i < 8; ++i) o.v[i] = i < N ? y[i] : 0.0;
    return o;
}

template <int N> PRR_HD Vec<N> vec_of(const double *y) {
    Vec<N> o;
    for (int i = 0; i < N; ++i) o.v[i] = y[i];
    return o;
}

// Termination tests after an accepted Euler step (reference semantics: the
// chord P(yo) -> P(yn), spherical wraps, horizon, first state beyond
// r_escape).  yn becomes the lane state.
template <class Mt> PRR_HD int euler_events(const JobDev &job, Lane<Mt> &L, const double *yo, double *yn) {
    const FrameDev &fr = job.frames[L.frame];
    if (fr.n_spheres > 0) {
        const D3 P0 = cart_of<Mt>(yo), P1 = cart_of<Mt>(yn);
        double t;
        const int k = chord_hit(fr, P0, P1, t);
        if (k >= 0) {
            // the lane's position becomes the chord's hit point
            const D3 H = P0 + t * (P1 - P0);
            for (int i = 0; i < Mt::N; ++i) L.y[i] = yn[i];
            if (Mt::spherical) {
                const double r = sqrt(dot3(H, H));
                if (Mt::mino) {
                    L.y[0] = 1.0 / r;
                    L.y[2] = H.z / r;
                    L.y[4] = atan2(H.y, H.x);
                } else {
                    L.y[0] = r;
                    L.y[1] = acos(H.z / r);
                    L.y[2] = atan2(H.y, H.x);
                }
            } else {
                L.y[0] = H.x;
                L.y[1] = H.y;
                L.y[2] = H.z;
            }
            L.hit = k;
            return PRR_SPHERE;
        }
    }
    if (Mt::kind == PRR_SCHW_SPH) sph_post_step(yn);
    for (int i = 0; i < Mt::N; ++i) L.y[i] = yn[i];
    const double r = radius_of<Mt>(yn);
    if (r <= fr.r_hor) return PRR_HORIZON;
    if (r >= job.tr.r_escape) return PRR_ESCAPE;
    return -1;
}

// Termination tests after an accepted RKF45 step yo -> yn of size h, with
// path-located sphere hits and escape crossings.
template <class Mt>
PRR_HD int rkf_events(const JobDev &job, Lane<Mt> &L, const double *yo, double *yn, double h) {
    const FrameDev &fr = job.frames[L.frame];
    if (fr.n_spheres > 0) {
        float P1[3];
        cart_of_f32<Mt>(yn, P1);
        unsigned mask = 0u;
        for (int k = 0; k < fr.n_spheres; ++k)
            if (sphere_screen(fr, k, L.P0, P1)) mask |= 1u << k;
        if (mask) {
            const Located<Mt::N> ev =
                sphere_locate<Mt>(fr, mask, vec_of<Mt::N>(yo), vec_of<Mt::N>(yn), h, L.cs);
            if (ev.k >= 0) {
                for (int i = 0; i < Mt::N; ++i) L.y[i] = ev.y.v[i];
                L.hit = ev.k;
                if (job.out.event)
                    record_event(job.out.event + L.rid, widen8<Mt::N>(yo), h, ev.theta, PRR_SPHERE, ev.k);
                return PRR_SPHERE;
            }
        }
        L.P0[0] = P1[0];
        L.P0[1] = P1[1];
        L.P0[2] = P1[2];
    }
    const double r = radius_of<Mt>(yn);
    if (r >= job.tr.r_escape) {
        const Located<Mt::N> ev = escape_locate<Mt>(vec_of<Mt::N>(yo), vec_of<Mt::N>(yn), h, L.cs, job.tr.r_escape);
        for (int i = 0; i < Mt::N; ++i) L.y[i] = ev.y.v[i];
        if (Mt::kind == PRR_SCHW_SPH) sph_post_step(L.y);
        if (job.out.event) record_event(job.out.event + L.rid, widen8<Mt::N>(yo), h, ev.theta, PRR_ESCAPE, -1);
        return PRR_ESCAPE;
    }
    if (Mt::kind == PRR_SCHW_SPH) {
        sph_post_step(yn);
        if (fr.n_spheres > 0) cart_of_f32<Mt>(yn, L.P0);
    }
    for (int i = 0; i < Mt::N; ++i) L.y[i] = yn[i];
    if (r <= fr.r_hor) return PRR_HORIZON;
    return -1;
}

// One Euler step or one RKF45 attempt.  Returns the class when the ray ends,
// -1 otherwise.
template <class Mt, int SOLVER, bool MIXED> PRR_HD int lane_attempt(const JobDev &job, Lane<Mt> &L) {
    constexpr int N = Mt::N;
    const TracerDev &tr = job.tr;
    if (L.steps >= tr.max_steps || L.lam >= tr.lambda_max) return PRR_STEP_CAP;
    const double r = radius_of<Mt>(L.y);
    if (SOLVER == PRR_EULER) {
        const bool fp64 = !MIXED || r < tr.r_fp64;
        double f[N];
        if (fp64) euler_rhs<Mt, double>(L.y, L.cs, f);
        else euler_rhs<Mt, float>(L.y, L.cs, f);
        bool ok = true;
#pragma unroll
        for (int i = 0; i < N; ++i) ok = ok && __builtin_isfinite(f[i]);
        if (!ok) return PRR_NAN;
        const double h = Mt::mino ? tr.h / lambda_rate<Mt>(L.y, L.cs) : tr.h;
        double yn[N];
#pragma unroll
        for (int i = 0; i < N; ++i) yn[i] = L.y[i] + h * f[i];
        L.steps++;
        L.attempts++;
        L.att64 += fp64 ? 1 : 0;
        L.lam = Mt::mino ? yn[5] : L.lam + h;
        if (!finite_state<Mt>(yn)) return PRR_NAN;
        double yo[N];
#pragma unroll
        for (int i = 0; i < N; ++i) yo[i] = L.y[i];
        return euler_events<Mt>(job, L, yo, yn);
    }
    if (L.attempts >= tr.max_attempts) return PRR_STEP_CAP;
    L.attempts++;
    bool at_min;
    const double h = rkf_step_size<Mt>(tr, L.y, L.cs, L.h, at_min);
    // FP64 stages: the plain form; in the FP32-bulk form inside r_fp64 and,
    // for Kerr, where the step can come within ~13 degrees of the polar axis
    // (1 - mu^2 and L / (1 - mu^2) lose their relative precision in FP32)
    bool fp64 = !MIXED || r < tr.r_fp64;
    if (Mt::mino && !fp64) {
        const double mu_reach = fabs(L.y[2]) + h * fabs(L.y[3]);
        fp64 = 1.0 - mu_reach * mu_reach < kAxisFP64;
    }
    L.att64 += fp64 ? 1 : 0;
    double d4[N], e[N];
    if (fp64) rkf_stages<Mt, double>(L.y, h, L.cs, d4, e);
    else if (Mt::mino) kerr_stages_fp32(L.y, h, L.cs, d4, e);
    else rkf_stages<Mt, float>(L.y, h, L.cs, d4, e);
    double errn = 0.0;
    double yn[N];
    // Kerr: mu = cos th and mu' are measured in units of th (divided by
    // sin th), so that the tolerance keeps its meaning near the poles, where a
    // fixed error in mu is a large relative error in 1 - mu^2
    const double ist = Mt::mino ? 1.0 / fmax(sqrt(fmax(1.0 - L.y[2] * L.y[2], 0.0)), 1e-12) : 1.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        yn[i] = L.y[i] + d4[i];
        if (i < Mt::NERR) {
            const double sc = tr.atol + tr.rtol * fmax(fabs(L.y[i]), fabs(yn[i]));
            const double ei = (Mt::mino && (i == 2 || i == 3)) ? fabs(e[i]) * ist : fabs(e[i]);
            errn = fmax(errn, ei / sc);
        }
    }
    // Kerr: a step may not reach the polar axis (|cos th| >= 1), where the chart is singular
    if (Mt::mino && !(fabs(yn[2]) < 1.0)) errn = fmax(errn, 1024.0);
    // _h = clamp(0.9 h (eps/maxerr)^0.2, hmin, hmax) (kernels/solver.cuh:175-176)
    double hn;
    if (MIXED) hn = 0.9 * h * double(powf(float(errn), -0.2f));
    else hn = 0.9 * h * pow(errn, -0.2);
    if (!Mt::mino) hn = fmin(fmax(hn, tr.hmin), tr.hmax);
    if (errn > 1.0 && !at_min) { // reject (:185-188)
        L.h = hn;
        return -1;
    }
    if (!finite_state<Mt>(yn)) return PRR_NAN;
    if constexpr (Mt::mino) Kerr::project(yn, L.cs);
    L.steps++;
    L.lam = Mt::mino ? yn[5] : L.lam + h;
    L.h = hn;
    double yo[N];
#pragma unroll
    for (int i = 0; i < N; ++i) yo[i] = L.y[i];
    return rkf_events<Mt>(job, L, yo, yn, h);
}

} // namespace prr
