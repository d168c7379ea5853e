// lane.hpp -- the per-ray integrator: ray set-up, one Euler step / RKF45
// attempt / event-location sub-step per loop trip, and shading, for one ray
// held by one lane.
//
// The kernels (kernels.hip) run these as a pipeline over a ray-state array:
// init_ray (k_init) -> lane_iterate in a persistent loop (k_integrate) ->
// shade_ray (k_shade), so that the integrator's loop carries no calls and no
// once-per-ray code.  Everything here is __host__ __device__: the same code
// runs on the host for selected rays (prr_debug_trace_rays).
//
// Replaces initRays, trace_Euler / trace_RKF and rayToRGB
// (kernels/solver.cuh:26-220, metrics/schwarzschild.cuh:55-79,151-183).
// For the adaptive solver the two events whose position depends on the step
// sequence are located on the integrated path instead of at the first step
// beyond them, by FP64 RK45 sub-steps of theta h from the step's start, one
// sub-step per loop trip (the lane's search state lives in a LocateCtx in
// device memory):
//   * escape: the state where r crosses r_escape (Newton on theta);
//   * sphere hit: the first point of the path inside a sphere, for steps
//     whose chord passes within a margin of it: the path's end or the minimum
//     of |P(theta) - c|^2 (successive parabolic interpolation) brackets the
//     entry, which Illinois regula falsi refines.
// The fixed-step Euler solver keeps the reference's chord and first-step
// semantics.
#pragma once

#include "geodesic.hpp"

namespace prr {

// Explicit Runge-Kutta tableaux: S stages, a(s, j), c(s) = sum_j a(s, j), the
// propagated weights b(s) and the error weights e(s) (propagated minus
// embedded; all zero for a fixed-step method).

// Fehlberg 4(5) tableau (kernels/solver.cuh:125-131; Fehlberg 1969, NASA TR
// R-315); the 4th order is propagated, as in the reference.
struct Fehl {
    static constexpr int S = 6;
    static constexpr bool adaptive = true;
    static PRR_HD constexpr double b(int s) { return b4(s); }
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

// Dormand-Prince 5(4) (Dormand and Prince 1980, J. Comput. Appl. Math. 6,
// 19-26), the reference's "dormand" (main/Parameters.h:45): 7 stages, the 5th
// order is propagated and the embedded 4th order gives the error estimate.
struct Dopri {
    static constexpr int S = 7;
    static constexpr bool adaptive = true;
    static PRR_HD constexpr double a(int s, int j) {
        constexpr double A[7][6] = {
            {0, 0, 0, 0, 0, 0},
            {1.0 / 5.0, 0, 0, 0, 0, 0},
            {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0, 0},
            {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 0, 0, 0},
            {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 0, 0},
            {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0, 0},
            {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0}};
        return A[s][j];
    }
    static PRR_HD constexpr double c(int s) {
        constexpr double C[7] = {0.0, 1.0 / 5.0, 3.0 / 10.0, 4.0 / 5.0, 8.0 / 9.0, 1.0, 1.0};
        return C[s];
    }
    static PRR_HD constexpr double b(int s) {
        constexpr double B[7] = {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0.0};
        return B[s];
    }
    // b5 - b4* (the embedded 4th-order weights 5179/57600, 0, 7571/16695,
    // 393/640, -92097/339200, 187/2100, 1/40)
    static PRR_HD constexpr double e(int s) {
        constexpr double E[7] = {35.0 / 384.0 - 5179.0 / 57600.0,     0.0,
                                 500.0 / 1113.0 - 7571.0 / 16695.0,   125.0 / 192.0 - 393.0 / 640.0,
                                 -2187.0 / 6784.0 + 92097.0 / 339200.0, 11.0 / 84.0 - 187.0 / 2100.0,
                                 -1.0 / 40.0};
        return E[s];
    }
};

// Classic 4th-order Runge-Kutta (Kutta 1901), the reference's "rk"
// (main/Parameters.h:41): fixed step, no error estimate.
struct Rk4 {
    static constexpr int S = 4;
    static constexpr bool adaptive = false;
    static PRR_HD constexpr double a(int s, int j) { return (j == s - 1) ? (s == 3 ? 1.0 : 0.5) : 0.0; }
    static PRR_HD constexpr double c(int s) {
        constexpr double C[4] = {0.0, 0.5, 0.5, 1.0};
        return C[s];
    }
    static PRR_HD constexpr double b(int s) {
        constexpr double B[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
        return B[s];
    }
    static PRR_HD constexpr double e(int) { return 0.0; }
};

// The tableau of an RK solver.
template <int SOLVER> struct TableauOf { using type = Fehl; };
template <> struct TableauOf<PRR_DORMAND> { using type = Dopri; };
template <> struct TableauOf<PRR_RK4> { using type = Rk4; };

// 1 - cos^2 th below which Kerr's FP32-bulk attempts run their stages in FP64.
constexpr double kAxisFP64 = 0.05;
// Largest turn of a Kerr step, in units of the ray's angular distance from the polar axis.
constexpr double kAxisStep = 0.25;

// Kerr's right-hand-side coefficients of the lane's ray, computed once when
// the lane takes the ray instead of once per attempt (other metrics: none).
template <bool MINO> struct LaneCoef {
    PRR_HD void set(const Consts<double> &) {}
};
template <> struct LaneCoef<true> {
    KerrCoef<double> k;
    PRR_HD void set(const Consts<double> &c) { k = Kerr::coef<double>(c); }
};

// Lane state of one ray.
// While the lane locates an escape (MODE_LOCATE), three fields that the ray
// no longer needs hold the location's state, so that it stays in registers:
// h is the size of the step being searched, lam the fraction theta of the
// pending sub-step and hit the Newton iteration count.
template <class Mt> struct Lane {
    double y[Mt::N];
    Consts<double> cs;
    LaneCoef<Mt::mino> kc;
    double h;   // next step size in the integration variable (Mino time for Kerr)
    double lam; // affine parameter
    int32_t steps, attempts, att64;
    int32_t frame;
    int32_t hit; // sphere index of a SPHERE termination
    int64_t rid;
    // Sphere screen at the start of the step: the FP32 Cartesian position; for
    // Kerr its (rho, z) in P0[0..1], and in gap a path length the ray can go
    // before any sphere's FP32 screen can pass (not above 0: screen the step)
    float P0[3];
    float gap;
    int32_t seg;   // adjoint replay: the segment being replayed
    int64_t rec_i; // adjoint replay: the task's record column
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

// Frame of a queue position (frames' queue ranges start at multiples of 64).
PRR_HD int find_frame_q(const JobDev &job, int64_t q) {
    int lo = 0, hi = job.nframes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (job.frames[mid].queue_begin <= q) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// k-th index in [0, n) counted from the middle outwards, alternating sides:
// n = 4 gives 2, 1, 3, 0 and n = 5 gives 2, 3, 1, 4, 0.
PRR_HD int64_t centre_out(int64_t k, int64_t n) {
    const int64_t h = n >> 1;
    if (n & 1) return (k & 1) ? h + ((k + 1) >> 1) : h - (k >> 1);
    return (k & 1) ? h - 1 - (k >> 1) : h + (k >> 1);
}

// Queue position -> ray id, or -1 for the padding after a frame whose ray
// count is not a multiple of 64.  The queue hands a wave 64 consecutive positions;
// within each frame they map to a tile of pixels (all samples of each pixel)
// rather than to a run of one image row, so that a wave's rays are close
// together on the sky and follow similar paths; tiles are ordered from the
// middle of the image outwards (below).  Tiles hold 64 / s pixels
// (s samples per pixel, a divisor of 64), Th x Tw with Th the largest power
// of two whose square fits; the part of the image that whole tiles do not
// cover (a right strip, then a bottom strip) keeps row order.  A bijection on
// each frame's ids (j W + i) s + k; frames whose s does not divide 64 keep
// the identity.
PRR_HD int64_t queue_ray(const JobDev &job, int64_t q) {
    if (job.explicit_x0u0) return q;
    const FrameDev &fr = job.frames[find_frame_q(job, q)];
    const int64_t s = fr.s;
    const int64_t local = q - fr.queue_begin;
    if (local >= int64_t(fr.W) * fr.H * s) return -1;
    if (s <= 0 || 64 % s != 0) return fr.ray_begin + local;
    const int64_t P = 64 / s;
    int64_t th = 1;
    while ((th * 2) * (th * 2) <= P) th *= 2;
    const int64_t tw = P / th;
    const int64_t W = fr.W, H = fr.H;
    const int64_t wc = (W / tw) * tw, hc = (H / th) * th;
    const int64_t ncore = wc * hc * s;
    int64_t i, j, k;
    if (local < ncore) {
        const int64_t per = P * s, q_t = local / per, w = local - q_t * per;
        // tiles are handed out a column at a time, columns from the middle
        // of the image outwards and, within a column, rows from the middle
        // outwards: the rays that take the most steps (those that pass close
        // to the polar axis, whose image is the middle column when the spin
        // axis projects upright, and those near the photon ring around the
        // shadow) start first, and the cheap sky at the left and right edges
        // fills the end of the launch
        const int64_t tiles_x = wc / tw, tiles_y = hc / th;
        const int64_t kc = q_t / tiles_y, kr = q_t - kc * tiles_y;
        const int64_t tx = centre_out(kc, tiles_x), ty = centre_out(kr, tiles_y);
        const int64_t p = w / s;
        k = w - p * s;
        i = tx * tw + p % tw;
        j = ty * th + p / tw;
    } else {
        int64_t r = local - ncore;
        const int64_t wr = W - wc, nright = wr * hc * s;
        if (r < nright) { // right strip: x >= wc, y < hc
            const int64_t p = r / s;
            k = r - p * s;
            i = wc + p % wr;
            j = p / wr;
        } else { // bottom strip: y >= hc
            r -= nright;
            const int64_t p = r / s;
            k = r - p * s;
            i = p % W;
            j = hc + p / W;
        }
    }
    return fr.ray_begin + (j * W + i) * s + k;
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
    return rcp_t(y[0] * y[0]) + cs.a * cs.a * y[2] * y[2];
}
// 1 / Sigma = u^2 / (1 + a^2 mu^2 u^2), with one reciprocal.
template <class Mt> PRR_HD double inv_lambda_rate(const double *y, const Consts<double> &cs) {
#pragma clang fp contract(off)
    if (!Mt::mino) return 1.0;
    const double u2 = y[0] * y[0];
    return u2 * rcp_t(1.0 + cs.a * cs.a * y[2] * y[2] * u2);
}

// Near the polar axis a Kerr step may turn the ray by at most kAxisStep
// times its angular distance from the axis (sin th): the L / sin^2 th spike
// of phi' is then resolved by several steps, which the embedded error
// estimate alone does not guarantee (angular speed sqrt(Q + L^2 + a^2 E^2
// mu^2)).  A bound, so FP32 is enough.
PRR_HD double axis_step_limit(const double *y, const Consts<double> &cs) {
#pragma clang fp contract(off)
    const double s2 = 1.0 - y[2] * y[2];
    const double K = cs.Q + cs.L * cs.L + cs.a * cs.a * cs.E * cs.E * y[2] * y[2];
    return double(float(kAxisStep) * sqrt_f32(float(fmax(s2, 0.0)) * rcp_t(float(K))));
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
    const double iS = inv_lambda_rate<Mt>(y, cs);
    const double lo = tr.hmin * iS;
    h = fmin(fmax(h, lo), fmin(tr.hmax * iS, fmax(axis_step_limit(y, cs), lo)));
    at_min = h <= lo;
    return h;
}

// Step of the fixed-step RK4 solver: h, in Mino time h / Sigma, within the
// polar-axis limit.
template <class Mt> PRR_HD double fixed_step_size(const TracerDev &tr, const double *y, const Consts<double> &cs) {
#pragma clang fp contract(off)
    if (!Mt::mino) return tr.h;
    return fmin(tr.h * inv_lambda_rate<Mt>(y, cs), axis_step_limit(y, cs));
}

// The RKF45 step factor 0.9 errn^(-1/5) (kernels/solver.cuh:175): an FP32
// estimate from the hardware log2 / exp2 on the device, refined by one Newton
// step in FP64 for y^-5 = x, y <- y (6 - x y^5) / 5, which squares the
// estimate's relative error (~1e-7 -> ~3e-14) at a tenth of the cost of pow.
// errn is clamped to [1e-30, 1e30] (errn = 0 gives the factor 0.9e6; the
// step bounds cap it).
PRR_HD double step_factor(double errn) {
#pragma clang fp contract(off)
    const double x = fmin(fmax(errn, 1e-30), 1e30);
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = double(__builtin_amdgcn_exp2f(-0.2f * __builtin_amdgcn_logf(float(x))));
#else
    const double y = double(exp2f(-0.2f * log2f(float(x))));
#endif
    const double y2 = y * y;
    const double y5 = y2 * y2 * y;
    return 0.9 * (y * (1.2 - 0.2 * (x * y5)));
}

// True when c holds for any lane of the wave that executes this call (on the
// host, which runs one lane at a time: c itself).
PRR_HD bool wave_any(bool c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __ballot(c) != 0;
#else
    return c;
#endif
}

// Six RK stages in precision T from the FP64 state y (plain or FP32 form of
// the 8-component metrics, and Kerr's FP64 form); the 4th-order increment d4
// and the error vector e in FP64.
template <class Mt, class T, class Tab = Fehl, class C>
PRR_HD void rkf_stages_c(const double *y, double h, const C &c, double *d4, double *e) {
    constexpr int N = Mt::N, S = Tab::S;
    // The slopes f_s stay unscaled: h enters through the attempt's coefficients
    // h a_sj, h b_s and h e_s, one multiply per coefficient instead of one per
    // slope component.
    T yT[N];
#pragma unroll
    for (int i = 0; i < N; ++i) yT[i] = T(y[i]);
    T f[S][N];
    Mt::template rhs<T>(yT, c, f[0]);
#pragma unroll
    for (int s = 1; s < S; ++s) {
        T ys[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            T acc = yT[i];
#pragma unroll
            for (int j = 0; j < s; ++j)
                if (Tab::a(s, j) != 0.0) acc += T(h * Tab::a(s, j)) * f[j][i];
            ys[i] = acc;
        }
        Mt::template rhs<T>(ys, c, f[s]);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double a4 = 0.0, ae = 0.0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (Tab::b(s) != 0.0) a4 += (h * Tab::b(s)) * double(f[s][i]);
            if (Tab::e(s) != 0.0) ae += (h * Tab::e(s)) * double(f[s][i]);
        }
        d4[i] = a4;
        e[i] = ae;
    }
}
template <class Mt, class T, class Tab = Fehl>
PRR_HD void rkf_stages(const double *y, double h, const Consts<double> &cd, double *d4, double *e) {
    rkf_stages_c<Mt, T, Tab>(y, h, Mt::template coef<T>(cd), d4, e);
}

// Kerr's FP32-bulk attempt: stage 1 and the difference coefficients in FP64,
// stages 2..S evaluate g_s = f(y + d_s) - f(y) in FP32 with
// d_s = h (c_s F + sum_j a_sj g_j).  Since sum b = 1 and sum e = 0 (for
// RKF45: b = b4, e = b5 - b4),
//   d4 = h F + h sum b_s g_s,   e = h sum e_s g_s,
// so FP32 rounding enters only through the small differences g_s and the
// error estimate never sees the rounding of f itself.
template <class Tab = Fehl>
PRR_HD void kerr_stages_fp32_c(const double *y, double h, const KerrCoef<double> &kc, double *d4, double *e) {
    constexpr int N = Kerr::N, S = Tab::S;
    double F[N];
    KerrDelta P;
    Kerr::delta_prep(y, kc, F, P);
    const float hf = float(h);
    float Ff[N];
#pragma unroll
    for (int i = 0; i < N; ++i) Ff[i] = float(F[i]);
    float g[S][N];
#pragma unroll
    for (int i = 0; i < N; ++i) g[0][i] = 0.0f;
#pragma unroll
    for (int s = 1; s < S; ++s) {
        float d[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            // h enters through the coefficients h c_s, h a_sj (one multiply each)
            float acc = (hf * float(Tab::c(s))) * Ff[i];
#pragma unroll
            for (int j = 1; j < s; ++j)
                if (Tab::a(s, j) != 0.0) acc += (hf * float(Tab::a(s, j))) * g[j][i];
            d[i] = acc;
        }
        Kerr::delta_rhs(d, P, g[s]);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        float a4 = 0.0f, ae = 0.0f;
#pragma unroll
        for (int s = 1; s < S; ++s) {
            if (Tab::b(s) != 0.0) a4 += (hf * float(Tab::b(s))) * g[s][i];
            if (Tab::e(s) != 0.0) ae += (hf * float(Tab::e(s))) * g[s][i];
        }
        d4[i] = h * F[i] + double(a4);
        e[i] = double(ae);
    }
}
template <class Tab = Fehl>
PRR_HD void kerr_stages_fp32(const double *y, double h, const Consts<double> &cd, double *d4, double *e) {
    kerr_stages_fp32_c<Tab>(y, h, Kerr::coef<double>(cd), d4, e);
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

// One FP64 RK45 step (4th-order solution) of size h from y.
template <class Mt> PRR_HD void rk_substep(const double *y, double h, const Consts<double> &cs, double *out) {
    double d4[Mt::N], e[Mt::N];
    rkf_stages<Mt, double>(y, h, cs, d4, e);
    for (int i = 0; i < Mt::N; ++i) out[i] = y[i] + d4[i];
}

// Escape radius on the state's side: beyond a wormhole's throat (l < 0)
// rays escape at the second sky's radius (rout2).
template <class Mt> PRR_HD double escape_radius(const JobDev &job, const FrameDev &fr, const double *y) {
    return (Mt::kind == PRR_ELLIS && y[0] < 0.0) ? fr.r_esc2 : job.tr.r_escape;
}

// The radius tests of state y.  Kerr's state holds u = 1 / r: it compares u
// with the reciprocals of the radii, computed once on the host, instead of
// taking a reciprocal of u per step.
template <class Mt> PRR_HD bool beyond_escape(const JobDev &job, const FrameDev &fr, const double *y) {
    if constexpr (Mt::mino) return y[0] <= job.tr.inv_r_escape;
    else return radius_of<Mt>(y) >= escape_radius<Mt>(job, fr, y);
}
template <class Mt> PRR_HD bool inside_horizon(const FrameDev &fr, const double *y) {
    if constexpr (Mt::mino) return y[0] >= fr.inv_r_hor;
    else return radius_of<Mt>(y) <= fr.r_hor;
}
template <class Mt> PRR_HD bool inside_r_fp64(const TracerDev &tr, const double *y) {
    if constexpr (Mt::mino) return y[0] > tr.inv_r_fp64;
    else return radius_of<Mt>(y) < tr.r_fp64;
}

// Escape crossing function g (g < 0 inside r_escape) and dg/dt along the
// integration variable.
template <class Mt> PRR_HD void escape_g(const double *y, double r_esc, double &g, double &dg) {
    if (Mt::mino) {
        g = 1.0 / r_esc - y[0];
        dg = -y[1];
    } else if (Mt::kind == PRR_ELLIS) { // |l| = r_escape on either side of the throat
        g = fabs(y[0]) - r_esc;
        dg = y[0] < 0.0 ? -y[4] : y[4];
    } else if (Mt::spherical) {
        g = y[0] - r_esc;
        dg = y[4];
    } else {
        const double r = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
        g = r - r_esc;
        dg = (y[0] * y[4] + y[1] * y[5] + y[2] * y[6]) / r;
    }
}

// Sphere k: G = |P - c|^2 - R^2 at a state.
template <class Mt> PRR_HD double sphere_G(const FrameDev &fr, int k, const double *y) {
    const D3 m = cart_of<Mt>(y) - d3(fr.sph[k][0], fr.sph[k][1], fr.sph[k][2]);
    return dot3(m, m) - fr.sph[k][3] * fr.sph[k][3];
}

// Path length from Q, the end of a screened chord, before sphere k's FP32
// screen can pass: |Q - c| - R less a margin for the screen's relative margin
// and the FP32 rounding of the positions over the steps it then covers (the
// screen's scale grows with the distance travelled, at most |Q - c|).
PRR_HD float sphere_gap(const FrameDev &fr, int k, const float Q[3]) {
    const float mx = Q[0] - fr.sphf[k][0], my = Q[1] - fr.sphf[k][1], mz = Q[2] - fr.sphf[k][2];
    return sqrt_f32(mx * mx + my * my + mz * mz) - fr.sphf[k][3] -
           5e-4f * (fr.sphf[k][4] + fabsf(Q[0]) + fabsf(Q[1]) + fabsf(Q[2]));
}

// FP32 screen: can the path of this step come near sphere k?  The chord's
// closest distance to the centre against R plus half the chord length (a
// bound on the path's bulge for any step that turns by less than ~1 rad)
// plus a relative margin for FP32 rounding.
PRR_HD bool sphere_screen(const FrameDev &fr, int k, const float P0[3], const float P1[3]) {
    const float cx = fr.sphf[k][0], cy = fr.sphf[k][1], cz = fr.sphf[k][2], R = fr.sphf[k][3];
    const float mx = P0[0] - cx, my = P0[1] - cy, mz = P0[2] - cz;
    const float dx = P1[0] - P0[0], dy = P1[1] - P0[1], dz = P1[2] - P0[2];
    const float A = dx * dx + dy * dy + dz * dz;
    // |cx| + |cy| + |cz| + R + |P0|_1, the sphere's part summed on the host
    const float scale = fr.sphf[k][4] + fabsf(P0[0]) + fabsf(P0[1]) + fabsf(P0[2]);
    // every point of the chord is at least |m| - |d| from the centre: steps that
    // start farther than R + 1.5 |d| + the margin pass nowhere near (most steps)
    const float far = R + 1.5f * sqrt_f32(A) + 2e-4f * scale;
    if (mx * mx + my * my + mz * mz > far * far) return false;
    float t = A > 0.0f ? -(mx * dx + my * dy + mz * dz) * rcp_t(A) : 0.0f;
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    const float qx = mx + t * dx, qy = my + t * dy, qz = mz + t * dz;
    const float dist = sqrt_f32(qx * qx + qy * qy + qz * qz);
    return dist <= R + 0.5f * sqrt_f32(A) + 1e-4f * scale;
}

// The screen's state where no budget carries over (a ray loaded, or resumed
// after a location): the position of L.y, and no budget.
template <class Mt> PRR_HD void screen_reset(const FrameDev &fr, Lane<Mt> &L) {
    L.P0[0] = L.P0[1] = L.P0[2] = 0.0f;
    L.gap = 0.0f;
    if (fr.n_spheres > 0) {
        if (Mt::mino) kerr_rz_f32(L.y, L.P0);
        else cart_of_f32<Mt>(L.y, L.P0);
    }
}

// ------------------------------------------------------------ ray state I/O
// Doubles per ray in the ray-state arrays: the state, then E, L, Q for Kerr.
template <class Mt> constexpr int rs_width() { return Mt::mino ? Mt::N + 3 : Mt::N; }

PRR_HD int64_t rs_index(const JobDev &job, int64_t id) { return id - job.rs_base; }

// k_init: the initial state of ray id (initRays, kernels/solver.cuh:26-56).
template <class Mt> PRR_HD void init_ray(const JobDev &job, int64_t id) {
    const int f = job.explicit_x0u0 ? 0 : find_frame(job, id);
    const FrameDev &fr = job.frames[f];
    double y[Mt::N];
    Consts<double> cs;
    if (job.explicit_x0u0) {
        const double *s = job.explicit_x0u0 + 8 * id;
        state_from_xu<Mt>(s, s + 4, fr.M, fr.a, y, cs);
    } else {
        init_state_camera<Mt>(fr, camera_ray(fr, id - fr.ray_begin), y, cs);
    }
    const int64_t T = job.rs_total, i = rs_index(job, id);
    for (int k = 0; k < Mt::N; ++k) job.rs.y[k * T + i] = y[k];
    if (Mt::mino) {
        job.rs.y[(Mt::N + 0) * T + i] = cs.E;
        job.rs.y[(Mt::N + 1) * T + i] = cs.L;
        job.rs.y[(Mt::N + 2) * T + i] = cs.Q;
    }
}

template <class Mt> PRR_HD void load_consts(const JobDev &job, int64_t id, const FrameDev &fr, Consts<double> &cs) {
    const int64_t T = job.rs_total, i = rs_index(job, id);
    cs.M = fr.M;
    cs.a = fr.a;
    if (Mt::mino) {
        cs.E = job.rs.y[(Mt::N + 0) * T + i];
        cs.L = job.rs.y[(Mt::N + 1) * T + i];
        cs.Q = job.rs.y[(Mt::N + 2) * T + i];
    } else {
        cs.E = cs.L = cs.Q = 0.0;
    }
}

// The lane takes ray id from the ray-state arrays.
template <class Mt> PRR_HD void lane_load(const JobDev &job, Lane<Mt> &L, int64_t id) {
    L.rid = id;
    L.frame = job.explicit_x0u0 ? 0 : find_frame(job, id);
    const FrameDev &fr = job.frames[L.frame];
    const int64_t T = job.rs_total, i = rs_index(job, id);
    for (int k = 0; k < Mt::N; ++k) L.y[k] = job.rs.y[k * T + i];
    load_consts<Mt>(job, id, fr, L.cs);
    L.kc.set(L.cs);
    L.h = job.tr.h / lambda_rate<Mt>(L.y, L.cs);
    L.lam = 0.0;
    L.steps = L.attempts = L.att64 = 0;
    L.hit = -1;
    screen_reset<Mt>(fr, L);
    L.seg = 0;
    L.rec_i = 0;
}

// Checkpoint j of ray id (8 doubles): slot id * cap + j in the first pass,
// which has no offset table; afterwards ckpt_off[id] + j in ckpt, or
// (-1 - ckpt_off[id]) + j in ckpt2 for the rays with more than cap segments.
PRR_HD double *ckpt_at(const JobDev &job, int64_t id, int j) {
    const AdjRunDev &a = job.adj;
    if (!a.ckpt_off) return a.ckpt + 8 * (id * a.cap + j);
    const int64_t o = a.ckpt_off[id];
    return o >= 0 ? a.ckpt + 8 * (o + j) : a.ckpt2 + 8 * (-1 - o + j);
}

// The lane takes ray id from its checkpoint j (adjoint replay).
template <class Mt> PRR_HD void lane_load_checkpoint(const JobDev &job, Lane<Mt> &L, int64_t id, int j) {
    L.rid = id;
    L.frame = find_frame(job, id);
    const FrameDev &fr = job.frames[L.frame];
    const double *c = ckpt_at(job, id, j);
    for (int k = 0; k < Mt::N; ++k) L.y[k] = c[k];
    load_consts<Mt>(job, id, fr, L.cs);
    L.kc.set(L.cs);
    L.h = c[6];
    L.attempts = int32_t(c[7]);
    L.lam = Mt::mino ? L.y[5] : 0.0;
    L.steps = j * job.adj.K;
    L.seg = j;
    L.att64 = 0;
    L.hit = -1;
    screen_reset<Mt>(fr, L);
}

PRR_HD void store_checkpoint(const JobDev &job, const double *y, int n, double h, int32_t attempts, int64_t id,
                             int j) {
    double *c = ckpt_at(job, id, j);
    for (int k = 0; k < 6; ++k) c[k] = k < n ? y[k] : 0.0;
    c[6] = h;
    c[7] = double(attempts);
}

// The lane's final state back into the ray-state arrays.
template <class Mt> PRR_HD void lane_store(const JobDev &job, const Lane<Mt> &L, int cls) {
    const int64_t T = job.rs_total, i = rs_index(job, L.rid);
    for (int k = 0; k < Mt::N; ++k) job.rs.y[k * T + i] = L.y[k];
    job.rs.cls[i] = int8_t(cls);
    job.rs.hit[i] = int8_t(L.hit);
    job.rs.steps[i] = L.steps;
    job.rs.attempts[i] = L.attempts;
}

// k_shade: outputs of a finished ray (rayToRGB, metrics/schwarzschild.cuh:151-183).
template <class Mt> PRR_HD void shade_ray(const JobDev &job, int64_t id) {
    const int f = job.explicit_x0u0 ? 0 : find_frame(job, id);
    const FrameDev &fr = job.frames[f];
    const int64_t T = job.rs_total, i = rs_index(job, id);
    double y[Mt::N];
    for (int k = 0; k < Mt::N; ++k) y[k] = job.rs.y[k * T + i];
    Consts<double> cs;
    load_consts<Mt>(job, id, fr, cs);
    const int cls = job.rs.cls[i], hit = job.rs.hit[i];
    D3 endpoint;
    float3 color;
    if (cls == PRR_SPHERE) {
        endpoint = normalize3(cart_of<Mt>(y) - d3(fr.sph[hit][0], fr.sph[hit][1], fr.sph[hit][2]));
        const int slot = fr.sph_tex[hit];
        color = sphere_color((slot > 0 && slot < PRR_MAX_TEXTURES) ? &job.tex[slot] : nullptr, endpoint);
    } else if (cls == PRR_ESCAPE) {
        const D3 n = escape_dir<Mt>(y, cs);
        endpoint = n;
        double th, ph;
        if (Mt::kind == PRR_SCHW_SPH || Mt::kind == PRR_ELLIS) {
            const double two_pi = 6.28318530717958647692, pi = 3.14159265358979323846;
            th = y[1];
            ph = y[2];
            if (ph == two_pi) ph = 0.0; // main/base.h:86-89
            if (th == pi) th = 0.0;
        } else {
            th = acos(fmin(fmax(n.z, -1.0), 1.0));
            ph = atan2(n.y, n.x);
            if (ph < 0.0) ph += 6.28318530717958647692;
        }
        // through a wormhole's throat (l < 0): the second sky (bimage2)
        const bool far_side = Mt::kind == PRR_ELLIS && y[0] < 0.0;
        color = sky_lookup(job.tex[far_side ? PRR_SKY2_SLOT : 0], th, ph);
    } else {
        endpoint = cart_of<Mt>(y);
        color = make_float3(0.f, 0.f, 0.f);
    }
    const OutDev &o = job.out; // per-ray arrays like the ray state (index i)
    if (o.color) {
        o.color[3 * i] = color.x;
        o.color[3 * i + 1] = color.y;
        o.color[3 * i + 2] = color.z;
    }
    if (o.cls) o.cls[i] = int8_t(cls);
    if (o.endpoint) {
        o.endpoint[3 * i] = endpoint.x;
        o.endpoint[3 * i + 1] = endpoint.y;
        o.endpoint[3 * i + 2] = endpoint.z;
    }
    if (o.steps) o.steps[i] = job.rs.steps[i];
    if (o.attempts) o.attempts[i] = job.rs.attempts[i];
    if (o.state8) {
        double x[4], u[4];
        xu_from_state<Mt>(y, cs, x, u);
        for (int c = 0; c < 4; ++c) {
            o.state8[8 * i + c] = x[c];
            o.state8[8 * i + 4 + c] = u[c];
        }
    }
}

// ------------------------------------------------------------ Euler
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
    chart_post_step<Mt>(yn);
    for (int i = 0; i < Mt::N; ++i) L.y[i] = yn[i];
    if (inside_horizon<Mt>(fr, yn)) return PRR_HORIZON;
    if (beyond_escape<Mt>(job, fr, yn)) return PRR_ESCAPE;
    return -1;
}

template <class Mt, bool MIXED> PRR_HD int euler_step(const JobDev &job, Lane<Mt> &L) {
    constexpr int N = Mt::N;
    const TracerDev &tr = job.tr;
    if (L.steps >= tr.max_steps || L.lam >= tr.lambda_max) return PRR_STEP_CAP;
    bool fp64 = !MIXED || inside_r_fp64<Mt>(tr, L.y);
    // FP32-bulk render with whole-wave refills: one precision per wave and
    // step, as for the Runge-Kutta attempts (lane_iterate)
    if constexpr (MIXED && !Mt::ckpt && !Mt::replay)
        if (job.refill >= 64) fp64 = wave_any(fp64);
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

// ------------------------------------------------------------ RKF45 events
// MODE_LOCATE: the escape crossing inside the last accepted step, located by
// Newton sub-steps with its state in the lane (Lane); MODE_SPHERE: a sphere
// hit, located through the lane's LocateCtx in device memory by out-of-line
// functions (sphere_prepare / sphere_update) that take the lane's state by
// value.  Sphere locations are rare (only steps whose chord passes near a
// sphere), and keeping their code out of the loop leaves the loop the
// registers of three waves per SIMD.
enum { MODE_IDLE = 0, MODE_ACTIVE = 1, MODE_LOCATE = 2, MODE_SPHERE = 3 };
enum { SPH_NEXT = 0, SPH_PROBE = 1, SPH_MIN = 2, SPH_BRACKET = 3 };
constexpr int kNeedSubstep = -2; // a sub-step of the returned size is due
constexpr int kResume = -3;      // the step had no event after all: the lane goes on
constexpr int kSphereHit = -4;   // sphere_prepare: the best hit (C.ybest, C.best_th, C.best_k) ends the ray
constexpr int kSphereNone = -5;  // sphere_prepare: no sphere was hit in the step (its end state is C.yn)

template <int N> PRR_HD void copy_n(double *d, const double *s) {
    for (int i = 0; i < N; ++i) d[i] = s[i];
}

// A state passed by value to the out-of-line sphere functions (the lane's
// own array stays in registers).
template <int N> struct StateN {
    double v[N];
};

// Adjoint checkpoint passes: a checkpoint every K accepted steps while the lane integrates.
template <class Mt> PRR_HD void maybe_checkpoint(const JobDev &job, const Lane<Mt> &L) {
    if (!Mt::ckpt || job.adj.mode != ADJ_CHECKPOINT) return;
    const int K = job.adj.K;
    if (L.steps % K != 0) return;
    // with known step counts, checkpoints at 0, K, 2K, ... < N_i; in the
    // first pass, the first cap of them
    if (job.adj.nsteps ? L.steps >= job.adj.nsteps[L.rid] : L.steps / K >= job.adj.cap) return;
    store_checkpoint(job, L.y, Mt::N, L.h, L.attempts, L.rid, L.steps / K);
}

// Adjoint replay: the step's start and size of every accepted step of the
// segment.  A first pass with a record area records every step of the ray
// instead (the reverse sweep then needs no replay).
template <class Mt> PRR_HD void maybe_record(const JobDev &job, const Lane<Mt> &L, double h) {
    if constexpr (Mt::ckpt) {
        if (job.adj.mode == ADJ_CHECKPOINT && job.adj.rec_all && L.steps <= job.adj.rec_cap) {
            double *r = job.adj.rec_all + (L.rid * job.adj.rec_cap + (L.steps - 1)) * 6;
            for (int c = 0; c < 5; ++c) r[c] = L.y[c];
            r[5] = h;
        }
        return;
    }
    if (!Mt::replay || job.adj.mode != ADJ_REPLAY) return;
    const int64_t slot = (L.steps - 1) - int64_t(L.seg) * job.adj.K; // L.steps already counts this step
    const int64_t T = job.adj.rec_stride;
    double *r = job.adj.rec + (slot * 6) * T + L.rec_i;
    for (int c = 0; c < 5; ++c) r[c * T] = L.y[c];
    r[5 * T] = h;
}

// A located event of the step of size h from L.y ends the ray: the lane
// takes the event's state ys (adjoint replay: the terminal step is recorded
// for the seed).
template <class Mt>
PRR_HD int finish_located(const JobDev &job, Lane<Mt> &L, double h, const double *ys, double theta, int cls,
                          int hit) {
    if ((Mt::replay && job.adj.mode == ADJ_REPLAY && job.adj.event) ||
        (Mt::ckpt && job.adj.mode == ADJ_CHECKPOINT && job.adj.rec_all)) {
        EventDev &ev = job.adj.event[L.rid];
        for (int i = 0; i < 8; ++i) ev.y_prev[i] = i < Mt::N ? L.y[i] : 0.0;
        ev.h = h;
        ev.theta = theta;
        ev.cls = cls;
        ev.hit = hit;
    }
    copy_n<Mt::N>(L.y, ys);
    L.hit = hit;
    if (cls == PRR_ESCAPE) chart_post_step<Mt>(L.y);
    return cls;
}

// Escape test of the accepted step L.y -> yn (size h); starts the location
// (MODE_LOCATE) when r crosses r_escape.  Then spherical wraps and the
// horizon.  Returns a class or -1 (go on: L.y = yn, or the location started).
template <class Mt>
PRR_HD int escape_wrap_horizon(const JobDev &job, Lane<Mt> &L, LocateCtx &C, const double *yn, double h,
                               int &mode, const float *P1 = nullptr, float gap = 0.0f) {
    const FrameDev &fr = job.frames[L.frame];
    if (beyond_escape<Mt>(job, fr, yn)) {
        const double r_esc = escape_radius<Mt>(job, fr, yn);
        double go, gn, dg;
        escape_g<Mt>(L.y, r_esc, go, dg);
        if (!(go < 0.0)) return finish_located(job, L, h, L.y, 0.0, PRR_ESCAPE, -1); // started outside
        escape_g<Mt>(yn, r_esc, gn, dg);
#if !defined(__HIP_DEVICE_COMPILE__)
        copy_n<Mt::N>(C.yn, yn); // the host-run loop logs the accepted state
#endif
        L.h = h;
        L.lam = fmin(fmax(go / (go - gn), 0.0), 1.0);
        L.hit = 0;
        mode = MODE_LOCATE;
        return -1;
    }
    double y[Mt::N];
    copy_n<Mt::N>(y, yn);
    chart_post_step<Mt>(y);
    copy_n<Mt::N>(L.y, y);
    if (fr.n_spheres > 0) {
        // P1 (and gap): the screen's state at yn, which is the new state's where the chart has no wraps
        constexpr bool wraps = Mt::kind == PRR_SCHW_SPH || Mt::kind == PRR_ELLIS;
        if (P1 && !wraps) {
            L.P0[0] = P1[0];
            L.P0[1] = P1[1];
            L.P0[2] = P1[2];
            L.gap = gap;
        } else {
            screen_reset<Mt>(fr, L);
        }
    }
    if (inside_horizon<Mt>(fr, yn)) return PRR_HORIZON;
    maybe_checkpoint<Mt>(job, L);
    return -1;
}

// Events of an accepted RKF45 step L.y -> yn of size h: a sphere screen that
// starts the sphere location, or the escape / wrap / horizon tests.
template <class Mt>
PRR_HD int rkf_events(const JobDev &job, Lane<Mt> &L, LocateCtx &C, const double *yn, double h, int &mode) {
    const FrameDev &fr = job.frames[L.frame];
    float P1[3];
    float gap = 0.0f;
    if (fr.n_spheres > 0) {
        unsigned mask = 0u;
        if constexpr (Mt::mino) {
            // Kerr: the screen runs only when the path since the last screen
            // may have used up the budget.  With (rho, z) of both ends,
            // |P1 - P0|^2 = dz^2 + drho^2 + 2 rho0 rho1 (1 - cos dphi) <= D^2 =
            // dz^2 + drho^2 + rho0 rho1 dphi^2, D inflated to cover the FP32
            // rounding of the screen's positions (relative 1e-5).  The step is
            // skipped only when gap - D > D / 2: the FP32 screen cannot pass
            // (any point of this chord is at least gap + R - D from every
            // centre, and the screen passes within R + |P1 - P0| / 2 plus its
            // margin), so the steps taken are those of screening every step.
            kerr_rz_f32(yn, P1);
            P1[2] = 0.0f;
            const float dphi = float(yn[4] - L.y[4]);
            const float dr = P1[0] - L.P0[0], dz = P1[1] - L.P0[1];
            const float D = sqrt_f32(dr * dr + dz * dz + (L.P0[0] * P1[0]) * (dphi * dphi)) * 1.00001f +
                            1e-5f * (P1[0] + fabsf(P1[1]) + L.P0[0] + fabsf(L.P0[1]));
            gap = L.gap - D;
            if (!(gap > 0.5f * D)) { // (also for a NaN D)
                float Q0[3], Q1[3];
                cart_of_f32<Mt>(L.y, Q0);
                cart_of_f32<Mt>(yn, Q1);
                gap = 3.0e38f;
                for (int k = 0; k < fr.n_spheres; ++k) {
                    if (sphere_screen(fr, k, Q0, Q1)) mask |= 1u << k;
                    gap = fminf(gap, sphere_gap(fr, k, Q1));
                }
            }
        } else {
            cart_of_f32<Mt>(yn, P1);
            for (int k = 0; k < fr.n_spheres; ++k)
                if (sphere_screen(fr, k, L.P0, P1)) mask |= 1u << k;
        }
        if (mask) {
            copy_n<Mt::N>(C.yn, yn);
            C.h = h;
            C.phase = SPH_NEXT;
            C.mask = int32_t(mask);
            C.k = -1;
            C.best_k = -1;
            C.best_th = 2.0;
            mode = MODE_SPHERE;
            return -1;
        }
    }
    return escape_wrap_horizon<Mt>(job, L, C, yn, h, mode, fr.n_spheres > 0 ? P1 : nullptr, gap);
}

// Result of sphere_prepare: kNeedSubstep (a sub-step of fraction C.th, size
// hs), kSphereHit or kSphereNone.
struct SphStep {
    int32_t code;
    int32_t pad;
    double hs;
};

// Sphere location before the sub-step of this trip; y is the step's start
// state.  Bookkeeping that needs no sub-step is done here.
template <class Mt>
PRR_HD PRR_NOINLINE SphStep sphere_prepare(const FrameDev *frp, StateN<Mt::N> y, LocateCtx *Cp) {
    constexpr int N = Mt::N;
    const FrameDev &fr = *frp;
    LocateCtx &C = *Cp;
    for (;;) {
        if (C.phase == SPH_BRACKET) {
            const double R2 = fr.sph[C.k][3] * fr.sph[C.k][3];
            const D3 c = d3(fr.sph[C.k][0], fr.sph[C.k][1], fr.sph[C.k][2]);
            const double gtol = 1e-14 * (R2 + dot3(c, c));
            if (C.it < 60 && fabs(C.Ghi) > gtol && C.hi - C.lo > 1e-15) {
                double t = (C.lo * C.Ghi - C.hi * C.Glo) / (C.Ghi - C.Glo);
                if (!(t > C.lo && t < C.hi)) t = 0.5 * (C.lo + C.hi);
                C.th = t;
                return SphStep{kNeedSubstep, 0, t * C.h};
            }
            if (C.hi < C.best_th) {
                C.best_th = C.hi;
                C.best_k = C.k;
                copy_n<N>(C.ybest, C.yb);
            }
            C.phase = SPH_NEXT;
        }
        if (C.phase == SPH_PROBE || C.phase == SPH_MIN) return SphStep{kNeedSubstep, 0, C.th * C.h};
        // SPH_NEXT: the next screened sphere
        int k = C.k + 1;
        while (k < fr.n_spheres && !((unsigned(C.mask) >> k) & 1u)) ++k;
        if (k >= fr.n_spheres) return SphStep{C.best_k >= 0 ? kSphereHit : kSphereNone, 0, 0.0};
        C.k = k;
        const D3 c = d3(fr.sph[k][0], fr.sph[k][1], fr.sph[k][2]);
        const double R2 = fr.sph[k][3] * fr.sph[k][3];
        const D3 P0 = cart_of<Mt>(y.v), P1 = cart_of<Mt>(C.yn);
        const D3 m = P0 - c;
        C.G0 = dot3(m, m) - R2;
        if (C.G0 <= 0.0) { // the step starts inside (reference: chord t = 0)
            if (0.0 < C.best_th) {
                C.best_th = 0.0;
                C.best_k = k;
                copy_n<N>(C.ybest, y.v);
            }
            continue;
        }
        const double G1 = dot3(P1 - c, P1 - c) - R2;
        if (G1 <= 0.0) { // the step ends inside: bracket [0, 1]
            C.lo = 0.0;
            C.Glo = C.G0;
            C.hi = 1.0;
            C.Ghi = G1;
            copy_n<N>(C.yb, C.yn);
            C.side = 0;
            C.it = 0;
            C.phase = C.hi < C.best_th ? SPH_BRACKET : SPH_NEXT;
            continue;
        }
        // probe the path at the chord's closest point
        const D3 dP = P1 - P0;
        const double A = dot3(dP, dP);
        double t1 = A > 0.0 ? -dot3(m, dP) / A : 0.5;
        C.t1 = fmin(fmax(t1, 0.02), 0.98);
        C.ta = 0.0;
        C.Ga = C.G0;
        C.tb = 1.0;
        C.Gb = G1;
        C.th = C.t1;
        C.phase = SPH_PROBE;
    }
}

// Vertex of the parabola through (ta, Ga), (t1, G1), (tb, Gb); false when
// it is not a new interior point.
PRR_HD bool parabola_vertex(const LocateCtx &C, double &tv) {
    const double p = (C.t1 - C.ta) * (C.G1 - C.Gb), q = (C.t1 - C.tb) * (C.G1 - C.Ga);
    const double den = 2.0 * (p - q);
    if (!(den != 0.0)) return false;
    tv = C.t1 - ((C.t1 - C.ta) * p - (C.t1 - C.tb) * q) / den;
    return tv > C.ta && tv < C.tb && fabs(tv - C.t1) >= 1e-9;
}

// The sub-step of fraction C.th gave yt: advance the sphere location.
template <class Mt> PRR_HD PRR_NOINLINE void sphere_update(const FrameDev *frp, StateN<Mt::N> yt, LocateCtx *Cp) {
    constexpr int N = Mt::N;
    LocateCtx &C = *Cp;
    const double G = sphere_G<Mt>(*frp, C.k, yt.v);
    if (C.phase == SPH_BRACKET) { // Illinois regula falsi
        if (G > 0.0) {
            C.lo = C.th;
            C.Glo = G;
            if (C.side == -1) C.Ghi *= 0.5;
            C.side = -1;
        } else {
            C.hi = C.th;
            C.Ghi = G;
            copy_n<N>(C.yb, yt.v);
            if (C.side == 1) C.Glo *= 0.5;
            C.side = 1;
        }
        C.it++;
        return;
    }
    if (G <= 0.0) { // the path is inside at this probe: bracket [0, th]
        C.lo = 0.0;
        C.Glo = C.G0;
        C.hi = C.th;
        C.Ghi = G;
        copy_n<N>(C.yb, yt.v);
        C.side = 0;
        C.it = 0;
        C.phase = C.hi < C.best_th ? SPH_BRACKET : SPH_NEXT;
        return;
    }
    if (C.phase == SPH_PROBE) {
        C.G1 = G;
        C.it = 0;
    } else { // SPH_MIN: fold the probe tv = C.th into the three points
        const double tv = C.th;
        if (G < C.G1) {
            if (tv < C.t1) {
                C.tb = C.t1;
                C.Gb = C.G1;
            } else {
                C.ta = C.t1;
                C.Ga = C.G1;
            }
            C.t1 = tv;
            C.G1 = G;
        } else if (tv < C.t1) {
            C.ta = tv;
            C.Ga = G;
        } else {
            C.tb = tv;
            C.Gb = G;
        }
        C.it++;
    }
    double tv;
    if (C.it < 4 && parabola_vertex(C, tv)) {
        C.th = tv;
        C.phase = SPH_MIN;
    } else {
        C.phase = SPH_NEXT; // the path stays outside: a miss
    }
}

// Event location before the sub-step of this trip (MODE_LOCATE or
// MODE_SPHERE).  Returns kNeedSubstep with hs set, a class (the ray ends),
// or kResume with the lane active again.
template <class Mt>
PRR_HD int locate_prepare(const JobDev &job, Lane<Mt> &L, LocateCtx &C, double &hs, int &mode) {
    if (mode == MODE_LOCATE) {
        hs = L.lam * L.h;
        return kNeedSubstep;
    }
    StateN<Mt::N> y;
    copy_n<Mt::N>(y.v, L.y);
    const SphStep r = sphere_prepare<Mt>(&job.frames[L.frame], y, &C);
    if (r.code == kNeedSubstep) {
        hs = r.hs;
        return kNeedSubstep;
    }
    if (r.code == kSphereHit) {
        mode = MODE_IDLE;
        double yb[Mt::N];
        copy_n<Mt::N>(yb, C.ybest);
        return finish_located(job, L, C.h, yb, C.best_th, PRR_SPHERE, C.best_k);
    }
    // no hit: the step's other events
    mode = MODE_ACTIVE;
    double yn[Mt::N];
    copy_n<Mt::N>(yn, C.yn);
    const int e = escape_wrap_horizon<Mt>(job, L, C, yn, C.h, mode);
    if (e >= 0) {
        mode = MODE_IDLE;
        return e;
    }
    if (mode == MODE_LOCATE) {
        hs = L.lam * L.h;
        return kNeedSubstep;
    }
    return kResume;
}

// The sub-step gave yt: advance the location.  Returns a class (located) or
// -1 (go on locating).
template <class Mt> PRR_HD int locate_update(const JobDev &job, Lane<Mt> &L, LocateCtx &C, const double *yt, int mode) {
    if (mode == MODE_LOCATE) {
        const double r_esc = escape_radius<Mt>(job, job.frames[L.frame], yt);
        double g, dg;
        escape_g<Mt>(yt, r_esc, g, dg);
        const double tol = 1e-13 * (Mt::mino ? 1.0 / r_esc : r_esc);
        if (fabs(g) <= tol || !(dg * L.h != 0.0) || L.hit >= 6)
            return finish_located(job, L, L.h, yt, L.lam, PRR_ESCAPE, -1);
        L.lam = fmin(fmax(L.lam - g / (dg * L.h), 0.0), 1.0);
        L.hit++;
        return -1;
    }
    StateN<Mt::N> y;
    copy_n<Mt::N>(y.v, yt);
    sphere_update<Mt>(&job.frames[L.frame], y, &C);
    return -1;
}

// The tracer as the loop trip reads it.  On the device it is reloaded from
// the kernel's argument segment each trip (scalar loads) instead of staying
// in scalar registers across the persistent loop, where the RK tableau's FP64
// constants leave no room for it and the compiler would spill it to vector
// lanes (a VALU v_readlane per value and trip).  Valid in k_integrate, whose
// only argument is the JobDev (at offset 0 of the argument segment).
PRR_HD TracerDev trip_tracer(const JobDev &job) {
#if defined(__HIP_DEVICE_COMPILE__)
    using KP = const __attribute__((address_space(4))) char *;
    KP kp = (KP)__builtin_amdgcn_kernarg_segment_ptr();
    __asm__ volatile("" : "+s"(kp));
    const __attribute__((address_space(4))) TracerDev *t =
        (const __attribute__((address_space(4))) TracerDev *)(kp + offsetof(JobDev, tr));
    TracerDev r;
    r.solver = t->solver;
    r.precision = t->precision;
    r.h = t->h;
    r.hmin = t->hmin;
    r.hmax = t->hmax;
    r.atol = t->atol;
    r.rtol = t->rtol;
    r.r_escape = t->r_escape;
    r.lambda_max = t->lambda_max;
    r.r_fp64 = t->r_fp64;
    r.inv_r_escape = t->inv_r_escape;
    r.inv_r_fp64 = t->inv_r_fp64;
    r.max_steps = t->max_steps;
    r.max_attempts = t->max_attempts;
    return r;
#else
    return job.tr;
#endif
}

// One loop trip of a lane: an Euler step, a Runge-Kutta attempt (RKF45,
// Dormand-Prince 5(4) or fixed-step RK4), or one sub-step of an event
// location.  Returns the class when the ray ends (mode becomes MODE_IDLE),
// -1 otherwise.
template <class Mt, int SOLVER, bool MIXED>
PRR_HD int lane_iterate(const JobDev &job, Lane<Mt> &L, LocateCtx &C, int &mode) {
    constexpr int N = Mt::N;
    using Tab = typename TableauOf<SOLVER>::type;
    if (SOLVER == PRR_EULER) {
        const int r = euler_step<Mt, MIXED>(job, L);
        if (r >= 0) mode = MODE_IDLE;
        return r;
    }
    // (the FP32-bulk instantiation keeps the tracer in registers: reloading it
    // there moves its stage registers into scratch memory)
    const TracerDev tr = MIXED ? job.tr : trip_tracer(job);
    double hs;
    bool use64 = true, at_min = false;
    const bool locating = mode == MODE_LOCATE || mode == MODE_SPHERE;
    if (locating) {
        const int r = locate_prepare<Mt>(job, L, C, hs, mode);
        if (r == kResume) return -1;
        if (r != kNeedSubstep) {
            if (r >= 0) mode = MODE_IDLE;
            return r;
        }
    } else {
        if (L.steps >= tr.max_steps || L.lam >= tr.lambda_max || L.attempts >= tr.max_attempts) {
            mode = MODE_IDLE;
            return PRR_STEP_CAP;
        }
        L.attempts++;
        if constexpr (Tab::adaptive) {
            hs = rkf_step_size<Mt>(tr, L.y, L.cs, L.h, at_min);
        } else {
            hs = fixed_step_size<Mt>(tr, L.y, L.cs);
            at_min = true; // no rejection
        }
        // FP64 stages: the plain form; in the FP32-bulk form inside r_fp64 and,
        // for Kerr, where the step can come within ~13 degrees of the polar
        // axis (1 - mu^2 and L / (1 - mu^2) lose their relative precision)
        // (the radius, not inside_r_fp64: with the latter the FP32-bulk
        // instantiation's register allocation spills twice as much)
        use64 = !MIXED || radius_of<Mt>(L.y) < tr.r_fp64;
        if (Mt::mino && !use64) {
            const double mu_reach = fabs(L.y[2]) + hs * fabs(L.y[3]);
            use64 = 1.0 - mu_reach * mu_reach < kAxisFP64;
        }
    }
    // FP32-bulk render with whole-wave refills (each wave takes 64 consecutive
    // queue entries at once): a trip takes the FP64 stages for all lanes of
    // the wave when any of them needs them, so that no trip executes both
    // stage forms (on a SIMD, both cost the same).  The decision depends only
    // on those 64 rays, so results do not depend on scheduling.  The adjoint's passes keep the per-lane choice: its
    // replay must repeat each ray's steps bit for bit outside its wave.
    if constexpr (MIXED && !Mt::ckpt && !Mt::replay)
        if (job.refill >= 64) use64 = wave_any(use64);
    if (!locating) L.att64 += use64 ? 1 : 0;
    double d4[N], e[N];
    if constexpr (Mt::mino) {
        if (use64) rkf_stages_c<Mt, double, Tab>(L.y, hs, L.kc.k, d4, e);
        else kerr_stages_fp32_c<Tab>(L.y, hs, L.kc.k, d4, e);
    } else {
        if (use64) rkf_stages<Mt, double, Tab>(L.y, hs, L.cs, d4, e);
        else rkf_stages<Mt, float, Tab>(L.y, hs, L.cs, d4, e);
    }
    double yn[N];
#pragma unroll
    for (int i = 0; i < N; ++i) yn[i] = L.y[i] + d4[i];
    if (locating) {
        const int r = locate_update<Mt>(job, L, C, yn, mode);
        if (r >= 0) mode = MODE_IDLE;
        return r;
    }
    double errn = 0.0; // fixed-step RK4: no estimate, no rejection
    if constexpr (Tab::adaptive) {
        // Kerr: mu = cos th and mu' are measured in units of th (divided by
        // sin th), so that the tolerance keeps its meaning near the poles, where a
        // fixed error in mu is a large relative error in 1 - mu^2 (a scale: FP32)
        const double ist = Mt::mino ? double(rsqrt_f32(fmaxf(float(1.0 - L.y[2] * L.y[2]), 1e-24f))) : 1.0;
        // errn = max_i ei / sc_i; NaN components are ignored as by fmax
#pragma unroll
        for (int i = 0; i < Mt::NERR; ++i) {
            const double sc = tr.atol + tr.rtol * fmax_abs(L.y[i], yn[i]);
            const double ei = (Mt::mino && (i == 2 || i == 3)) ? fabs(e[i]) * ist : fabs(e[i]);
            errn = fmax(errn, ei * rcp_t(sc));
        }
    }
    // Kerr: a step may not reach the polar axis (|cos th| >= 1), where the chart is singular
    if (Mt::mino && !(fabs(yn[2]) < 1.0)) errn = fmax(errn, 1024.0);
    // _h = clamp(0.9 h (eps/maxerr)^0.2, hmin, hmax) (kernels/solver.cuh:175-176)
    // a NaN errn (non-finite stages) is no rejection: the state check below ends the ray
    double hn = hs * step_factor(errn);
    if (!Mt::mino) hn = fmin(fmax(hn, tr.hmin), tr.hmax);
    if (errn > 1.0 && !at_min) { // reject (:185-188)
        L.h = hn;
        return -1;
    }
    if (!finite_state<Mt>(yn)) {
        mode = MODE_IDLE;
        return PRR_NAN;
    }
    if constexpr (Mt::mino) Kerr::project(yn, L.cs);
    L.steps++;
    L.lam = Mt::mino ? yn[5] : L.lam + hs;
    L.h = hn;
    maybe_record<Mt>(job, L, hs);
    const int r = rkf_events<Mt>(job, L, C, yn, hs, mode);
    if (r >= 0) mode = MODE_IDLE;
    return r;
}

} // namespace prr
