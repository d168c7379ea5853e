// debug.hip -- host instantiations of the device metric plug-ins and of the
// per-ray integrator (include/prr_debug.h).  Compiled as HIP so that the
// __host__ __device__ templates of geodesic.hpp and lane.hpp are available
// to host code.
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/prr_debug.h"
#include "adjoint.hpp"
#include "frame_host.hpp"
#include "prr_host.hpp"

using namespace prr;

namespace {

template <class F> prr_status with_metric(int32_t metric, F &&f) {
    switch (metric) {
    case PRR_FLAT: return f(Flat{});
    case PRR_SCHW_KS_CART: return f(SchwKS{});
    case PRR_SCHW_SPH: return f(SchwSph{});
    case PRR_KERR_BL: return f(Kerr{});
    case PRR_ELLIS: return f(Ellis{});
    default: set_global_error("unknown metric"); return PRR_EINVAL;
    }
}

} // namespace

namespace {

// One ray through the host run of the pipeline: init_ray, lane_iterate until
// the ray ends, shade_ray; `log` receives one row per accepted step.
template <class Mt>
prr_status host_ray(JobDev &job, const prr_tracer &tr, int64_t id, double *log, int max_n, int *n_log) {
    constexpr int W = rs_width<Mt>();
    double rsy[W];
    int8_t cls1 = 0, hit1 = -1;
    int32_t steps1 = 0, att1 = 0;
    job.rs = RayStateDev{rsy, &cls1, &hit1, &steps1, &att1};
    job.rs_total = 1;
    job.rs_base = id;
    init_ray<Mt>(job, id);
    Lane<Mt> L;
    lane_load<Mt>(job, L, id);
    LocateCtx C{};
    int mode = MODE_ACTIVE;
    const bool mixed = tr.precision == PRR_FP32_BULK;
    int n = 0;
    if (!finite_state<Mt>(L.y)) {
        lane_store<Mt>(job, L, PRR_NAN);
        mode = MODE_IDLE;
    }
    while (mode != MODE_IDLE) {
        const int32_t s0 = L.steps, a64 = L.att64;
        const double h_before = L.h;
        const int was = mode;
        int c;
        if (tr.solver == PRR_EULER)
            c = mixed ? lane_iterate<Mt, PRR_EULER, true>(job, L, C, mode)
                      : lane_iterate<Mt, PRR_EULER, false>(job, L, C, mode);
        else if (tr.solver == PRR_DORMAND)
            c = mixed ? lane_iterate<Mt, PRR_DORMAND, true>(job, L, C, mode)
                      : lane_iterate<Mt, PRR_DORMAND, false>(job, L, C, mode);
        else if (tr.solver == PRR_RK4)
            c = mixed ? lane_iterate<Mt, PRR_RK4, true>(job, L, C, mode)
                      : lane_iterate<Mt, PRR_RK4, false>(job, L, C, mode);
        else
            c = mixed ? lane_iterate<Mt, PRR_RKF45, true>(job, L, C, mode)
                      : lane_iterate<Mt, PRR_RKF45, false>(job, L, C, mode);
        if (log && was == MODE_ACTIVE && L.steps != s0 && n < max_n) {
            double *row = log + 12 * n++;
            row[0] = L.steps;
            row[1] = L.attempts;
            row[2] = h_before;
            row[3] = L.att64 - a64;
            // the accepted state (also while its events are being located)
            const double *yy = (mode == MODE_LOCATE || mode == MODE_SPHERE) ? C.yn : L.y;
            for (int i = 0; i < 8; ++i) row[4 + i] = i < Mt::N ? yy[i] : 0.0;
        }
        if (c >= 0) lane_store<Mt>(job, L, c);
    }
    shade_ray<Mt>(job, id);
    if (n_log) *n_log = n;
    return PRR_OK;
}

} // namespace

extern "C" {

int prr_debug_state_size(int32_t metric) { return metric == PRR_KERR_BL ? Kerr::N : 8; }

prr_status prr_debug_rhs(int32_t metric, double M, double a, double E, double L, double Q, const double *y,
                         int32_t fp32, double *f) {
    if (!y || !f) return PRR_EINVAL;
    const Consts<double> cs{M, a, E, L, Q};
    return with_metric(metric, [&](auto mt) {
        using Mt = decltype(mt);
        if (fp32) {
            float yf[Mt::N], ff[Mt::N];
            for (int i = 0; i < Mt::N; ++i) yf[i] = float(y[i]);
            Mt::template rhs<float>(yf, Mt::template coef<float>(cs), ff);
            for (int i = 0; i < Mt::N; ++i) f[i] = ff[i];
        } else {
            Mt::template rhs<double>(y, Mt::template coef<double>(cs), f);
        }
        return PRR_OK;
    });
}

prr_status prr_debug_kerr_delta_rhs(double M, double a, double E, double L, double Q, const double *y0,
                                    const double *d, double *g) {
    if (!y0 || !d || !g) return PRR_EINVAL;
    double F[6];
    KerrDelta P;
    Kerr::delta_prep(y0, Kerr::coef<double>(Consts<double>{M, a, E, L, Q}), F, P);
    float df[6], gf[6];
    for (int i = 0; i < 6; ++i) df[i] = float(d[i]);
    Kerr::delta_rhs(df, P, gf);
    for (int i = 0; i < 6; ++i) g[i] = gf[i];
    return PRR_OK;
}

prr_status prr_debug_kerr_stages(double M, double a, double E, double L, double Q, const double *y, double h,
                                 int32_t fp32, double *d4, double *e) {
    if (!y || !d4 || !e) return PRR_EINVAL;
    const Consts<double> cs{M, a, E, L, Q};
    if (fp32) kerr_stages_fp32(y, h, cs, d4, e);
    else rkf_stages<Kerr, double>(y, h, cs, d4, e);
    return PRR_OK;
}

prr_status prr_debug_rhs_vjp(double M, double a, double E, double L, double Q, const double *y, const double *fb,
                             double *yb, double *cb) {
    if (!y || !fb || !yb || !cb) return PRR_EINVAL;
    const Consts<double> cs{M, a, E, L, Q};
    KerrCoef<double> kb{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    Kerr::rhs_vjp(y, Kerr::coef<double>(cs), fb, yb, kb);
    Kerr::coef_vjp(cs, kb, cb);
    return PRR_OK;
}

prr_status prr_debug_orbit_init(double M, double a, double r_cam, double incl, double cam_phi, double vfov,
                                double aspect, double u, double v, double z[8], double dz[24]) {
    if (!z) return PRR_EINVAL;
    using D = Dual<3>;
    const D da = D::var(a, 0), dr = D::var(r_cam, 1), di = D::var(incl, 2);
    V3<D> pos, d;
    orbit_camera_ray<D>(dr, di, cam_phi, std::tan(vfov / 2.0), aspect, u, v, pos, d);
    D zz[8];
    kerr_init_t<D>(pos, d, M, da, zz);
    for (int i = 0; i < 8; ++i) {
        z[i] = zz[i].v;
        if (dz)
            for (int j = 0; j < 3; ++j) dz[3 * i + j] = zz[i].d[j];
    }
    return PRR_OK;
}

prr_status prr_debug_kerr_step(double M, double a, double E, double L, double Q, const double y[6], double h,
                               double out[6]) {
    if (!y || !out) return PRR_EINVAL;
    kerr_step(y, h, Consts<double>{M, a, E, L, Q}, out);
    return PRR_OK;
}

prr_status prr_debug_kerr_step_vjp(double M, double a, double E, double L, double Q, const double y[6], double h,
                                   double lam[9]) {
    if (!y || !lam) return PRR_EINVAL;
    kerr_step_vjp(y, h, Consts<double>{M, a, E, L, Q}, lam);
    return PRR_OK;
}

prr_status prr_debug_escape_color(double M, double a, double E, double L, double Q, const double y[5],
                                  const uint8_t *rgb8, int w, int h, double c[3], double dc[27]) {
    if (!y || !rgb8 || !c || w < 2 || h < 2) return PRR_EINVAL;
    std::vector<float4> tx(size_t(w) * h);
    for (size_t i = 0; i < tx.size(); ++i)
        tx[i] = make_float4(float(rgb8[3 * i]) / 255.0f, float(rgb8[3 * i + 1]) / 255.0f,
                            float(rgb8[3 * i + 2]) / 255.0f, 1.0f);
    TexDev t{tx.data(), w, h, PRR_FILTER_BILINEAR, 0};
    using D = Dual<9>;
    D yy[5];
    for (int i = 0; i < 5; ++i) yy[i] = D::var(y[i], i);
    const D dE = D::var(E, 5), dL = D::var(L, 6), dQ = D::var(Q, 7), da = D::var(a, 8);
    const V3<D> n = kerr_escape_dir_t<D>(yy, dE, dL, dQ, da, M);
    D col[3];
    sky_color_t<D>(t, n, col);
    for (int k = 0; k < 3; ++k) {
        c[k] = col[k].v;
        if (dc)
            for (int j = 0; j < 9; ++j) dc[9 * k + j] = col[k].d[j];
    }
    return PRR_OK;
}

prr_status prr_debug_state_from_xu(int32_t metric, double M, double a, const double x[4], const double u[4],
                                   double *y, double ELQ[3]) {
    if (!x || !u || !y) return PRR_EINVAL;
    return with_metric(metric, [&](auto mt) {
        using Mt = decltype(mt);
        Consts<double> cs;
        state_from_xu<Mt>(x, u, M, a, y, cs);
        if (ELQ) {
            ELQ[0] = cs.E;
            ELQ[1] = cs.L;
            ELQ[2] = cs.Q;
        }
        return PRR_OK;
    });
}

prr_status prr_debug_xu_from_state(int32_t metric, double M, double a, double E, double L, double Q, const double *y,
                                   double x[4], double u[4]) {
    if (!x || !u || !y) return PRR_EINVAL;
    return with_metric(metric, [&](auto mt) {
        using Mt = decltype(mt);
        xu_from_state<Mt>(y, Consts<double>{M, a, E, L, Q}, x, u);
        return PRR_OK;
    });
}

prr_status prr_debug_init_ray(int32_t metric, double M, double a, const double pos[3], const double d[3], double *y,
                              double ELQ[3]) {
    if (!pos || !d || !y) return PRR_EINVAL;
    return with_metric(metric, [&](auto mt) {
        using Mt = decltype(mt);
        Consts<double> cs;
        init_state<Mt>(d3(pos[0], pos[1], pos[2]), d3(d[0], d[1], d[2]), M, a, y, cs);
        if (ELQ) {
            ELQ[0] = cs.E;
            ELQ[1] = cs.L;
            ELQ[2] = cs.Q;
        }
        return PRR_OK;
    });
}

prr_status prr_debug_queue_rays(const prr_frame *f, const int64_t *q, int64_t n, int64_t *ids) {
    if (!f || !q || !ids || n < 0) return PRR_EINVAL;
    prr_tracer tr{};
    const FrameDev fr = frame_dev(PRR_FLAT, tr, *f, 0);
    JobDev job{};
    job.frames = &fr;
    job.nframes = 1;
    job.total = int64_t(f->width) * f->height * f->samples;
    job.qtotal = job.total;
    for (int64_t i = 0; i < n; ++i) {
        if (q[i] < 0 || q[i] >= job.total) return PRR_EINVAL;
        ids[i] = queue_ray(job, q[i]);
    }
    return PRR_OK;
}

prr_status prr_camera_basis(double alpha, double beta, double forward[3], double up[3], double right[3]) {
    if (!forward || !up || !right) return PRR_EINVAL;
    camera_basis(alpha, beta, forward, up, right);
    return PRR_OK;
}

prr_status prr_camera_dir(const prr_frame *f, double u, double v, double d[3]) {
    if (!f || !d || f->width <= 0 || f->height <= 0) return PRR_EINVAL;
    prr_tracer tr{};
    const FrameDev fr = frame_dev(PRR_FLAT, tr, *f, 0);
    const D3 r = camera_dir(fr, u, v);
    d[0] = r.x;
    d[1] = r.y;
    d[2] = r.z;
    return PRR_OK;
}

prr_status prr_debug_camera_ray(const prr_frame *f, int64_t local, double d[3]) {
    if (!f || !d) return PRR_EINVAL;
    prr_tracer tr{};
    const FrameDev fr = frame_dev(PRR_FLAT, tr, *f, 0);
    const D3 v = camera_ray(fr, local);
    d[0] = v.x;
    d[1] = v.y;
    d[2] = v.z;
    return PRR_OK;
}

prr_status prr_debug_trace_rays(int32_t metric, const prr_tracer *tracer, const prr_frame *f, const double *x0u0,
                                const int64_t *ids, int n, int32_t *cls, double *endpoint, int32_t *steps,
                                int32_t *attempts, double *state8) try {
    if (!tracer || !f || n < 0 || (!x0u0 && !ids)) return PRR_EINVAL;
    const bool adaptive = tracer->solver == PRR_RKF45 || tracer->solver == PRR_DORMAND;
    if ((adaptive && !(tracer->hmin > 0.0)) || (tracer->solver != PRR_EULER && tracer->max_attempts > 100000000LL))
        return PRR_EINVAL;
    const FrameDev fr = frame_dev(metric, *tracer, *f, 0);
    JobDev job{};
    job.frames = &fr;
    job.nframes = 1;
    job.refill = 1;
    job.total = x0u0 ? n : int64_t(f->width) * f->height * f->samples;
    job.explicit_x0u0 = x0u0;
    job.tr = tracer_dev(*tracer);
    int8_t c1;
    double e3[3], s8[8];
    int32_t st1, at1;
    job.out.cls = &c1;
    job.out.endpoint = e3;
    job.out.steps = &st1;
    job.out.attempts = &at1;
    job.out.state8 = s8;
    return with_metric(metric, [&](auto mt) {
        using Mt = decltype(mt);
        for (int i = 0; i < n; ++i) {
            const int64_t id = x0u0 ? i : ids[i];
            if (id < 0 || id >= job.total) return PRR_EINVAL;
            host_ray<Mt>(job, *tracer, id, nullptr, 0, nullptr);
            if (cls) cls[i] = c1;
            if (steps) steps[i] = st1;
            if (attempts) attempts[i] = at1;
            for (int k = 0; k < 3; ++k)
                if (endpoint) endpoint[3 * i + k] = e3[k];
            for (int k = 0; k < 8; ++k)
                if (state8) state8[8 * i + k] = s8[k];
        }
        return PRR_OK;
    });
} catch (...) {
    return prr::abi_exception(nullptr);
}

prr_status prr_debug_trace_log(int32_t metric, const prr_tracer *tracer, const prr_frame *f, int64_t id, int max_n,
                               double *log, int *n_out) try {
    if (!tracer || !f || !log || !n_out || max_n <= 0) return PRR_EINVAL;
    const FrameDev fr = frame_dev(metric, *tracer, *f, 0);
    JobDev job{};
    job.frames = &fr;
    job.nframes = 1;
    job.refill = 1;
    job.total = int64_t(f->width) * f->height * f->samples;
    job.tr = tracer_dev(*tracer);
    if (id < 0 || id >= job.total) return PRR_EINVAL;
    return with_metric(metric, [&](auto mt) {
        using Mt = decltype(mt);
        return host_ray<Mt>(job, *tracer, id, log, max_n, n_out);
    });
} catch (...) {
    return prr::abi_exception(nullptr);
}

} // extern "C"
