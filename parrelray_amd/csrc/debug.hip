// debug.hip -- host instantiations of the device metric plug-ins
// (include/prr_debug.h).  Compiled as HIP so that geodesic.hpp's
// __host__ __device__ templates are available to host code.
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/prr_debug.h"
#include "adjoint.hpp"
#include "geodesic.hpp"
#include "prr_host.hpp"

using namespace prr;

namespace {

template <class F> prr_status with_metric(int32_t metric, F &&f) {
    switch (metric) {
    case PRR_FLAT: return f(Flat{});
    case PRR_SCHW_KS_CART: return f(SchwKS{});
    case PRR_SCHW_SPH: return f(SchwSph{});
    case PRR_KERR_BL: return f(Kerr{});
    default: set_global_error("unknown metric"); return PRR_EINVAL;
    }
}

} // namespace

extern "C" {

int prr_debug_state_size(int32_t metric) { return metric == PRR_KERR_BL ? Kerr::N : 8; }

prr_status prr_debug_rhs(int32_t metric, double M, double a, double E, double L, const double *y, int32_t fp32,
                         double *f) {
    if (!y || !f) return PRR_EINVAL;
    return with_metric(metric, [&](auto mt) {
        using Mt = decltype(mt);
        if (fp32) {
            float yf[Mt::N], ff[Mt::N];
            for (int i = 0; i < Mt::N; ++i) yf[i] = float(y[i]);
            Mt::template rhs<float>(yf, Consts<float>{float(M), float(a), float(E), float(L)}, ff);
            for (int i = 0; i < Mt::N; ++i) f[i] = ff[i];
        } else {
            Mt::template rhs<double>(y, Consts<double>{M, a, E, L}, f);
        }
        return PRR_OK;
    });
}

prr_status prr_debug_rhs_vjp(int32_t metric, double M, double a, double E, double L, const double *y,
                             const double *fb, double *yb, double *cb) {
    if (metric != PRR_KERR_BL) {
        set_global_error("rhs_vjp: Kerr only");
        return PRR_EINVAL;
    }
    if (!y || !fb || !yb || !cb) return PRR_EINVAL;
    Kerr::rhs_vjp(y, Consts<double>{M, a, E, L}, fb, yb, cb);
    return PRR_OK;
}

prr_status prr_debug_orbit_init(double M, double a, double r_cam, double incl, double cam_phi, double vfov,
                                double aspect, double u, double v, double z[7], double dz[21]) {
    if (!z) return PRR_EINVAL;
    using D = Dual<3>;
    const D da = D::var(a, 0), dr = D::var(r_cam, 1), di = D::var(incl, 2);
    V3<D> pos, d;
    orbit_camera_ray<D>(dr, di, cam_phi, std::tan(vfov / 2.0), aspect, u, v, pos, d);
    D zz[7];
    kerr_init_t<D>(pos, d, M, da, zz);
    for (int i = 0; i < 7; ++i) {
        z[i] = zz[i].v;
        if (dz)
            for (int j = 0; j < 3; ++j) dz[3 * i + j] = zz[i].d[j];
    }
    return PRR_OK;
}

prr_status prr_debug_kerr_step(double M, double a, double E, double L, const double y[5], double h, double out[5]) {
    if (!y || !out) return PRR_EINVAL;
    kerr_step(y, h, Consts<double>{M, a, E, L}, out);
    return PRR_OK;
}

prr_status prr_debug_kerr_step_vjp(double M, double a, double E, double L, const double y[5], double h,
                                   double lam[8]) {
    if (!y || !lam) return PRR_EINVAL;
    kerr_step_vjp(y, h, Consts<double>{M, a, E, L}, lam);
    return PRR_OK;
}

prr_status prr_debug_escape_color(double M, double a, double E, double L, const double y[5], const uint8_t *rgb8,
                                  int w, int h, double c[3], double dc[24]) {
    if (!y || !rgb8 || !c || w < 2 || h < 2) return PRR_EINVAL;
    std::vector<float4> tx(size_t(w) * h);
    for (size_t i = 0; i < tx.size(); ++i)
        tx[i] = make_float4(float(rgb8[3 * i]) / 255.0f, float(rgb8[3 * i + 1]) / 255.0f,
                            float(rgb8[3 * i + 2]) / 255.0f, 1.0f);
    TexDev t{tx.data(), w, h, PRR_FILTER_BILINEAR, 0};
    using D = Dual<8>;
    D yy[5];
    for (int i = 0; i < 5; ++i) yy[i] = D::var(y[i], i);
    const D dE = D::var(E, 5), dL = D::var(L, 6), da = D::var(a, 7);
    const V3<D> n = kerr_escape_dir_t<D>(yy, dE, dL, da, M);
    D col[3];
    sky_color_t<D>(t, n, col);
    for (int k = 0; k < 3; ++k) {
        c[k] = col[k].v;
        if (dc)
            for (int j = 0; j < 8; ++j) dc[8 * k + j] = col[k].d[j];
    }
    return PRR_OK;
}

prr_status prr_debug_state_from_xu(int32_t metric, double M, double a, const double x[4], const double u[4],
                                   double *y, double EL[2]) {
    if (!x || !u || !y) return PRR_EINVAL;
    return with_metric(metric, [&](auto mt) {
        using Mt = decltype(mt);
        Consts<double> cs;
        state_from_xu<Mt>(x, u, M, a, y, cs);
        if (EL) {
            EL[0] = cs.E;
            EL[1] = cs.L;
        }
        return PRR_OK;
    });
}

prr_status prr_debug_xu_from_state(int32_t metric, double M, double a, double E, double L, const double *y,
                                   double x[4], double u[4]) {
    if (!x || !u || !y) return PRR_EINVAL;
    return with_metric(metric, [&](auto mt) {
        using Mt = decltype(mt);
        xu_from_state<Mt>(y, Consts<double>{M, a, E, L}, x, u);
        return PRR_OK;
    });
}

prr_status prr_debug_init_ray(int32_t metric, double M, double a, const double pos[3], const double d[3], double *y,
                              double EL[2]) {
    if (!pos || !d || !y) return PRR_EINVAL;
    return with_metric(metric, [&](auto mt) {
        using Mt = decltype(mt);
        Consts<double> cs;
        init_state<Mt>(d3(pos[0], pos[1], pos[2]), d3(d[0], d[1], d[2]), M, a, y, cs);
        if (EL) {
            EL[0] = cs.E;
            EL[1] = cs.L;
        }
        return PRR_OK;
    });
}

prr_status prr_debug_camera_ray(const prr_frame *f, int64_t local, double d[3]) {
    if (!f || !d) return PRR_EINVAL;
    FrameDev fr{};
    double F[3], U[3], R[3];
    camera_basis(f->alpha, f->beta, F, U, R);
    for (int i = 0; i < 3; ++i) {
        fr.F[i] = F[i];
        fr.U[i] = U[i];
        fr.R[i] = R[i];
    }
    fr.alpha = f->alpha;
    fr.beta = f->beta;
    fr.vfov = f->vfov;
    fr.tan_half = std::tan(f->vfov / 2.0);
    fr.aspect = double(f->width) / double(f->height);
    fr.W = f->width;
    fr.H = f->height;
    fr.s = f->samples;
    fr.n = (f->ssaa == PRR_SSAA_NONE || f->samples == 1) ? 0 : int(std::lround(std::sqrt(double(f->samples))));
    fr.projection = f->projection;
    const D3 v = camera_ray(fr, local);
    d[0] = v.x;
    d[1] = v.y;
    d[2] = v.z;
    return PRR_OK;
}

} // extern "C"
