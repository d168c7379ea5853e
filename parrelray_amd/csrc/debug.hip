// debug.hip -- host instantiations of the device metric plug-ins
// (include/prr_debug.h).  Compiled as HIP so that geodesic.hpp's
// __host__ __device__ templates are available to host code.
#include <hip/hip_runtime.h>

#include "../../include/prr_debug.h"
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
