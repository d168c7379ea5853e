// frame_host.cpp -- see frame_host.hpp.
#include "frame_host.hpp"

#include <algorithm>
#include <cmath>

#include "geodesic.hpp"
#include "prr_host.hpp"

namespace prr {

// Outer horizon: Kerr r+ = M + sqrt(M^2 - a^2); Reissner-Nordstrom (the
// Schwarzschild-coordinate metric with charge a = q) M + sqrt(M^2 - q^2).
double horizon_radius(int metric, double M, double a) {
    if (metric == PRR_FLAT || metric == PRR_ELLIS) return 0.0;
    if (metric == PRR_KERR_BL || metric == PRR_SCHW_SPH) return M + std::sqrt(std::max(M * M - a * a, 0.0));
    return 2.0 * M;
}

TracerDev tracer_dev(const prr_tracer &t) {
    TracerDev d{};
    d.solver = t.solver;
    d.precision = t.precision;
    d.h = t.h;
    d.hmin = t.hmin;
    d.hmax = t.hmax;
    d.atol = t.atol;
    d.rtol = t.rtol;
    d.r_escape = t.r_escape;
    d.lambda_max = t.lambda_max;
    d.r_fp64 = t.r_fp64;
    d.inv_r_escape = 1.0 / t.r_escape;
    d.inv_r_fp64 = 1.0 / t.r_fp64;
    d.max_steps = t.max_steps;
    d.max_attempts = t.max_attempts;
    return d;
}

FrameDev frame_dev(int metric, const prr_tracer &tr, const prr_frame &f, int64_t begin, double param,
                   double r_esc2) {
    FrameDev d{};
    for (int i = 0; i < 3; ++i) d.pos[i] = f.pos[i];
    double F[3], U[3], R[3];
    camera_basis(f.alpha, f.beta, F, U, R);
    for (int i = 0; i < 3; ++i) {
        d.F[i] = F[i];
        d.U[i] = U[i];
        d.R[i] = R[i];
    }
    d.alpha = f.alpha;
    d.beta = f.beta;
    d.vfov = f.vfov;
    d.tan_half = std::tan(f.vfov / 2.0);
    d.aspect = double(f.width) / double(f.height);
    d.W = f.width;
    d.H = f.height;
    d.s = f.samples;
    const bool grid = f.ssaa == PRR_SSAA_REGULAR || f.ssaa == PRR_SSAA_JITTERED;
    d.n = (!grid || (f.ssaa == PRR_SSAA_REGULAR && f.samples == 1)) ? 0
                                                                      : int(std::lround(std::sqrt(double(f.samples))));
    d.ssaa = f.ssaa;
    d.seed = f.seed;
    d.projection = f.projection;
    d.n_spheres = f.n_spheres;
    d.M = f.mass;
    // the metric's second parameter: Kerr's spin, the charge of the
    // Schwarzschild-coordinate metric (Reissner-Nordstrom), the Ellis throat radius
    d.a = metric == PRR_KERR_BL ? f.spin : ((metric == PRR_SCHW_SPH || metric == PRR_ELLIS) ? param : 0.0);
    const double rh = horizon_radius(metric, f.mass, d.a);
    d.r_hor = (metric == PRR_FLAT || metric == PRR_ELLIS) ? -1.0
                                      : rh * (1.0 + tr.horizon_eps_rel) + tr.horizon_eps_abs;
    d.r_esc2 = r_esc2 > 0.0 ? r_esc2 : tr.r_escape;
    d.inv_r_hor = 1.0 / d.r_hor;
    d.ray_begin = begin;
    d.queue_begin = begin;
    for (int k = 0; k < f.n_spheres; ++k) {
        for (int c = 0; c < 3; ++c) d.sph[k][c] = f.spheres[k].c[c];
        d.sph[k][3] = f.spheres[k].radius;
        d.sph_tex[k] = f.spheres[k].tex;
        for (int c = 0; c < 4; ++c) d.sphf[k][c] = float(d.sph[k][c]);
        d.sphf[k][4] = std::fabs(d.sphf[k][0]) + std::fabs(d.sphf[k][1]) + std::fabs(d.sphf[k][2]) + d.sphf[k][3];
    }
    // initRays' per-frame terms (the camera position through pos_cartTOspher,
    // cudamaths/transform.cuh:15-22, and the Kerr metric there)
    const double px = f.pos[0], py = f.pos[1], pz = f.pos[2];
    const double x2y2 = px * px + py * py, r2 = x2y2 + pz * pz, r = std::sqrt(r2);
    d.cam_r = r;
    d.cam_th = std::acos(pz / r);
    d.cam_ph = std::atan2(py, px);
    d.cam_st = std::sin(d.cam_th);
    d.cam_ct = std::cos(d.cam_th);
    d.cam_ir = 1.0 / r;
    d.cam_x2y2 = x2y2;
    d.cam_ix2y2 = 1.0 / x2y2;
    d.cam_iu1 = 1.0 / (r2 * std::sqrt(x2y2));
    if (metric == PRR_KERR_BL) {
        const KerrG g = kerr_g(f.mass, d.a, r, d.cam_th);
        d.g_tt = g.tt;
        d.g_tph = g.tph;
        d.g_rr = g.rr;
        d.g_sig = g.Sig;
        d.g_phph = g.phph;
    }
    return d;
}

} // namespace prr
