// oracle_capi.cpp -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
// C entry points of the CPU FP64 oracle (prr_oracle.hpp) for ctypes, used by
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg only.
#include <omp.h>

#include <chrono>

#include "prr_oracle.hpp"

using namespace prro;

namespace {
int threads_or_default(int n) { return n > 0 ? n : omp_get_max_threads(); }
// Rays per OpenMP chunk: 64 (the schedule(dynamic, 64) of the CPU baseline plan)
// unless that leaves threads idle on a small ray set.
int chunk_for(int64_t n, int threads) {
    return int(std::max<int64_t>(1, std::min<int64_t>(64, n / (8 * int64_t(threads)))));
}
} // namespace

extern "C" {

void *orc_scene_new(int metric, double M, double a, const prr_tracer *tr) {
    Scene *S = new Scene();
    S->metric = metric;
    S->M = M;
    S->a = a;
    if (tr) S->tr = *tr;
    return S;
}

void orc_scene_free(void *s) { delete static_cast<Scene *>(s); }

void orc_scene_set_far_escape(void *s, double r) { static_cast<Scene *>(s)->r_esc2 = r; }

// 1: integrate in FP32 (the reference-like precision of the CPU baseline); 0: FP64.
void orc_scene_set_fp32(void *s, int on) { static_cast<Scene *>(s)->fp32 = on != 0; }

void orc_scene_set_tracer(void *s, const prr_tracer *tr) { static_cast<Scene *>(s)->tr = *tr; }

void orc_scene_set_texture(void *s, int slot, const uint8_t *rgb8, int w, int h, int filter) {
    if (slot < 0 || slot >= PRR_MAX_TEXTURES) return;
    texture_from_rgb8(static_cast<Scene *>(s)->tex[slot], rgb8, w, h, filter);
}

void orc_scene_set_spheres(void *s, const prr_sphere *sph, int n) {
    static_cast<Scene *>(s)->spheres.assign(sph, sph + n);
}

// Rays of one frame (ray_ids == NULL: all rays in order).  Any output may be NULL.
// Returns 0; *seconds receives the wall time of integration + shading.
int orc_render_rays(void *scene, const prr_frame *f, const int64_t *ray_ids, int64_t n, double delta, int axis,
                    int8_t *cls, double *endpoint, int32_t *steps, int32_t *attempts, float *color, double *state8,
                    int nthreads, double *seconds, double *color64) {
    const Scene &S = *static_cast<Scene *>(scene);
    const Basis B = camera_basis(f->alpha, f->beta);
    const int64_t total = int64_t(f->width) * f->height * f->samples;
    if (!ray_ids) n = total;
    const int threads = threads_or_default(nthreads);
    const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic, chunk_for(n, threads)) num_threads(threads)
    for (int64_t i = 0; i < n; ++i) {
        const int64_t ray = ray_ids ? ray_ids[i] : i;
        const RayResult R = render_ray(S, *f, B, ray, delta, axis);
        if (cls) cls[i] = int8_t(R.cls);
        if (endpoint) {
            endpoint[3 * i] = R.endpoint.x;
            endpoint[3 * i + 1] = R.endpoint.y;
            endpoint[3 * i + 2] = R.endpoint.z;
        }
        if (steps) steps[i] = int32_t(R.steps);
        if (attempts) attempts[i] = int32_t(R.attempts);
        if (color) {
            color[3 * i] = R.color.r;
            color[3 * i + 1] = R.color.g;
            color[3 * i + 2] = R.color.b;
        }
        if (color64) {
            color64[3 * i] = R.color64.r;
            color64[3 * i + 1] = R.color64.g;
            color64[3 * i + 2] = R.color64.b;
        }
        if (state8)
            for (int c = 0; c < 4; ++c) {
                state8[8 * i + c] = R.x[c];
                state8[8 * i + 4 + c] = R.u[c];
            }
    }
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    return 0;
}

// Integrator alone from explicit (x, u) initial states, with the scene's
// spacetime, tracer and spheres.
int orc_trace(void *scene, const double *x0u0, int64_t n, double *xend, int32_t *cls, int32_t *steps,
              int32_t *attempts, int nthreads) {
    const Scene &S = *static_cast<Scene *>(scene);
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads_or_default(nthreads))
    for (int64_t i = 0; i < n; ++i) {
        RayResult R = trace(S, S.M, S.a, S.spheres, x0u0 + 8 * i, x0u0 + 8 * i + 4);
        if (xend)
            for (int c = 0; c < 4; ++c) {
                xend[8 * i + c] = R.x[c];
                xend[8 * i + 4 + c] = R.u[c];
            }
        if (cls) cls[i] = R.cls;
        if (steps) steps[i] = int32_t(R.steps);
        if (attempts) attempts[i] = int32_t(R.attempts);
    }
    return 0;
}

void orc_ssaa(const float *ray_rgb, int W, int H, int s, float *img) { ssaa_combine(ray_rgb, W, H, s, img); }
void orc_fxaa(const float *img, int W, int H, float *out) { fxaa(img, W, H, out); }
void orc_rgb8(const float *img, int64_t n, uint8_t *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = to_rgb8(img[i]);
}

// Whole frame: rays, SSAA combine, optional FXAA, RGB8.
int orc_render_frame(void *scene, const prr_frame *f, float *ray_color, float *img, uint8_t *rgb8, int nthreads) {
    const int64_t total = int64_t(f->width) * f->height * f->samples;
    std::vector<float> col(size_t(total) * 3);
    orc_render_rays(scene, f, nullptr, total, 0.0, 0, nullptr, nullptr, nullptr, nullptr, col.data(), nullptr,
                    nthreads, nullptr, nullptr);
    if (ray_color) std::copy(col.begin(), col.end(), ray_color);
    std::vector<float> ss(size_t(f->width) * f->height * 3);
    ssaa_combine(col.data(), f->width, f->height, f->samples, ss.data());
    if (f->fxaa)
        fxaa(ss.data(), f->width, f->height, img);
    else
        std::copy(ss.begin(), ss.end(), img);
    if (rgb8) orc_rgb8(img, int64_t(f->width) * f->height * 3, rgb8);
    return 0;
}

void orc_metric(int kind, double M, double a, const double *x, double *g) { metric(kind, M, a, x, g); }
void orc_dmetric_kerr(double M, double a, const double *x, double *dg) {
    double d[4][16];
    dmetric_kerr(x, M, a, d);
    for (int k = 0; k < 4; ++k)
        for (int i = 0; i < 16; ++i) dg[16 * k + i] = d[k][i];
}
void orc_christoffel_kerr(double M, double a, const double *x, double *G) {
    double g[4][4][4];
    christoffel_kerr(x, M, a, g);
    std::memcpy(G, g, sizeof g);
}
void orc_geod(int kind, double M, double a, const double *x, const double *u, double *du) {
    geod(kind, M, a, x, u, du);
}
void orc_kerr_constants(double M, double a, const double *x, const double *u, double *out3) {
    kerr_constants(x, u, M, a, out3);
}
void orc_camera_dir(const prr_frame *f, double u, double v, double *d3) {
    const V3 d = ray_dir(*f, camera_basis(f->alpha, f->beta), u, v);
    d3[0] = d.x;
    d3[1] = d.y;
    d3[2] = d.z;
}
void orc_camera_basis(double alpha, double beta, double *fur9) {
    const Basis B = camera_basis(alpha, beta);
    const V3 v[3] = {B.F, B.U, B.R};
    for (int i = 0; i < 3; ++i) {
        fur9[3 * i] = v[i].x;
        fur9[3 * i + 1] = v[i].y;
        fur9[3 * i + 2] = v[i].z;
    }
}
void orc_sample_uv(const prr_frame *f, int64_t ray, double *uv) { sample_uv(*f, ray, uv[0], uv[1]); }
void orc_init_ray(int kind, double M, double a, const double *pos, const double *d, double *x, double *u) {
    init_ray(kind, M, a, v3(pos[0], pos[1], pos[2]), v3(d[0], d[1], d[2]), x, u);
}
void orc_sky_lookup(void *scene, double theta, double phi, float *rgb) {
    const RGB c = sky_lookup(static_cast<Scene *>(scene)->tex[0], theta, phi);
    rgb[0] = c.r;
    rgb[1] = c.g;
    rgb[2] = c.b;
}
void orc_bilinear(void *scene, int slot, double u, double v, float *rgb) {
    const RGB c = bilinear(static_cast<Scene *>(scene)->tex[slot], u, v);
    rgb[0] = c.r;
    rgb[1] = c.g;
    rgb[2] = c.b;
}
int orc_chord_hit(const prr_sphere *sph, int n, const double *P0, const double *P1, double *t) {
    std::vector<prr_sphere> v(sph, sph + n);
    return chord_hit(v, v3(P0[0], P0[1], P0[2]), v3(P1[0], P1[1], P1[2]), *t);
}
// Tableau of a solver: which 0 = A[i][j], 1 = propagated b[i], 2 = embedded bh[i], 3 = stages.
double orc_tableau(int solver, int which, int i, int j) {
    const Tableau &T = tableau_of(solver);
    if (which == 0) return (i < 7 && j < 6) ? T.A[i][j] : 0.0;
    if (which == 1) return T.b[i];
    if (which == 2) return T.bh[i];
    return T.S;
}
}
