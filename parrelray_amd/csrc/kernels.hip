// kernels.hip -- gfx950 kernels: the persistent geodesic integrator and the
// SSAA / FXAA / RGB8 passes.
//
// Integrator (replaces initRays + trace_Euler/trace_RKF + rayToRGB,
// kernels/solver.cuh:26-220, metrics/schwarzschild.cuh:151-183, launched by
// prr.cpp:279-323 as one thread per ray in 64-row bands):
//   * one persistent grid sized to the CUs; every lane holds one ray's state
//     in VGPRs and runs one Euler step / RKF45 attempt per loop trip;
//   * when at least `refill` lanes of a wave are idle the wave claims that
//     many ray ids from a device queue with a single atomic (ballot + mbcnt
//     prefix sum) and initialises them in place, so lanes never wait for the
//     slowest ray of their wave;
//   * rays are generated from their id (no ray buffer), coloured on
//     termination, and only the colour (plus optional diagnostics) is written.
#include <hip/hip_runtime.h>

#include "adjoint.hpp"
#include "kernels.hpp"
#include "lane.hpp"

// Register budget of the integrator: waves per SIMD the compiler must fit
// (build-time knob, -DPRR_WAVES_PER_EU=n; unset = the compiler's choice).
#ifdef PRR_WAVES_PER_EU
#define PRR_OCCUPANCY __attribute__((amdgpu_waves_per_eu(PRR_WAVES_PER_EU, 8)))
#else
#define PRR_OCCUPANCY
#endif

namespace prr {

__device__ inline unsigned long long shfl_u64(unsigned long long v, int src) {
    const int lo = __shfl(int(v & 0xffffffffull), src);
    const int hi = __shfl(int(v >> 32), src);
    return (unsigned long long)(unsigned(lo)) | ((unsigned long long)(unsigned(hi)) << 32);
}

struct KerrStepOut {
    Lane<Kerr> L;
    int cls;
};
template <bool MIXED> __device__ __noinline__ KerrStepOut kerr_attempt_shared(const JobDev &job, const Lane<Kerr> L);

// SHARED: call the adjoint's out-of-line attempt function (Kerr, RKF45).
template <class Mt, int SOLVER, bool MIXED, bool SHARED>
__global__ void __launch_bounds__(256) PRR_OCCUPANCY k_integrate(JobDev job) {
    const int lane = int(threadIdx.x & 63);
    Lane<Mt> L;
    bool active = false, exhausted = false;
    unsigned long long acc_att = 0, acc_ok = 0, acc_64 = 0, acc_rays = 0;
    while (true) {
        const unsigned long long idle = __ballot(!active);
        const int nidle = __popcll(idle);
        if (!exhausted && nidle >= job.refill) {
            const int leader = __ffsll((long long)idle) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(&job.ctr->queue, (unsigned long long)nidle);
            base = shfl_u64(base, leader);
            if (base + (unsigned long long)nidle >= (unsigned long long)job.total) exhausted = true;
            if (!active) {
                const unsigned rank =
                    __builtin_amdgcn_mbcnt_hi(unsigned(idle >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(idle), 0u));
                const int64_t id = int64_t(base) + rank;
                if (id < job.total) {
                    L = lane_start<Mt>(job, id);
                    active = true;
                    if (!finite_state<Mt>(L.y)) {
                        lane_finish<Mt>(job, L, PRR_NAN);
                        acc_rays++;
                        active = false;
                    }
                }
            }
        }
        if (__ballot(active) == 0ull) break;
        if (active) {
            int cls;
            if constexpr (SHARED) {
                const KerrStepOut r = kerr_attempt_shared<MIXED>(job, L);
                L = r.L;
                cls = r.cls;
            }
            else cls = lane_attempt<Mt, SOLVER, MIXED>(job, L);
            if (cls >= 0) {
                lane_finish<Mt>(job, L, cls);
                acc_att += (unsigned long long)L.attempts;
                acc_ok += (unsigned long long)L.steps;
                acc_64 += (unsigned long long)L.att64;
                acc_rays++;
                active = false;
            }
        }
    }
    // one atomic per wave for the counters
    for (int off = 32; off > 0; off >>= 1) {
        acc_att += shfl_u64(acc_att, lane ^ off);
        acc_ok += shfl_u64(acc_ok, lane ^ off);
        acc_64 += shfl_u64(acc_64, lane ^ off);
        acc_rays += shfl_u64(acc_rays, lane ^ off);
    }
    if (lane == 0) {
        atomicAdd(&job.ctr->attempts, acc_att);
        atomicAdd(&job.ctr->accepted, acc_ok);
        atomicAdd(&job.ctr->attempts_fp64, acc_64);
        atomicAdd(&job.ctr->rays, acc_rays);
    }
}

// Counter reset by a kernel (no memset node, so the sequence can be captured
// in a graph).
__global__ void k_reset_counters(CountersDev *c) {
    if (threadIdx.x == 0) {
        c->queue = 0;
        c->attempts = 0;
        c->accepted = 0;
        c->attempts_fp64 = 0;
        c->rays = 0;
        c->mismatch = 0;
    }
}

// ----------------------------------------------------------------- AA passes
__device__ inline float clamp01(float c) { return fminf(fmaxf(c, 0.0f), 1.0f); }

__device__ inline uint8_t to_rgb8(float c) {
    float v = roundf(255.0f * c); // cudamaths/image.cu:31
    v = fminf(fmaxf(v, 0.0f), 255.0f);
    return uint8_t(v);
}

// combineSamples (kernels/ssaa.cuh:109-131): one thread per pixel, frames
// along grid.y.  Writes RGB8 directly when no FXAA follows.
__global__ void __launch_bounds__(256) k_ssaa(const PostFrameDev *pfs) {
#pragma clang fp contract(off)
    const PostFrameDev pf = pfs[blockIdx.y];
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t npix = int64_t(pf.W) * pf.H;
    if (p >= npix) return;
    const int s = pf.s;
    float out[3];
    if (s == 1) {
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) out[ch] = clamp01(pf.ray_color[3 * p + ch]);
    } else {
        float acc[3] = {0.f, 0.f, 0.f};
        const float *src = pf.ray_color + 3 * p * s;
        for (int k = 0; k < s; ++k) {
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
                const float c = src[3 * k + ch];
                acc[ch] += clamp01(c * c);
            }
        }
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) out[ch] = sqrtf(acc[ch] / float(s));
    }
    float *dst = pf.fxaa ? pf.img : pf.final_rgbf;
    if (dst) {
        dst[3 * p] = out[0];
        dst[3 * p + 1] = out[1];
        dst[3 * p + 2] = out[2];
    }
    if (!pf.fxaa && pf.rgb8) {
        pf.rgb8[3 * p] = to_rgb8(out[0]);
        pf.rgb8[3 * p + 1] = to_rgb8(out[1]);
        pf.rgb8[3 * p + 2] = to_rgb8(out[2]);
    }
}

// luma (kernels/ssaa.cuh:141-143): double literals, float result.
__device__ inline float luma(float r, float g, float b) {
#pragma clang fp contract(off)
    const double k = double(1.0f / 255.0f);
    return float(double(r) * 0.299 * k + double(g) * 0.587 * k + double(b) * 0.114 * k);
}

__device__ inline float3 load3(const float *im, int W, int x, int y) {
    const float *p = im + 3 * (int64_t(y) * W + x);
    return make_float3(p[0], p[1], p[2]);
}

// Clamp-to-edge bilinear tap at pixel-space (X, Y).
__device__ inline float3 fxaa_tap(const float *im, int W, int H, float X, float Y) {
#pragma clang fp contract(off)
    const float fx0 = floorf(X), fy0 = floorf(Y);
    const float fx = X - fx0, fy = Y - fy0;
    const int x0 = int(fx0), y0 = int(fy0);
    const int xa = min(max(x0, 0), W - 1), xb = min(max(x0 + 1, 0), W - 1);
    const int ya = min(max(y0, 0), H - 1), yb = min(max(y0 + 1, 0), H - 1);
    const float3 c00 = load3(im, W, xa, ya), c10 = load3(im, W, xb, ya), c01 = load3(im, W, xa, yb),
                 c11 = load3(im, W, xb, yb);
    const float w00 = (1.0f - fx) * (1.0f - fy), w10 = fx * (1.0f - fy), w01 = (1.0f - fx) * fy, w11 = fx * fy;
    return make_float3(c00.x * w00 + c10.x * w10 + c01.x * w01 + c11.x * w11,
                       c00.y * w00 + c10.y * w10 + c01.y * w01 + c11.y * w11,
                       c00.z * w00 + c10.z * w10 + c01.z * w01 + c11.z * w11);
}

// fxaa (kernels/ssaa.cuh:150-182) on every pixel with clamp-to-edge taps,
// fused with the RGB8 quantisation.  16x16 tiles; neighbour taps hit L1/L2.
__global__ void __launch_bounds__(256) k_fxaa(const PostFrameDev *pfs) {
#pragma clang fp contract(off)
    const PostFrameDev pf = pfs[blockIdx.z];
    const int x = int(blockIdx.x) * 16 + int(threadIdx.x & 15);
    const int y = int(blockIdx.y) * 16 + int(threadIdx.x >> 4);
    const int W = pf.W, H = pf.H;
    if (!pf.fxaa || x >= W || y >= H) return;
    const float *im = pf.img;
    auto L = [&](int xx, int yy) {
        xx = min(max(xx, 0), W - 1);
        yy = min(max(yy, 0), H - 1);
        const float3 c = load3(im, W, xx, yy);
        return luma(c.x, c.y, c.z);
    };
    const float lNW = L(x - 1, y - 1), lNE = L(x + 1, y - 1), lSW = L(x - 1, y + 1), lSE = L(x + 1, y + 1);
    const float lM = L(x, y);
    const float lMin = fminf(lM, fminf(fminf(lNW, lNE), fminf(lSW, lSE)));
    const float lMax = fmaxf(lM, fmaxf(fmaxf(lNW, lNE), fmaxf(lSW, lSE)));
    float dx = -((lNW + lNE) - (lSW + lSE));
    float dy = ((lNW + lSW) - (lNE + lSE));
    const float reduce = fmaxf((lNW + lNE + lSW + lSE) * (0.25f * (1.0f / 8.0f)), 1.0f / 128.0f);
    const float rcp = 1.0f / (fminf(fabsf(dx), fabsf(dy)) + reduce);
    dx = fminf(fmaxf(dx * rcp, -8.0f), 8.0f);
    dy = fminf(fmaxf(dy * rcp, -8.0f), 8.0f);
    const float k1 = 1.0f / 3.0f - 0.5f, k2 = 2.0f / 3.0f - 0.5f;
    const float fxp = float(x), fyp = float(y);
    const float3 t1 = fxaa_tap(im, W, H, fxp + dx * k1, fyp + dy * k1);
    const float3 t2 = fxaa_tap(im, W, H, fxp + dx * k2, fyp + dy * k2);
    const float3 t3 = fxaa_tap(im, W, H, fxp + dx * -0.5f, fyp + dy * -0.5f);
    const float3 t4 = fxaa_tap(im, W, H, fxp + dx * 0.5f, fyp + dy * 0.5f);
    const float3 A = make_float3(t1.x * 0.5f + t2.x * 0.5f, t1.y * 0.5f + t2.y * 0.5f, t1.z * 0.5f + t2.z * 0.5f);
    const float3 B = make_float3(0.5f * A.x + (t3.x * 0.25f + t4.x * 0.25f), 0.5f * A.y + (t3.y * 0.25f + t4.y * 0.25f),
                                 0.5f * A.z + (t3.z * 0.25f + t4.z * 0.25f));
    const float lB = luma(B.x, B.y, B.z);
    const float3 o = (lB < lMin || lB > lMax) ? A : B;
    const int64_t p = int64_t(y) * W + x;
    if (pf.final_rgbf) {
        pf.final_rgbf[3 * p] = o.x;
        pf.final_rgbf[3 * p + 1] = o.y;
        pf.final_rgbf[3 * p + 2] = o.z;
    }
    if (pf.rgb8) {
        pf.rgb8[3 * p] = to_rgb8(o.x);
        pf.rgb8[3 * p + 1] = to_rgb8(o.y);
        pf.rgb8[3 * p + 2] = to_rgb8(o.z);
    }
}

// ----------------------------------------------------------------- adjoint (Kerr)
// Pass 1 renders with step counts and colours, pass 2 re-runs the
// integration and stores a checkpoint every K accepted steps, pass 3 walks
// each ray's segments backwards: replay K steps from a checkpoint, then the
// reverse sweep kerr_step_vjp over them.  All three passes call one
// out-of-line attempt function, so their accepted steps agree bit for bit.
// (lane data passed by value, so the callers' lane state stays in registers)
template <bool MIXED> __device__ __noinline__ KerrStepOut kerr_attempt_shared(const JobDev &job, const Lane<Kerr> L) {
    KerrStepOut o;
    o.L = L;
    o.cls = lane_attempt<Kerr, PRR_RKF45, MIXED>(job, o.L);
    return o;
}

// Weights dL/dc of one ray's colour from dL/dI of its pixel through the SSAA
// combine (kernels/ssaa.cuh:116-130): s = 1, I = clamp(c); s > 1,
// I = sqrt(mean(clamp(c^2))).
__device__ inline void color_weights(const AdjDev &ad, const FrameDev &fr, int64_t local, const float *c,
                                     double w[3]) {
    const int64_t pix = local / fr.s;
    for (int ch = 0; ch < 3; ++ch) {
        const double g = double(ad.dLdI[3 * pix + ch]);
        const double cv = double(c[ch]);
        if (fr.s == 1) {
            w[ch] = (cv > 0.0 && cv < 1.0) ? g : 0.0;
        } else {
            const double I = double(ad.img[3 * pix + ch]);
            w[ch] = (I > 0.0 && cv * cv < 1.0) ? g * cv / (fr.s * I) : 0.0;
        }
    }
}

__device__ inline void ckpt_store(double *c, const Lane<Kerr> &L) {
    for (int i = 0; i < 6; ++i) c[i] = L.y[i];
    c[6] = L.h;
    c[7] = double(L.attempts);
}

__device__ inline void ckpt_load(const JobDev &job, const double *c, Lane<Kerr> &L, int32_t steps) {
    for (int i = 0; i < 6; ++i) L.y[i] = c[i];
    L.h = c[6];
    L.attempts = int32_t(c[7]);
    L.lam = L.y[5];
    L.steps = steps;
    L.hit = -1;
    if (job.frames[L.frame].n_spheres > 0) cart_of_f32<Kerr>(L.y, L.P0);
}

template <bool MIXED> __global__ void __launch_bounds__(256) k_adj_forward(JobDev job, AdjDev ad) {
    const int lane = int(threadIdx.x & 63);
    Lane<Kerr> L;
    int32_t N = 0;
    bool active = false, exhausted = false;
    while (true) {
        const unsigned long long idle = __ballot(!active);
        const int nidle = __popcll(idle);
        if (!exhausted && nidle >= job.refill) {
            const int leader = __ffsll((long long)idle) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(&job.ctr->queue, (unsigned long long)nidle);
            base = shfl_u64(base, leader);
            if (base + (unsigned long long)nidle >= (unsigned long long)job.total) exhausted = true;
            if (!active) {
                const unsigned rank =
                    __builtin_amdgcn_mbcnt_hi(unsigned(idle >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(idle), 0u));
                const int64_t id = int64_t(base) + rank;
                if (id < job.total) {
                    L = lane_start<Kerr>(job, id);
                    N = ad.steps[id];
                    ckpt_store(ad.ckpt + 8 * ad.ckpt_off[id], L);
                    active = N > 0 && finite_state<Kerr>(L.y);
                }
            }
        }
        if (__ballot(active) == 0ull) break;
        if (active) {
            const int32_t before = L.steps;
            const KerrStepOut r = kerr_attempt_shared<MIXED>(job, L);
            L = r.L;
            const int cls = r.cls;
            if (cls < 0 && L.steps != before && L.steps < N && L.steps % ad.K == 0)
                ckpt_store(ad.ckpt + 8 * (ad.ckpt_off[L.rid] + L.steps / ad.K), L);
            if (cls >= 0 || L.steps >= N) active = false;
        }
    }
}

// Terminal seed: lambda at y_{N-1} (state and constants) and dL/dphi_s from
// the located event of the last step (see adjoint.hpp).
__device__ __noinline__ void adj_terminal(const JobDev &job, const AdjDev &ad, const Lane<Kerr> L,
                                          const EventDev &ev, const double w[3], double lam[9], double &gphs) {
    const FrameDev &fr = job.frames[L.frame];
    const double tau = ev.theta * ev.h;
    double ys[6];
    rk_substep<Kerr>(ev.y_prev, tau, L.cs, ys);
    using D = Dual<10>;
    D y[5];
    for (int i = 0; i < 5; ++i) y[i] = D::var(ys[i], i);
    const D E = D::var(L.cs.E, 5), Lz = D::var(L.cs.L, 6), Q = D::var(L.cs.Q, 7), a = D::var(L.cs.a, 8);
    D col[3], G;
    if (ev.cls == PRR_ESCAPE) {
        const V3<D> n = kerr_escape_dir_t<D>(y, E, Lz, Q, a, L.cs.M);
        sky_color_t<D>(job.tex[0], n, col);
        G = D(1.0 / job.tr.r_escape) - y[0];
    } else {
        const int k = ev.hit;
        V3<D> c = v3t<D>(D(fr.sph[k][0]), D(fr.sph[k][1]), D(fr.sph[k][2]));
        if (k == 0) { // the orbit's sphere: centre (R_orbit cos phi_s, R_orbit sin phi_s, 0)
            const double phs = atan2(fr.sph[0][1], fr.sph[0][0]);
            D sp, cp;
            sincos(D::var(phs, 9), &sp, &cp);
            c = v3t<D>(cp * ad.sphere_r, sp * ad.sphere_r, D(0.0));
        }
        const V3<D> m = vsub(kerr_cart_t<D>(y), c);
        const int slot = fr.sph_tex[k];
        sphere_color_t<D>((slot > 0 && slot < PRR_MAX_TEXTURES) ? &job.tex[slot] : nullptr, vnormalize(m), col);
        G = vdot(m, m) - fr.sph[k][3] * fr.sph[k][3];
    }
    double ls[9];
    for (int i = 0; i < 9; ++i) ls[i] = w[0] * col[0].d[i] + w[1] * col[1].d[i] + w[2] * col[2].d[i];
    gphs = w[0] * col[0].d[9] + w[1] * col[1].d[9] + w[2] * col[2].d[9];
    double f[6];
    Kerr::rhs<double>(ys, Kerr::coef<double>(L.cs), f);
    double num = 0.0, den = 0.0;
    for (int i = 0; i < 5; ++i) {
        num += ls[i] * f[i];
        den += G.d[i] * f[i];
    }
    double kappa = num / den;
    if (!isfinite(kappa)) kappa = 0.0;
    for (int i = 0; i < 9; ++i) lam[i] = ls[i] - kappa * G.d[i];
    gphs -= kappa * G.d[9];
    kerr_step_vjp(ev.y_prev, tau, L.cs, lam);
}

// Ray set-up Jacobian d(y_0, E, L, Q)/d(a, r_cam, incl) contracted with lambda_0.
__device__ __noinline__ void adj_init_grad(const JobDev &job, const AdjDev &ad, const Lane<Kerr> L,
                                           const double lam[9], double g[4]) {
    const FrameDev &fr = job.frames[L.frame];
    double u, v;
    ray_uv(fr, L.rid - fr.ray_begin, u, v);
    const double r_cam = sqrt(fr.pos[0] * fr.pos[0] + fr.pos[1] * fr.pos[1] + fr.pos[2] * fr.pos[2]);
    const double incl = acos(fr.pos[2] / r_cam);
    using D = Dual<3>;
    V3<D> pos, d;
    orbit_camera_ray<D>(D::var(r_cam, 1), D::var(incl, 2), ad.cam_phi, fr.tan_half, fr.aspect, u, v, pos, d);
    D z[8];
    kerr_init_t<D>(pos, d, fr.M, D::var(fr.a, 0), z);
    for (int k = 0; k < 3; ++k) {
        double acc = k == 0 ? lam[8] : 0.0; // a is itself one of the constants
        for (int i = 0; i < 8; ++i) acc += lam[i] * z[i].d[k];
        g[k] = acc;
    }
}

template <bool MIXED> __global__ void __launch_bounds__(256) k_adj_backward(JobDev job, AdjDev ad) {
    const int lane = int(threadIdx.x & 63);
    const int64_t gl = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; // this lane's segment buffer
    Lane<Kerr> L;
    bool active = false, exhausted = false;
    while (true) {
        const unsigned long long idle = __ballot(!active);
        const int nidle = __popcll(idle);
        if (!exhausted && nidle >= 1) {
            const int leader = __ffsll((long long)idle) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(&job.ctr->queue, (unsigned long long)nidle);
            base = shfl_u64(base, leader);
            if (base + (unsigned long long)nidle >= (unsigned long long)job.total) exhausted = true;
            if (!active) {
                const unsigned rank =
                    __builtin_amdgcn_mbcnt_hi(unsigned(idle >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(idle), 0u));
                const int64_t id = int64_t(base) + rank;
                if (id < job.total) {
                    L.rid = id;
                    active = true;
                }
            }
        }
        if (__ballot(active) == 0ull) break;
        if (!active) continue;
        // one whole ray per lane: replay + reverse every segment
        const int64_t id = L.rid;
        const int cls = ad.cls[id];
        const int32_t N = ad.steps[id];
        double g[4] = {0.0, 0.0, 0.0, 0.0};
        double w[3] = {0.0, 0.0, 0.0};
        if ((cls == PRR_ESCAPE || cls == PRR_SPHERE) && N > 0) {
            L = lane_start<Kerr>(job, id); // constants E, L, Q and the frame
            const FrameDev &fr = job.frames[L.frame];
            color_weights(ad, fr, id - fr.ray_begin, ad.ray_color + 3 * id, w);
        }
        if (w[0] != 0.0 || w[1] != 0.0 || w[2] != 0.0) {
            double lam[9];
            double gphs = 0.0;
            const int64_t c0 = ad.ckpt_off[id];
            bool ok = true;
            for (int32_t j = (N - 1) / ad.K; j >= 0 && ok; --j) {
                ckpt_load(job, ad.ckpt + 8 * (c0 + j), L, j * ad.K);
                const int32_t end = min((j + 1) * ad.K, N);
                // replay: record (y_n, h_n) of every accepted step of the segment
                int rc = -1;
                while (L.steps < end) {
                    double ys[5];
                    for (int i = 0; i < 5; ++i) ys[i] = L.y[i];
                    const double hb = L.h;
                    const int32_t before = L.steps;
                    const KerrStepOut r = kerr_attempt_shared<MIXED>(job, L);
                    L = r.L;
                    rc = r.cls;
                    if (L.steps != before) {
                        double *slot = ad.seg + (int64_t(before - j * ad.K) * 6) * ad.seg_lanes + gl;
                        for (int i = 0; i < 5; ++i) slot[i * ad.seg_lanes] = ys[i];
                        bool at_min;
                        slot[5 * ad.seg_lanes] = rkf_step_size<Kerr>(job.tr, ys, L.cs, hb, at_min);
                    }
                    if (rc >= 0) break;
                }
                // the ray must end exactly at step N with its class, and nowhere earlier
                const bool last = end == N;
                if (L.steps != end || (last ? rc != cls : rc >= 0)) {
                    atomicAdd(ad.replay_mismatch, 1ull);
                    ok = false;
                    break;
                }
                if (!last) { // the replay must land on the next checkpoint
                    const double *nx = ad.ckpt + 8 * (c0 + j + 1);
                    bool same = true;
                    for (int i = 0; i < 6; ++i) same = same && L.y[i] == nx[i];
                    if (!same) {
                        atomicAdd(ad.replay_mismatch, 1ull);
                        ok = false;
                        break;
                    }
                } else {
                    adj_terminal(job, ad, L, job.out.event[id], w, lam, gphs);
                }
                // reverse sweep (the terminal step N-1 is the located sub-step above)
                for (int32_t n = min(end, N - 1) - 1; n >= j * ad.K; --n) {
                    const double *slot = ad.seg + (int64_t(n - j * ad.K) * 6) * ad.seg_lanes + gl;
                    double yn[6];
                    for (int i = 0; i < 5; ++i) yn[i] = slot[i * ad.seg_lanes];
                    yn[5] = 0.0;
                    kerr_step_vjp(yn, slot[5 * ad.seg_lanes], L.cs, lam);
                }
            }
            if (ok) {
                adj_init_grad(job, ad, L, lam, g);
                g[3] = gphs;
                bool fin = true;
                for (int k = 0; k < 4; ++k) fin = fin && isfinite(g[k]);
                if (!fin) g[0] = g[1] = g[2] = g[3] = 0.0;
            }
        }
        double *out = ad.grad_ray + 4 * id;
        for (int k = 0; k < 4; ++k) out[k] = g[k];
        active = false;
    }
}

// Deterministic sum of the per-ray gradients: one workgroup, fixed order.
__global__ void __launch_bounds__(1024) k_grad_reduce(const double *grad_ray, int64_t n, double *out) {
    __shared__ double part[4][1024];
    const int t = int(threadIdx.x);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t i = t; i < n; i += 1024)
        for (int k = 0; k < 4; ++k) acc[k] += grad_ray[4 * i + k];
    for (int k = 0; k < 4; ++k) part[k][t] = acc[k];
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (t < s)
            for (int k = 0; k < 4; ++k) part[k][t] += part[k][t + s];
        __syncthreads();
    }
    if (t == 0)
        for (int k = 0; k < 4; ++k) out[k] = part[k][0];
}

hipError_t launch_adjoint(int pass, int precision, const JobDev &job, const AdjDev &ad, int grid, int block,
                          hipStream_t st) {
    const bool mixed = precision == PRR_FP32_BULK;
    if (pass == 1) {
        if (mixed) hipLaunchKernelGGL((k_integrate<Kerr, PRR_RKF45, true, true>), dim3(grid), dim3(block), 0, st, job);
        else hipLaunchKernelGGL((k_integrate<Kerr, PRR_RKF45, false, true>), dim3(grid), dim3(block), 0, st, job);
    } else if (pass == 2) {
        if (mixed) hipLaunchKernelGGL((k_adj_forward<true>), dim3(grid), dim3(block), 0, st, job, ad);
        else hipLaunchKernelGGL((k_adj_forward<false>), dim3(grid), dim3(block), 0, st, job, ad);
    } else {
        if (mixed) hipLaunchKernelGGL((k_adj_backward<true>), dim3(grid), dim3(block), 0, st, job, ad);
        else hipLaunchKernelGGL((k_adj_backward<false>), dim3(grid), dim3(block), 0, st, job, ad);
    }
    return hipGetLastError();
}

hipError_t launch_grad_reduce(const double *grad_ray, int64_t n, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_grad_reduce, dim3(1), dim3(1024), 0, st, grad_ray, n, out);
    return hipGetLastError();
}

// ----------------------------------------------------------------- dispatch
template <class Mt>
static hipError_t launch_metric(int solver, int precision, const JobDev &job, dim3 grid, dim3 block,
                                hipStream_t st) {
    const bool mixed = precision == PRR_FP32_BULK;
    if (solver == PRR_EULER) {
        if (mixed) hipLaunchKernelGGL((k_integrate<Mt, PRR_EULER, true, false>), grid, block, 0, st, job);
        else hipLaunchKernelGGL((k_integrate<Mt, PRR_EULER, false, false>), grid, block, 0, st, job);
    } else {
        if (mixed) hipLaunchKernelGGL((k_integrate<Mt, PRR_RKF45, true, false>), grid, block, 0, st, job);
        else hipLaunchKernelGGL((k_integrate<Mt, PRR_RKF45, false, false>), grid, block, 0, st, job);
    }
    return hipGetLastError();
}

hipError_t launch_integrate(int metric, const JobDev &job, int grid, int block, hipStream_t st) {
    const dim3 g(grid), b(block);
    switch (metric) {
    case PRR_FLAT: return launch_metric<Flat>(job.tr.solver, job.tr.precision, job, g, b, st);
    case PRR_SCHW_KS_CART: return launch_metric<SchwKS>(job.tr.solver, job.tr.precision, job, g, b, st);
    case PRR_SCHW_SPH: return launch_metric<SchwSph>(job.tr.solver, job.tr.precision, job, g, b, st);
    case PRR_KERR_BL: return launch_metric<Kerr>(job.tr.solver, job.tr.precision, job, g, b, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_reset(CountersDev *c, hipStream_t st) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, st, c);
    return hipGetLastError();
}

hipError_t launch_ssaa(const PostFrameDev *pfs, int nframes, int64_t max_pixels, hipStream_t st) {
    const dim3 g(unsigned((max_pixels + 255) / 256), unsigned(nframes)), b(256);
    hipLaunchKernelGGL(k_ssaa, g, b, 0, st, pfs);
    return hipGetLastError();
}

hipError_t launch_fxaa(const PostFrameDev *pfs, int nframes, int max_w, int max_h, hipStream_t st) {
    const dim3 g(unsigned((max_w + 15) / 16), unsigned((max_h + 15) / 16), unsigned(nframes)), b(256);
    hipLaunchKernelGGL(k_fxaa, g, b, 0, st, pfs);
    return hipGetLastError();
}

} // namespace prr
