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
#include "geodesic.hpp"
#include "kernels.hpp"

// Register budget of the integrator: waves per SIMD the compiler must fit
// (build-time knob, -DPRR_WAVES_PER_EU=n; unset = the compiler's choice).
#ifdef PRR_WAVES_PER_EU
#define PRR_OCCUPANCY __attribute__((amdgpu_waves_per_eu(PRR_WAVES_PER_EU, 8)))
#else
#define PRR_OCCUPANCY
#endif

namespace prr {

// Fehlberg 4(5) tableau (kernels/solver.cuh:125-131).
__device__ constexpr double kA[6][5] = {
    {0, 0, 0, 0, 0},
    {1.0 / 4.0, 0, 0, 0, 0},
    {3.0 / 32.0, 9.0 / 32.0, 0, 0, 0},
    {1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0, 0, 0},
    {439.0 / 216.0, -8.0, 3680.0 / 513.0, -845.0 / 4104.0, 0},
    {-8.0 / 27.0, 2.0, -3544.0 / 2565.0, 1859.0 / 4104.0, -11.0 / 40.0}};
__device__ constexpr double kB4[6] = {25.0 / 216.0, 0.0, 1408.0 / 2565.0, 2197.0 / 4104.0, -1.0 / 5.0, 0.0};
// b5 - b4: the embedded error estimate.
__device__ constexpr double kE[6] = {16.0 / 135.0 - 25.0 / 216.0,       0.0,
                                     6656.0 / 12825.0 - 1408.0 / 2565.0, 28561.0 / 56430.0 - 2197.0 / 4104.0,
                                     -9.0 / 50.0 + 1.0 / 5.0,            2.0 / 55.0};

// Six RK stages in precision T from the FP64 state y; returns the 4th-order
// increment d4 and the error vector e in FP64.
template <class Mt, class T>
__device__ inline void rkf_stages(const double *y, double h, const Consts<double> &cd, double *d4, double *e) {
    constexpr int N = Mt::N;
    const Consts<T> c{T(cd.M), T(cd.a), T(cd.E), T(cd.L)};
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
            for (int j = 0; j < s; ++j) acc += T(kA[s][j]) * k[j][i];
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
            if (kB4[s] != 0.0) a4 += kB4[s] * double(k[s][i]); // zero terms folded at compile time
            if (kE[s] != 0.0) ae += kE[s] * double(k[s][i]);
        }
        d4[i] = a4;
        e[i] = ae;
    }
}

template <class Mt, class T>
__device__ inline void euler_rhs(const double *y, const Consts<double> &cd, double *f) {
    constexpr int N = Mt::N;
    const Consts<T> c{T(cd.M), T(cd.a), T(cd.E), T(cd.L)};
    T yT[N], fT[N];
#pragma unroll
    for (int i = 0; i < N; ++i) yT[i] = T(y[i]);
    Mt::template rhs<T>(yT, c, fT);
#pragma unroll
    for (int i = 0; i < N; ++i) f[i] = double(fT[i]);
}

__device__ inline int find_frame(const JobDev &job, int64_t id) {
    int lo = 0, hi = job.nframes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (job.frames[mid].ray_begin <= id) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Lane state of one ray.
template <class Mt> struct Lane {
    double y[Mt::N];
    Consts<double> cs;
    double h, lam;
    int32_t steps, attempts, att64;
    int32_t frame;
    int32_t hit; // sphere index of a SPHERE termination
    int64_t rid;
    D3 P0; // Cartesian position at the start of the step (sphere chord); the hit normal after a hit
};

// lane_init and lane_finish run once per ray; they are kept out of line so
// that their transcendental constants and temporaries do not occupy
// registers across the integration loop.
template <class Mt>
__device__ __noinline__ void lane_init(const JobDev &job, Lane<Mt> &L, int64_t id) {
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
    L.h = job.tr.h;
    L.lam = 0.0;
    L.steps = L.attempts = L.att64 = 0;
    if (fr.n_spheres > 0) L.P0 = cart_of<Mt>(L.y);
}

template <class Mt>
__device__ inline bool finite_state(const double *y) {
    bool ok = true;
#pragma unroll
    for (int i = 0; i < Mt::N; ++i) ok = ok && isfinite(y[i]);
    return ok;
}

// Writes the outputs of a finished ray (rayToRGB, metrics/schwarzschild.cuh:151-183).
template <class Mt>
__device__ __noinline__ void lane_finish(const JobDev &job, const Lane<Mt> &L, int cls) {
    const int64_t id = L.rid;
    D3 endpoint;
    float3 color;
    if (cls == PRR_SPHERE) {
        endpoint = L.P0;
        const int slot = job.frames[L.frame].sph_tex[L.hit];
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

// Per-step termination tests after an accepted step from yo (User_UPDATE /
// User_ENDLOOP, metrics/schwarzschild.cuh:55-79): sphere chord, spherical
// wraps, horizon, escape.  Returns the class or -1 to continue.
template <class Mt>
__device__ inline int after_step(const JobDev &job, Lane<Mt> &L) {
    const FrameDev &fr = job.frames[L.frame];
    if (fr.n_spheres > 0) {
        const D3 P1 = cart_of<Mt>(L.y);
        double t;
        const int k = chord_hit(fr, L.P0, P1, t);
        if (k >= 0) {
            const D3 H = L.P0 + t * (P1 - L.P0);
            L.P0 = (1.0 / fr.sph[k][3]) * (H - d3(fr.sph[k][0], fr.sph[k][1], fr.sph[k][2]));
            L.hit = k;
            return PRR_SPHERE;
        }
        L.P0 = P1;
    }
    if (Mt::kind == PRR_SCHW_SPH) {
        sph_post_step(L.y);
        if (fr.n_spheres > 0) L.P0 = cart_of<Mt>(L.y);
    }
    const double r = radius_of<Mt>(L.y);
    if (r <= fr.r_hor) return PRR_HORIZON;
    if (r >= job.tr.r_escape) return PRR_ESCAPE;
    return -1;
}

// One Euler step or one RKF45 attempt.  Returns the class when the ray ends.
template <class Mt, int SOLVER, bool MIXED>
__device__ inline int lane_advance(const JobDev &job, Lane<Mt> &L) {
    constexpr int N = Mt::N;
    const TracerDev &tr = job.tr;
    if (L.steps >= tr.max_steps || L.lam >= tr.lambda_max) return PRR_STEP_CAP;
    const bool fp64 = !MIXED || radius_of<Mt>(L.y) < tr.r_fp64;
    if (SOLVER == PRR_EULER) {
        double f[N];
        if (fp64) euler_rhs<Mt, double>(L.y, L.cs, f);
        else euler_rhs<Mt, float>(L.y, L.cs, f);
        bool ok = true;
#pragma unroll
        for (int i = 0; i < N; ++i) ok = ok && isfinite(f[i]);
        if (!ok) return PRR_NAN;
        const double h = L.h;
#pragma unroll
        for (int i = 0; i < N; ++i) L.y[i] += h * f[i];
        L.steps++;
        L.attempts++;
        L.att64 += fp64 ? 1 : 0;
        L.lam += h;
        if (!finite_state<Mt>(L.y)) return PRR_NAN;
        return after_step<Mt>(job, L);
    }
    if (L.attempts >= tr.max_attempts) return PRR_STEP_CAP;
    L.attempts++;
    L.att64 += fp64 ? 1 : 0;
    double d4[N], e[N];
    if (fp64) rkf_stages<Mt, double>(L.y, L.h, L.cs, d4, e);
    else rkf_stages<Mt, float>(L.y, L.h, L.cs, d4, e);
    double errn = 0.0;
    double y4[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        y4[i] = L.y[i] + d4[i];
        const double sc = tr.atol + tr.rtol * fmax(fabs(L.y[i]), fabs(y4[i]));
        errn = fmax(errn, fabs(e[i]) / sc);
    }
    // _h = clamp(0.9 h (eps/maxerr)^0.2, hmin, hmax) (kernels/solver.cuh:175-176)
    double hn;
    if (MIXED) hn = 0.9 * L.h * double(powf(float(errn), -0.2f));
    else hn = 0.9 * L.h * pow(errn, -0.2);
    hn = fmin(fmax(hn, tr.hmin), tr.hmax);
    if (errn > 1.0 && L.h != tr.hmin) { // reject (:185-188)
        L.h = hn;
        return -1;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) L.y[i] = y4[i];
    if (!finite_state<Mt>(L.y)) return PRR_NAN;
    L.steps++;
    L.lam += L.h;
    L.h = hn;
    return after_step<Mt>(job, L);
}

__device__ inline unsigned long long shfl_u64(unsigned long long v, int src) {
    const int lo = __shfl(int(v & 0xffffffffull), src);
    const int hi = __shfl(int(v >> 32), src);
    return (unsigned long long)(unsigned(lo)) | ((unsigned long long)(unsigned(hi)) << 32);
}

template <class Mt, int SOLVER, bool MIXED>
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
                    lane_init<Mt>(job, L, id);
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
            const int cls = lane_advance<Mt, SOLVER, MIXED>(job, L);
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
// Pass 1 is k_integrate itself (step counts and colours).  Pass 2 re-runs
// the same integration, stores a checkpoint every K accepted steps and the
// terminal seed dL/dz_N.  Pass 3 walks each ray's segments backwards:
// replay K steps from a checkpoint (same lane_advance, so the same accepted
// steps bit for bit), then the reverse sweep kerr_step_vjp over them.

// Weights dL/dc of one ray's colour from dL/dI of its pixel through the SSAA
// combine (kernels/ssaa.cuh:116-130): s = 1, I = clamp(c); s > 1,
// I = sqrt(mean(clamp(c^2))).
__device__ inline void color_weights(const JobDev &job, const AdjDev &ad, const FrameDev &fr, int64_t local,
                                     const float3 c, double w[3]) {
    const int64_t pix = local / fr.s;
    const float cc[3] = {c.x, c.y, c.z};
    for (int ch = 0; ch < 3; ++ch) {
        const double g = double(ad.dLdI[3 * pix + ch]);
        const double cv = double(cc[ch]);
        if (fr.s == 1) {
            w[ch] = (cv > 0.0 && cv < 1.0) ? g : 0.0;
        } else {
            const double I = double(ad.img[3 * pix + ch]);
            w[ch] = (I > 0.0 && cv * cv < 1.0) ? g * cv / (fr.s * I) : 0.0;
        }
    }
}

// Terminal seed of a finished ray: dL/dz_N (8), dL/dz_{N-1} (r, mu, phi) and
// dL/dphi_s, by forward-mode duals through the colour of its class.
__device__ __noinline__ void adj_seed(const JobDev &job, const AdjDev &ad, const Lane<Kerr> &L, int cls,
                                      const double yp[3]) {
    double *s = ad.seed + 12 * L.rid;
    for (int i = 0; i < 12; ++i) s[i] = 0.0;
    if (cls != PRR_ESCAPE && cls != PRR_SPHERE) return;
    const FrameDev &fr = job.frames[L.frame];
    const float *rc = ad.ray_color + 3 * L.rid;
    double w[3];
    color_weights(job, ad, fr, L.rid - fr.ray_begin, make_float3(rc[0], rc[1], rc[2]), w);
    if (w[0] == 0.0 && w[1] == 0.0 && w[2] == 0.0) return;
    using D = Dual<8>;
    D col[3];
    if (cls == PRR_ESCAPE) {
        D y[5];
        for (int i = 0; i < 5; ++i) y[i] = D::var(L.y[i], i);
        const V3<D> n = kerr_escape_dir_t<D>(y, D::var(L.cs.E, 5), D::var(L.cs.L, 6), D::var(L.cs.a, 7), L.cs.M);
        sky_color_t<D>(job.tex[0], n, col);
        for (int i = 0; i < 8; ++i) s[i] = w[0] * col[0].d[i] + w[1] * col[1].d[i] + w[2] * col[2].d[i];
        return;
    }
    // SPHERE: directions 0-2 = (r, mu, phi) at step N-1, 3-5 = at step N, 6 = phi_s
    D y0[5], y1[5];
    for (int i = 0; i < 5; ++i) {
        y0[i] = D(0.0);
        y1[i] = D(L.y[i]);
    }
    for (int i = 0; i < 3; ++i) {
        y0[i] = D::var(yp[i], i);
        y1[i] = D::var(L.y[i], 3 + i);
    }
    const int k = L.hit;
    V3<D> c = v3t<D>(D(fr.sph[k][0]), D(fr.sph[k][1]), D(fr.sph[k][2]));
    if (k == 0) { // the orbit's sphere: centre (R_orbit cos phi_s, R_orbit sin phi_s, 0)
        const double phs = atan2(fr.sph[0][1], fr.sph[0][0]);
        D sp, cp;
        sincos(D::var(phs, 6), &sp, &cp);
        c = v3t<D>(cp * ad.sphere_r, sp * ad.sphere_r, D(0.0));
    }
    const int slot = fr.sph_tex[k];
    sphere_hit_color_t<D>((slot > 0 && slot < PRR_MAX_TEXTURES) ? &job.tex[slot] : nullptr, kerr_cart_t<D>(y0),
                          kerr_cart_t<D>(y1), c, fr.sph[k][3], col);
    for (int i = 0; i < 3; ++i) {
        s[8 + i] = w[0] * col[0].d[i] + w[1] * col[1].d[i] + w[2] * col[2].d[i];
        s[i] = w[0] * col[0].d[3 + i] + w[1] * col[1].d[3 + i] + w[2] * col[2].d[3 + i];
    }
    s[11] = w[0] * col[0].d[6] + w[1] * col[1].d[6] + w[2] * col[2].d[6];
}

__device__ inline void ckpt_store(double *c, const Lane<Kerr> &L) {
    for (int i = 0; i < 5; ++i) c[i] = L.y[i];
    c[5] = L.h;
    c[6] = L.lam;
    c[7] = double(L.attempts);
}

__device__ inline void ckpt_load(const double *c, Lane<Kerr> &L, int32_t steps) {
    for (int i = 0; i < 5; ++i) L.y[i] = c[i];
    L.h = c[5];
    L.lam = c[6];
    L.attempts = int32_t(c[7]);
    L.steps = steps;
    L.P0 = cart_of<Kerr>(L.y);
}

template <bool MIXED>
__global__ void __launch_bounds__(256) k_adj_forward(JobDev job, AdjDev ad) {
    const int lane = int(threadIdx.x & 63);
    Lane<Kerr> L;
    double yp[3] = {0.0, 0.0, 0.0};
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
                    lane_init<Kerr>(job, L, id);
                    active = true;
                    ckpt_store(ad.ckpt + 8 * ad.ckpt_off[id], L);
                    if (!finite_state<Kerr>(L.y)) {
                        adj_seed(job, ad, L, PRR_NAN, yp);
                        active = false;
                    }
                }
            }
        }
        if (__ballot(active) == 0ull) break;
        if (active) {
            const double y0 = L.y[0], y1 = L.y[1], y2 = L.y[2];
            const int32_t before = L.steps;
            const int cls = lane_advance<Kerr, PRR_RKF45, MIXED>(job, L);
            if (L.steps != before) {
                yp[0] = y0;
                yp[1] = y1;
                yp[2] = y2;
                if (L.steps % ad.K == 0 && cls < 0) ckpt_store(ad.ckpt + 8 * (ad.ckpt_off[L.rid] + L.steps / ad.K), L);
            }
            if (cls >= 0) {
                adj_seed(job, ad, L, cls, yp);
                active = false;
            }
        }
    }
}

// Ray set-up Jacobian dz_0/d(a, r_cam, incl) contracted with lambda_0.
__device__ __noinline__ void adj_init_grad(const JobDev &job, const AdjDev &ad, const Lane<Kerr> &L,
                                           const double lam[8], double g[4]) {
    const FrameDev &fr = job.frames[L.frame];
    double u, v;
    ray_uv(fr, L.rid - fr.ray_begin, u, v);
    const double r_cam = sqrt(fr.pos[0] * fr.pos[0] + fr.pos[1] * fr.pos[1] + fr.pos[2] * fr.pos[2]);
    const double incl = acos(fr.pos[2] / r_cam);
    using D = Dual<3>;
    V3<D> pos, d;
    orbit_camera_ray<D>(D::var(r_cam, 1), D::var(incl, 2), ad.cam_phi, fr.tan_half, fr.aspect, u, v, pos, d);
    D z[7];
    kerr_init_t<D>(pos, d, fr.M, D::var(fr.a, 0), z);
    for (int k = 0; k < 3; ++k) {
        double acc = k == 0 ? lam[7] : 0.0; // a is itself part of z
        for (int i = 0; i < 7; ++i) acc += lam[i] * z[i].d[k];
        g[k] = acc;
    }
}

template <bool MIXED>
__global__ void __launch_bounds__(256) k_adj_backward(JobDev job, AdjDev ad) {
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
        const double *sd = ad.seed + 12 * id;
        double g[4] = {0.0, 0.0, 0.0, sd[11]};
        bool any = false;
        for (int i = 0; i < 11; ++i) any = any || sd[i] != 0.0;
        const int32_t N = ad.steps[id];
        if (any && N > 0) {
            lane_init<Kerr>(job, L, id); // constants E, L and the frame
            double lam[8];
            for (int i = 0; i < 8; ++i) lam[i] = sd[i];
            const int64_t c0 = ad.ckpt_off[id];
            bool ok = true;
            for (int32_t j = (N - 1) / ad.K; j >= 0; --j) {
                ckpt_load(ad.ckpt + 8 * (c0 + j), L, j * ad.K);
                const int32_t end = min((j + 1) * ad.K, N);
                // replay: record (y_n, h_n) of every accepted step of the segment
                while (L.steps < end) {
                    double ys[5];
                    for (int i = 0; i < 5; ++i) ys[i] = L.y[i];
                    const double hs = L.h;
                    const int32_t before = L.steps;
                    const int cls = lane_advance<Kerr, PRR_RKF45, MIXED>(job, L);
                    if (L.steps != before) {
                        double *slot = ad.seg + (int64_t(before - j * ad.K) * 6) * ad.seg_lanes + gl;
                        for (int i = 0; i < 5; ++i) slot[i * ad.seg_lanes] = ys[i];
                        slot[5 * ad.seg_lanes] = hs;
                    }
                    // a ray only ends at step N; ending earlier means the replay diverged
                    if (cls >= 0 && L.steps < end) {
                        atomicAdd(ad.replay_mismatch, 1ull);
                        break;
                    }
                }
                if (L.steps < end) { // diverged: this ray contributes nothing
                    ok = false;
                    g[3] = 0.0;
                    break;
                }
                if (end < N && end % ad.K == 0) { // the replay must land on the next checkpoint
                    const double *nx = ad.ckpt + 8 * (c0 + j + 1);
                    bool same = L.steps == end;
                    for (int i = 0; i < 5; ++i) same = same && L.y[i] == nx[i];
                    if (!same) atomicAdd(ad.replay_mismatch, 1ull);
                }
                // reverse sweep
                for (int32_t n = end - 1; n >= j * ad.K; --n) {
                    const double *slot = ad.seg + (int64_t(n - j * ad.K) * 6) * ad.seg_lanes + gl;
                    double yn[5];
                    for (int i = 0; i < 5; ++i) yn[i] = slot[i * ad.seg_lanes];
                    kerr_step_vjp(yn, slot[5 * ad.seg_lanes], L.cs, lam);
                    if (n == N - 1) {
                        lam[0] += sd[8];
                        lam[1] += sd[9];
                        lam[2] += sd[10];
                    }
                }
            }
            if (ok) adj_init_grad(job, ad, L, lam, g);
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
    if (pass == 2) {
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
        if (mixed) hipLaunchKernelGGL((k_integrate<Mt, PRR_EULER, true>), grid, block, 0, st, job);
        else hipLaunchKernelGGL((k_integrate<Mt, PRR_EULER, false>), grid, block, 0, st, job);
    } else {
        if (mixed) hipLaunchKernelGGL((k_integrate<Mt, PRR_RKF45, true>), grid, block, 0, st, job);
        else hipLaunchKernelGGL((k_integrate<Mt, PRR_RKF45, false>), grid, block, 0, st, job);
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
