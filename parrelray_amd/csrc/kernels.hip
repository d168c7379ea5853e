// kernels.hip -- gfx950 kernels: the ray pipeline (set-up, persistent
// integrator, shading), the SSAA / FXAA / RGB8 passes and the Kerr adjoint.
//
// Replaces initRays + trace_Euler/trace_RKF + rayToRGB (kernels/solver.cuh:
// 26-220, metrics/schwarzschild.cuh:151-183, launched by prr.cpp:279-323 as
// one thread per ray in 64-row bands) with three kernels over a ray-state
// array in HBM:
//   * k_init: one thread per ray, camera ray and initial state;
//   * k_integrate: one persistent grid sized to the CUs; every lane holds one
//     ray's state in VGPRs and runs one Euler step / RKF45 attempt / event
//     sub-step per loop trip.  When at least `refill` lanes of a wave are
//     idle, the wave claims that many ray ids with a single atomic (ballot +
//     mbcnt prefix sum).  The loop body carries no calls and no once-per-ray
//     code, so the state stays in registers;
//   * k_shade: one thread per ray, endpoint and colour.
// The Kerr adjoint's checkpoint and replay passes run k_integrate itself
// (AdjRunDev), so that every pass takes the same accepted steps bit for bit.
#include <hip/hip_runtime.h>

#include "adjoint.hpp"
#include "kernels.hpp"
#include "lane.hpp"

// Register budget of the integrator: waves per SIMD the compiler must fit.
// PRR_WAVES_PER_EU (build-time knob, default 3: at most 168 VGPRs) for the
// renders of the 6-component Kerr state and of the flat and Kerr-Schild
// metrics with Euler, RK4 and RKF45, which fit it (escape location stays in
// the loop, sphere location is out of line) with a few spills in branches
// taken once per ray.  2 waves (at most 256 VGPRs) for the rest: the
// 8-component spherical charts, Dormand-Prince's seven stages and the
// adjoint's checkpointing and replay instantiations carry more live state
// and would spill inside the loop.
#ifndef PRR_WAVES_PER_EU
#define PRR_WAVES_PER_EU 3
#endif
template <class Mt, int SOLVER> struct WavesOf {
    static constexpr bool lean = !Mt::ckpt && !Mt::replay && SOLVER != PRR_DORMAND &&
                                 (Mt::kind == PRR_KERR_BL || Mt::kind == PRR_FLAT || Mt::kind == PRR_SCHW_KS_CART);
    static constexpr int value = lean ? PRR_WAVES_PER_EU : 2;
};
#define PRR_OCCUPANCY __attribute__((amdgpu_waves_per_eu(WavesOf<Mt, SOLVER>::value, 8)))

namespace prr {

__device__ inline unsigned long long shfl_u64(unsigned long long v, int src) {
    const int lo = __shfl(int(v & 0xffffffffull), src);
    const int hi = __shfl(int(v >> 32), src);
    return (unsigned long long)(unsigned(lo)) | ((unsigned long long)(unsigned(hi)) << 32);
}

template <class Mt> __global__ void __launch_bounds__(256) k_init(JobDev job) {
    const int64_t id = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id < job.total) init_ray<Mt>(job, id);
}

template <class Mt> __global__ void __launch_bounds__(256) k_shade(JobDev job) {
    const int64_t id = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id < job.total) shade_ray<Mt>(job, id);
}

// The persistent integrator.  Adjoint modes: ADJ_CHECKPOINT stores a
// checkpoint every K accepted steps; ADJ_REPLAY runs segment j of each ray
// of job.adj.rays from its checkpoint and records its steps.
template <class Mt, int SOLVER, bool MIXED>
__global__ void __launch_bounds__(256) PRR_OCCUPANCY k_integrate(JobDev job) {
    const int lane = int(threadIdx.x & 63);
    LocateCtx &C = job.loc[int64_t(blockIdx.x) * blockDim.x + threadIdx.x];
    const bool replay = Mt::replay && job.adj.mode == ADJ_REPLAY;
    Lane<Mt> L;
    int mode = MODE_IDLE;
    int32_t end = 0; // replay: stop at this step count
    bool exhausted = false;
    unsigned acc_att = 0, acc_ok = 0, acc_64 = 0, acc_rays = 0;
    // queue length: the render queue has padding between frames; adjoint
    // passes and explicit initial states run over job.total entries
    const int64_t qn = (Mt::ckpt || Mt::replay || job.explicit_x0u0) ? job.total : job.qtotal;
    while (true) {
        const unsigned long long idle = __ballot(mode == MODE_IDLE);
        const int nidle = __popcll(idle);
        if (!exhausted && nidle >= job.refill) {
            const int leader = __ffsll((long long)idle) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(&job.ctr->queue, (unsigned long long)nidle);
            base = shfl_u64(base, leader);
            if (base + (unsigned long long)nidle >= (unsigned long long)qn) exhausted = true;
            if (mode == MODE_IDLE) {
                const unsigned rank =
                    __builtin_amdgcn_mbcnt_hi(unsigned(idle >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(idle), 0u));
                const int64_t idx = int64_t(base) + rank;
                if (idx < qn) {
                    if (replay) { // task idx: segment adj.seg[idx] of ray adj.rays[idx]
                        const int64_t rid = job.adj.rays[idx];
                        const int32_t seg = job.adj.seg[idx];
                        lane_load_checkpoint<Mt>(job, L, rid, seg);
                        L.rec_i = idx;
                        end = min((seg + 1) * job.adj.K, job.adj.nsteps[rid]);
                        mode = MODE_ACTIVE;
                    } else {
                        // a checkpoint pass over some of the rays takes them from adj.rays
                        const bool some = Mt::ckpt && job.adj.mode == ADJ_CHECKPOINT && job.adj.rays;
                        const int64_t rid = some ? job.adj.rays[idx] : queue_ray(job, idx);
                        if (rid >= 0) { // -1: padding between frames, the lane stays idle
                            lane_load<Mt>(job, L, rid);
                            if (Mt::ckpt && job.adj.mode == ADJ_CHECKPOINT)
                                store_checkpoint(job, L.y, Mt::N, L.h, L.attempts, L.rid, 0);
                            mode = MODE_ACTIVE;
                        }
                    }
                    if (mode == MODE_ACTIVE && !finite_state<Mt>(L.y)) {
                        if (!replay) lane_store<Mt>(job, L, PRR_NAN);
                        acc_rays++;
                        mode = MODE_IDLE;
                    }
                }
            }
        }
        if (__ballot(mode != MODE_IDLE) == 0ull) break;
        if (mode != MODE_IDLE) {
            const int cls = lane_iterate<Mt, SOLVER, MIXED>(job, L, C, mode);
            if (cls >= 0 || (replay && mode == MODE_ACTIVE && L.steps >= end)) {
                if (!replay) {
                    lane_store<Mt>(job, L, cls);
                } else if (cls < 0) { // a segment's replay must land on the next checkpoint bit for bit
                    const double *nx = ckpt_at(job, L.rid, L.seg + 1);
                    bool same = L.steps == end;
                    for (int i = 0; i < Mt::N; ++i) same = same && L.y[i] == nx[i];
                    if (!same) atomicAdd(job.adj.mismatch, 1ull);
                }
                acc_att += unsigned(L.attempts);
                acc_ok += unsigned(L.steps);
                acc_64 += unsigned(L.att64);
                acc_rays++;
                mode = MODE_IDLE;
            }
        }
    }
    // one atomic per wave for the counters
    unsigned long long t_att = acc_att, t_ok = acc_ok, t_64 = acc_64, t_rays = acc_rays;
    for (int off = 32; off > 0; off >>= 1) {
        t_att += shfl_u64(t_att, lane ^ off);
        t_ok += shfl_u64(t_ok, lane ^ off);
        t_64 += shfl_u64(t_64, lane ^ off);
        t_rays += shfl_u64(t_rays, lane ^ off);
    }
    if (lane == 0) {
        atomicAdd(&job.ctr->attempts, t_att);
        atomicAdd(&job.ctr->accepted, t_ok);
        atomicAdd(&job.ctr->attempts_fp64, t_64);
        atomicAdd(&job.ctr->rays, t_rays);
    }
}

// Counter reset by a kernel (no memset node, so the sequence can be captured
// in a graph).
// all = 0 keeps the adjoint's mismatch count across launches.
__global__ void k_reset_counters(CountersDev *c, int all) {
    if (threadIdx.x == 0) {
        c->queue = 0;
        c->attempts = 0;
        c->accepted = 0;
        c->attempts_fp64 = 0;
        c->rays = 0;
        if (all) c->mismatch = 0;
    }
}

// ----------------------------------------------------------------- AA passes
__device__ inline float clamp01(float c) { return fminf(fmaxf(c, 0.0f), 1.0f); }

__device__ inline uint8_t to_rgb8(float c) {
    float v = roundf(255.0f * c); // cudamaths/image.cu:31
    v = fminf(fmaxf(v, 0.0f), 255.0f);
    return uint8_t(v);
}

// combineSamples (kernels/ssaa.cuh:109-131) of pixel p: the root mean square
// of its s sample colours, each clamped to [0, 1] after squaring (one sample:
// the clamped colour).  The result lies in [0, 1].
__device__ inline float3 ssaa_pixel(const PostFrameDev &pf, int64_t p) {
#pragma clang fp contract(off)
    const int s = pf.s;
    if (s == 1) {
        const float *src = pf.ray_color + 3 * p;
        return make_float3(clamp01(src[0]), clamp01(src[1]), clamp01(src[2]));
    }
    float acc[3] = {0.f, 0.f, 0.f};
    const float *src = pf.ray_color + 3 * p * s;
    for (int k = 0; k < s; ++k) {
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            const float c = src[3 * k + ch];
            acc[ch] += clamp01(c * c);
        }
    }
    return make_float3(sqrtf(acc[0] / float(s)), sqrtf(acc[1] / float(s)), sqrtf(acc[2] / float(s)));
}

// The SSAA pass of the frames without FXAA: one thread per pixel, frames
// along grid.y, the linear image and RGB8 written directly.  (FXAA frames
// combine their samples inside k_fxaa.)
__global__ void __launch_bounds__(256) k_ssaa(const PostFrameDev *pfs) {
    const PostFrameDev pf = pfs[blockIdx.y];
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t npix = int64_t(pf.W) * pf.H;
    if (pf.fxaa || p >= npix) return;
    const float3 o = ssaa_pixel(pf, p);
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

// luma (kernels/ssaa.cuh:141-143): double literals, float result.
__device__ inline float luma(float r, float g, float b) {
#pragma clang fp contract(off)
    const double k = double(1.0f / 255.0f);
    return float(double(r) * 0.299 * k + double(g) * 0.587 * k + double(b) * 0.114 * k);
}

// fxaa (kernels/ssaa.cuh:150-182) on every pixel with clamp-to-edge taps,
// fused with the SSAA combine before it and the RGB8 quantisation after it,
// so that the combined image never goes through HBM.  A workgroup of 512
// threads takes a 32 x 32 tile, two pixels per thread (rows ty and
// ty + 16): the combined colours of its pixels and of a 1-pixel
// halo (clamped to the image) are computed from the ray colours into LDS with
// their lumas, once per pixel, the loads of a thread all in flight together.
// A bilinear tap reads LDS when its 2 x 2 footprint lies in the staged
// 34 x 34 block; with luma / 255 (D14) and combined colours in [0, 1] the
// search span is at most about one pixel, so every tap does.  A footprint
// outside it would combine its pixels' samples from the ray colours again.
// The arithmetic is the same either way.
constexpr int kFxThreads = 512, kFxTile = 32, kFxHalo = kFxTile + 2, kFxRows = kFxThreads / kFxTile;

// The 2 x 2 footprint at (x0, y0) of a tap outside the staged block, combined
// from the ray colours (clamp-to-edge); out of line: no render reaches it.
struct Footprint {
    float3 c00, c10, c01, c11;
};
__device__ __noinline__ Footprint fxaa_footprint(const float *ray_color, int W, int H, int s, int x0, int y0) {
    PostFrameDev pf{};
    pf.ray_color = ray_color;
    pf.s = s;
    const int xa = min(max(x0, 0), W - 1), xb = min(max(x0 + 1, 0), W - 1);
    const int ya = min(max(y0, 0), H - 1), yb = min(max(y0 + 1, 0), H - 1);
    return Footprint{ssaa_pixel(pf, int64_t(ya) * W + xa), ssaa_pixel(pf, int64_t(ya) * W + xb),
                     ssaa_pixel(pf, int64_t(yb) * W + xa), ssaa_pixel(pf, int64_t(yb) * W + xb)};
}

__device__ inline float3 fxaa_tap_lds(const PostFrameDev &pf, const float (*tile)[kFxHalo][3], int ox, int oy,
                                      float X, float Y) {
#pragma clang fp contract(off)
    const float fx0 = floorf(X), fy0 = floorf(Y);
    const float fx = X - fx0, fy = Y - fy0;
    const int x0 = int(fx0), y0 = int(fy0);
    // The staged block holds image columns ox - 1 .. ox + 32 and rows oy - 1 ..
    // oy + 32 already clamped to the image, so a footprint inside it reads the
    // clamp-to-edge texels without clamping its own coordinates.
    const int lx = x0 - ox + 1, ly = y0 - oy + 1;
    float3 c00, c10, c01, c11;
    if (unsigned(lx) < unsigned(kFxHalo - 1) && unsigned(ly) < unsigned(kFxHalo - 1)) {
        c00 = make_float3(tile[ly][lx][0], tile[ly][lx][1], tile[ly][lx][2]);
        c10 = make_float3(tile[ly][lx + 1][0], tile[ly][lx + 1][1], tile[ly][lx + 1][2]);
        c01 = make_float3(tile[ly + 1][lx][0], tile[ly + 1][lx][1], tile[ly + 1][lx][2]);
        c11 = make_float3(tile[ly + 1][lx + 1][0], tile[ly + 1][lx + 1][1], tile[ly + 1][lx + 1][2]);
    } else {
        const Footprint f = fxaa_footprint(pf.ray_color, pf.W, pf.H, pf.s, x0, y0);
        c00 = f.c00;
        c10 = f.c10;
        c01 = f.c01;
        c11 = f.c11;
    }
    const float w00 = (1.0f - fx) * (1.0f - fy), w10 = fx * (1.0f - fy), w01 = (1.0f - fx) * fy, w11 = fx * fy;
    return make_float3(c00.x * w00 + c10.x * w10 + c01.x * w01 + c11.x * w11,
                       c00.y * w00 + c10.y * w10 + c01.y * w01 + c11.y * w11,
                       c00.z * w00 + c10.z * w10 + c01.z * w01 + c11.z * w11);
}

// combineSamples of S samples held in v (3 S floats): the same arithmetic
// as ssaa_pixel.
template <int S> __device__ inline float3 ssaa_combine(const float *v) {
#pragma clang fp contract(off)
    if constexpr (S == 1) {
        return make_float3(clamp01(v[0]), clamp01(v[1]), clamp01(v[2]));
    } else {
        float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < S; ++k) {
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) acc[ch] += clamp01(v[3 * k + ch] * v[3 * k + ch]);
        }
        return make_float3(sqrtf(acc[0] / float(S)), sqrtf(acc[1] / float(S)), sqrtf(acc[2] / float(S)));
    }
}

// The 3 S floats of a pixel's samples at src into v, 16 bytes per load
// (S a multiple of 4, src 16-byte aligned) or 4 (S = 1).
template <int S> __device__ inline void load_samples(const float *src, float *v) {
    if constexpr (S % 4 == 0) {
#pragma unroll
        for (int i = 0; i < 3 * S / 4; ++i) {
            const float4 q = reinterpret_cast<const float4 *>(src)[i];
            v[4 * i] = q.x;
            v[4 * i + 1] = q.y;
            v[4 * i + 2] = q.z;
            v[4 * i + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 3 * S; ++i) v[i] = src[i];
    }
}

// This thread's pixels of the 34 x 34 block (1156 pixels, at most three per
// thread) of the tile at (ox, oy), coordinates clamped to the image.  S: the
// frame's sample count when it is 1, 4 or 16 (the last two with 16-byte-
// aligned ray colours), else 0 (any count).  For S = 1 and 4 the raw samples
// are kept, all of a thread's loads issued before fxaa_put combines them;
// otherwise (16 samples would not fit the registers) the combined colours.
constexpr int kFxStage = kFxHalo * kFxHalo, kFxPer = (kFxStage + kFxThreads - 1) / kFxThreads;
template <int S> struct FxStaged {
    static constexpr bool raw = S == 1 || S == 4;
    float v[kFxPer][raw ? 3 * S : 3];
};

template <int S> __device__ inline void fxaa_get(const PostFrameDev &pf, int ox, int oy, FxStaged<S> &st) {
    constexpr int kUnroll = S == 16 ? 1 : kFxPer;
    const int W = pf.W, H = pf.H;
#pragma unroll kUnroll
    for (int r = 0; r < kFxPer; ++r) {
        const int t = int(threadIdx.x) + kFxThreads * r;
        if (t < kFxStage) {
            const int ly = t / kFxHalo, lx = t - ly * kFxHalo;
            const int xx = min(max(ox + lx - 1, 0), W - 1), yy = min(max(oy + ly - 1, 0), H - 1);
            const int64_t p = int64_t(yy) * W + xx;
            if constexpr (FxStaged<S>::raw) {
                load_samples<S>(pf.ray_color + 3 * S * p, st.v[r]);
            } else {
                float3 c;
                if constexpr (S == 16) {
                    float v[3 * S];
                    load_samples<S>(pf.ray_color + 3 * S * p, v);
                    c = ssaa_combine<S>(v);
                } else {
                    c = ssaa_pixel(pf, p);
                }
                st.v[r][0] = c.x;
                st.v[r][1] = c.y;
                st.v[r][2] = c.z;
            }
        }
    }
}

// The staged pixels, combined, into LDS with their lumas.
template <int S>
__device__ inline void fxaa_put(const FxStaged<S> &st, float (*tile)[kFxHalo][3], float (*lum)[kFxHalo]) {
#pragma unroll
    for (int r = 0; r < kFxPer; ++r) {
        const int t = int(threadIdx.x) + kFxThreads * r;
        if (t < kFxStage) {
            const int ly = t / kFxHalo, lx = t - ly * kFxHalo;
            float3 c;
            if constexpr (FxStaged<S>::raw)
                c = ssaa_combine<S>(st.v[r]);
            else
                c = make_float3(st.v[r][0], st.v[r][1], st.v[r][2]);
            tile[ly][lx][0] = c.x;
            tile[ly][lx][1] = c.y;
            tile[ly][lx][2] = c.z;
            lum[ly][lx] = luma(c.x, c.y, c.z);
        }
    }
}

// FXAA of the staged tile at (ox, oy): two pixels per thread (rows ty and
// ty + 16), written as the linear image and RGB8.
__device__ inline void fxaa_filter(const PostFrameDev &pf, const float (*tile)[kFxHalo][3], const float (*lum)[kFxHalo],
                                   int ox, int oy) {
#pragma clang fp contract(off)
    const int W = pf.W, H = pf.H;
    const int tx = int(threadIdx.x) % kFxTile, ty0 = int(threadIdx.x) / kFxTile;
    const int x = ox + tx;
    if (x >= W) return;
#pragma unroll
    for (int k = 0; k < kFxTile / kFxRows; ++k) {
        const int ty = ty0 + kFxRows * k, y = oy + ty;
        if (y >= H) break;
        const float lNW = lum[ty][tx], lNE = lum[ty][tx + 2], lSW = lum[ty + 2][tx], lSE = lum[ty + 2][tx + 2];
        const float lM = lum[ty + 1][tx + 1];
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
        const float3 t1 = fxaa_tap_lds(pf, tile, ox, oy, fxp + dx * k1, fyp + dy * k1);
        const float3 t2 = fxaa_tap_lds(pf, tile, ox, oy, fxp + dx * k2, fyp + dy * k2);
        const float3 t3 = fxaa_tap_lds(pf, tile, ox, oy, fxp + dx * -0.5f, fyp + dy * -0.5f);
        const float3 t4 = fxaa_tap_lds(pf, tile, ox, oy, fxp + dx * 0.5f, fyp + dy * 0.5f);
        const float3 A = make_float3(t1.x * 0.5f + t2.x * 0.5f, t1.y * 0.5f + t2.y * 0.5f, t1.z * 0.5f + t2.z * 0.5f);
        const float3 B = make_float3(0.5f * A.x + (t3.x * 0.25f + t4.x * 0.25f),
                                     0.5f * A.y + (t3.y * 0.25f + t4.y * 0.25f),
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
}

// One tile: stage, then filter.
template <int S>
__device__ inline void fxaa_tile(const PostFrameDev &pf, float (*tile)[kFxHalo][3], float (*lum)[kFxHalo], int ox,
                                 int oy) {
    FxStaged<S> c;
    fxaa_get<S>(pf, ox, oy, c);
    fxaa_put<S>(c, tile, lum);
    __syncthreads();
    fxaa_filter(pf, tile, lum, ox, oy);
}

// fxaa with the SSAA combine (see above): grid (tiles in x, tiles in y,
// frames), one 32 x 32 tile per workgroup.
__global__ void __launch_bounds__(kFxThreads) k_fxaa(const PostFrameDev *pfs) {
    __shared__ float tile[kFxHalo][kFxHalo][3];
    __shared__ float lum[kFxHalo][kFxHalo];
    const PostFrameDev pf = pfs[blockIdx.z];
    const int ox = int(blockIdx.x) * kFxTile, oy = int(blockIdx.y) * kFxTile;
    if (!pf.fxaa || ox >= pf.W || oy >= pf.H) return; // uniform over the workgroup
    const bool al16 = (reinterpret_cast<uintptr_t>(pf.ray_color) & 15) == 0;
    switch (pf.s) {
    case 1: fxaa_tile<1>(pf, tile, lum, ox, oy); break;
    case 4: al16 ? fxaa_tile<4>(pf, tile, lum, ox, oy) : fxaa_tile<0>(pf, tile, lum, ox, oy); break;
    case 16: al16 ? fxaa_tile<16>(pf, tile, lum, ox, oy) : fxaa_tile<0>(pf, tile, lum, ox, oy); break;
    default: fxaa_tile<0>(pf, tile, lum, ox, oy); break;
    }
}

// ----------------------------------------------------------------- adjoint (Kerr)
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

// Terminal seed: lambda at y_{N-1} (state and constants) and dL/dphi_s from
// the located event of the last step (see adjoint.hpp).
__device__ void adj_terminal(const JobDev &job, const AdjDev &ad, const FrameDev &fr, const Consts<double> &cs,
                             const EventDev &ev, const double w[3], double lam[9], double &gphs) {
    const double tau = ev.theta * ev.h;
    double ys[6];
    rk_substep<Kerr>(ev.y_prev, tau, cs, ys);
    using D = Dual<10>;
    D y[5];
    for (int i = 0; i < 5; ++i) y[i] = D::var(ys[i], i);
    const D E = D::var(cs.E, 5), Lz = D::var(cs.L, 6), Q = D::var(cs.Q, 7), a = D::var(cs.a, 8);
    D col[3], G;
    if (ev.cls == PRR_ESCAPE) {
        const V3<D> n = kerr_escape_dir_t<D>(y, E, Lz, Q, a, cs.M);
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
    Kerr::rhs<double>(ys, Kerr::coef<double>(cs), f);
    double num = 0.0, den = 0.0;
    for (int i = 0; i < 5; ++i) {
        num += ls[i] * f[i];
        den += G.d[i] * f[i];
    }
    double kappa = num / den;
    if (!isfinite(kappa)) kappa = 0.0;
    for (int i = 0; i < 9; ++i) lam[i] = ls[i] - kappa * G.d[i];
    gphs -= kappa * G.d[9];
    kerr_step_vjp(ev.y_prev, tau, cs, lam);
}

// Reverse sweep over segment j of every ray of this launch: the terminal seed
// on the ray's last segment, then lambda_n = lambda_{n+1} dPhi_n/dy_n over
// the replayed steps.  lambda lives in ad.lam between launches.
// Seed of the reverse sweep at each ray's terminal event: lambda from the
// pixel weights of its colour (forward-mode duals through the escape
// direction and sky lookup, or the sphere's texture and normal, plus the
// located event's fraction) into ad.lam, for the rays whose last segment lies
// in this chunk (ranks seed_lo..count-1).  A kernel of its own, so that the
// duals' registers do not weigh on the sweep.
__global__ void __launch_bounds__(256) k_adj_seed(JobDev job, AdjDev ad) {
    const int64_t idx = ad.seed_lo + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= ad.count) return;
    const int64_t id = ad.rays[idx];
    double *lm = ad.lam + 10 * id;
    const FrameDev &fr = job.frames[find_frame(job, id)];
    Consts<double> cs;
    load_consts<Kerr>(job, id, fr, cs);
    const int cls = ad.cls[id];
    double w[3] = {0.0, 0.0, 0.0};
    if (cls == PRR_ESCAPE || cls == PRR_SPHERE) color_weights(ad, fr, id - fr.ray_begin, ad.ray_color + 3 * id, w);
    if (ad.event[id].cls != cls && (cls == PRR_ESCAPE || cls == PRR_SPHERE)) atomicAdd(ad.mismatch, 1ull);
    if ((w[0] == 0.0 && w[1] == 0.0 && w[2] == 0.0) || ad.event[id].cls != cls) {
        for (int i = 0; i < 10; ++i) lm[i] = 0.0;
        return;
    }
    double lam[9], gphs;
    adj_terminal(job, ad, fr, cs, ad.event[id], w, lam, gphs);
    for (int i = 0; i < 9; ++i) lm[i] = lam[i];
    lm[9] = gphs;
}

// Reverse sweep over the segment levels jhi..jlo of one replayed chunk: one
// thread per ray (rank idx in the decreasing-segment-count order), from its
// highest segment in the chunk down to jlo, reading segment j's steps from
// task levelstart[j] + idx of the chunk's records.  lambda enters from ad.lam
// (the seed, or the previous chunk) and leaves through it.
// Record of step n (in segment j) of ray id at rank idx: y[0..4] and h at
// p[0], p[s], ..., p[5 s] (the replayed chunk's task-major records, or the
// first pass's per-ray records).
struct RecRef {
    const double *p;
    int64_t s;
};
__device__ inline RecRef adj_rec(const AdjDev &ad, int64_t id, int64_t idx, int32_t j, int32_t n) {
    if (ad.rec_cap > 0) return RecRef{ad.rec + (id * ad.rec_cap + n) * 6, 1};
    const int64_t T = ad.rec_stride;
    return RecRef{ad.rec + (int64_t(n - j * ad.K) * 6) * T + (ad.levelstart[j] - ad.t0 + idx), T};
}

// Column k of the transposed Jacobian of segment j of the ray at rank idx:
// the unit cotangent e_k swept backwards through the segment's K steps.
__device__ void adj_segment_jacobian(const JobDev &job, const AdjDev &ad, int64_t task, int k) {
    int lo = 0, hi = ad.jlevels - 1; // the level: the largest j with jacstart[j] <= task
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ad.jacstart[mid] <= task) lo = mid;
        else hi = mid - 1;
    }
    const int32_t j = lo;
    const int64_t idx = task - ad.jacstart[j];
    const int64_t id = ad.rays[idx];
    const FrameDev &fr = job.frames[find_frame(job, id)];
    Consts<double> cs;
    load_consts<Kerr>(job, id, fr, cs);
    double lam[9];
    for (int i = 0; i < 9; ++i) lam[i] = i == k ? 1.0 : 0.0;
    const int K = ad.K;
    for (int32_t n = (j + 1) * K - 1; n >= j * K; --n) { // a segment below the last: K steps
        const RecRef r = adj_rec(ad, id, idx, j, n);
        double yn[6];
        for (int c = 0; c < 5; ++c) yn[c] = r.p[c * r.s];
        yn[5] = 0.0;
        kerr_step_vjp(yn, r.p[5 * r.s], cs, lam);
    }
    double *J = ad.jac + 81 * task + 9 * k;
    for (int i = 0; i < 9; ++i) J[i] = lam[i];
}

// Blocks below ceil(count / 256) sweep rays; the rest form the long rays'
// segment Jacobians (9 threads per segment), concurrently with the sweeps.
__global__ void __launch_bounds__(256) k_adj_reverse(JobDev job, AdjDev ad) {
    const int64_t ray_blocks = (ad.count + 255) / 256;
    if (int64_t(blockIdx.x) >= ray_blocks) {
        const int64_t t = (int64_t(blockIdx.x) - ray_blocks) * blockDim.x + threadIdx.x;
        if (t < 9 * ad.njac) adj_segment_jacobian(job, ad, t / 9, int(t % 9));
        return;
    }
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= ad.count) return;
    const int64_t id = ad.rays[idx];
    const int32_t N = ad.nsteps[id];
    const int K = ad.K;
    const int32_t last = (N - 1) / K; // the ray's last segment (the rank order puts it at >= jlo)
    double *lm = ad.lam + 10 * id;
    double lam[9];
    for (int i = 0; i < 9; ++i) lam[i] = lm[i];
    bool zero = true;
    for (int i = 0; i < 9; ++i) zero = zero && lam[i] == 0.0;
    if (zero) return; // nothing flows back (lm stays zero)
    const FrameDev &fr = job.frames[find_frame(job, id)];
    Consts<double> cs;
    load_consts<Kerr>(job, id, fr, cs);
    const int32_t jtop = min(last, ad.jhi);
    // a long ray sweeps its last segment only; the chain applies the others
    const int32_t jbot = idx < ad.nlong ? jtop : ad.jlo;
    for (int32_t j = jtop; j >= jbot; --j) {
        const int32_t end = min((j + 1) * K, N - 1); // step N-1 is the located sub-step above
        for (int32_t n = end - 1; n >= j * K; --n) {
            const RecRef r = adj_rec(ad, id, idx, j, n);
            double yn[6];
            for (int c = 0; c < 5; ++c) yn[c] = r.p[c * r.s];
            yn[5] = 0.0;
            kerr_step_vjp(yn, r.p[5 * r.s], cs, lam);
        }
    }
    for (int i = 0; i < 9; ++i) lm[i] = lam[i];
}

// The long rays' lambda through their segments below the last, from the
// highest: lambda <- J_j^T lambda.
__global__ void __launch_bounds__(256) k_adj_chain(AdjDev ad) {
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= ad.nlong) return;
    const int64_t id = ad.rays[idx];
    double *lm = ad.lam + 10 * id;
    double lam[9];
    bool zero = true;
    for (int i = 0; i < 9; ++i) {
        lam[i] = lm[i];
        zero = zero && lam[i] == 0.0;
    }
    if (zero) return; // nothing flows back (as in the sweep)
    const int32_t last = (ad.nsteps[id] - 1) / ad.K;
    for (int32_t j = last - 1; j >= 0; --j) {
        const double *J = ad.jac + 81 * (ad.jacstart[j] + idx);
        double out[9];
        for (int i = 0; i < 9; ++i) out[i] = 0.0;
        for (int k = 0; k < 9; ++k)
            for (int i = 0; i < 9; ++i) out[i] += lam[k] * J[9 * k + i];
        for (int i = 0; i < 9; ++i) lam[i] = out[i];
    }
    for (int i = 0; i < 9; ++i) lm[i] = lam[i];
}

// Ray set-up Jacobian d(y_0, E, L, Q)/d(a, r_cam, incl) contracted with
// lambda_0: the per-ray gradient.
__global__ void __launch_bounds__(256) k_adj_init_grad(JobDev job, AdjDev ad) {
    const int64_t id = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= job.total) return;
    const double *lam = ad.lam + 10 * id;
    double *g = ad.grad_ray + 4 * id;
    bool zero = true;
    for (int i = 0; i < 10; ++i) zero = zero && lam[i] == 0.0;
    if (zero || ad.nsteps[id] <= 0) {
        for (int k = 0; k < 4; ++k) g[k] = 0.0;
        return;
    }
    const FrameDev &fr = job.frames[find_frame(job, id)];
    double u, v;
    ray_uv(fr, id - fr.ray_begin, u, v);
    const double r_cam = sqrt(fr.pos[0] * fr.pos[0] + fr.pos[1] * fr.pos[1] + fr.pos[2] * fr.pos[2]);
    const double incl = acos(fr.pos[2] / r_cam);
    using D = Dual<3>;
    V3<D> pos, d;
    orbit_camera_ray<D>(D::var(r_cam, 1), D::var(incl, 2), ad.cam_phi, fr.tan_half, fr.aspect, u, v, pos, d);
    D z[8];
    kerr_init_t<D>(pos, d, fr.M, D::var(fr.a, 0), z);
    double out[4];
    for (int k = 0; k < 3; ++k) {
        double acc = k == 0 ? lam[8] : 0.0; // a is itself one of the constants
        for (int i = 0; i < 8; ++i) acc += lam[i] * z[i].d[k];
        out[k] = acc;
    }
    out[3] = lam[9];
    bool fin = true;
    for (int k = 0; k < 4; ++k) fin = fin && isfinite(out[k]);
    for (int k = 0; k < 4; ++k) g[k] = fin ? out[k] : 0.0;
}

// ------------------------------------------------------------ adjoint plan
__device__ unsigned long long wave_sum(unsigned long long v) {
    const int lane = int(threadIdx.x & 63);
    for (int o = 32; o > 0; o >>= 1) v += shfl_u64(v, lane ^ o);
    return v;
}

__global__ void __launch_bounds__(256) k_adj_plan_clear(AdjPlanDev p) {
    const int i = int(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < p.nbins) p.hist[i] = 0ull;
    if (i < 4) p.ctr[i] = 0ull;
}

// Histogram bin of a ray with st accepted steps: its segment count n, then
// the length class of its last segment (1..K steps in L classes, longer in
// higher bins).  Ordered by decreasing bin, the rays with more than j
// segments stay a prefix of the order, and at each level the rays whose last
// segment it is follow the full segments from the longest down, so that a
// wave of replay or reverse-sweep tasks holds segments of similar length.
__device__ inline int plan_bin(const AdjPlanDev &p, int32_t st) {
    const int32_t n = st > 0 ? (st - 1) / p.K + 1 : 0;
    const int32_t c = n > 0 ? (st - (n - 1) * p.K - 1) * p.L / p.K : 0;
    return min(n * p.L + c, p.nbins - 1);
}

// Counts are gathered per workgroup in LDS and flushed with one global atomic
// per non-empty bin, so that the few bins do not serialise every wave.
__global__ void __launch_bounds__(256) k_adj_plan_count(AdjPlanDev p) {
    __shared__ unsigned lcount[kAdjPlanBins];
    __shared__ unsigned long long lsum[4];
    __shared__ unsigned lmax[4];
    for (int b = threadIdx.x; b < p.nbins; b += blockDim.x) lcount[b] = 0u;
    __syncthreads();
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool on = i < p.T;
    const int32_t st = on ? p.steps[i] : 0;
    const int32_t n = st > 0 ? (st - 1) / p.K + 1 : 0;
    if (on) {
        p.nseg[i] = n;
        if (n <= p.cap) {
            p.off[i] = i * p.cap;
        } else { // rare: checkpoints in the overflow area
            const unsigned long long base = atomicAdd(&p.ctr[0], (unsigned long long)n);
            p.off[i] = -1 - int64_t(base);
            p.over[atomicAdd(&p.ctr[1], 1ull)] = i;
        }
        atomicAdd(&lcount[plan_bin(p, st)], 1u);
    }
    const unsigned long long sum = wave_sum((unsigned long long)st);
    unsigned mx = unsigned(n);
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, unsigned(__shfl_xor(int(mx), o)));
    const int w = int(threadIdx.x >> 6);
    if ((threadIdx.x & 63) == 0) {
        lsum[w] = sum;
        lmax[w] = mx;
    }
    __syncthreads();
    for (int b = threadIdx.x; b < p.nbins; b += blockDim.x)
        if (lcount[b]) atomicAdd(&p.hist[b], (unsigned long long)lcount[b]);
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        unsigned m = 0;
        for (int k = 0; k < int(blockDim.x >> 6); ++k) {
            s += lsum[k];
            m = max(m, lmax[k]);
        }
        atomicAdd(&p.ctr[2], s);
        atomicMax(&p.ctr[3], (unsigned long long)m);
    }
}

// Each workgroup claims one run of positions per bin it holds; its rays take
// consecutive positions of that run.
__global__ void __launch_bounds__(256) k_adj_plan_scatter(AdjPlanDev p) {
    __shared__ unsigned lcount[kAdjPlanBins];
    __shared__ unsigned long long lbase[kAdjPlanBins];
    for (int b = threadIdx.x; b < p.nbins; b += blockDim.x) lcount[b] = 0u;
    __syncthreads();
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool on = i < p.T;
    const int n = on ? plan_bin(p, p.steps[i]) : 0;
    const unsigned rank = on ? atomicAdd(&lcount[n], 1u) : 0u;
    __syncthreads();
    for (int b = threadIdx.x; b < p.nbins; b += blockDim.x)
        if (lcount[b]) lbase[b] = atomicAdd(&p.cursor[b], (unsigned long long)lcount[b]);
    __syncthreads();
    if (on) p.order[lbase[n] + rank] = i;
}

__global__ void __launch_bounds__(256) k_adj_plan_tasks(AdjPlanDev p) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= p.ntask) return;
    // level j of task t: the smallest j with levelstart[j] <= t (levelstart falls with j)
    int lo = 0, hi = p.nlevels - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (p.levelstart[mid] <= t) hi = mid;
        else lo = mid + 1;
    }
    p.task_ray[t] = p.order[t - p.levelstart[lo]];
    p.task_seg[t] = lo;
}

hipError_t launch_adj_plan_count(const AdjPlanDev &p, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_plan_clear, dim3(unsigned((p.nbins + 255) / 256)), dim3(256), 0, st, p);
    if (p.T > 0) hipLaunchKernelGGL(k_adj_plan_count, dim3(unsigned((p.T + 255) / 256)), dim3(256), 0, st, p);
    return hipGetLastError();
}

hipError_t launch_adj_plan_scatter(const AdjPlanDev &p, hipStream_t st) {
    if (p.T > 0) hipLaunchKernelGGL(k_adj_plan_scatter, dim3(unsigned((p.T + 255) / 256)), dim3(256), 0, st, p);
    return hipGetLastError();
}

hipError_t launch_adj_plan_tasks(const AdjPlanDev &p, hipStream_t st) {
    if (p.ntask > 0)
        hipLaunchKernelGGL(k_adj_plan_tasks, dim3(unsigned((p.ntask + 255) / 256)), dim3(256), 0, st, p);
    return hipGetLastError();
}

// Deterministic sum of the per-ray gradients: one workgroup, fixed order.
// Two levels: kGradParts workgroups each sum a fixed contiguous range of
// rays into partial[b], then one workgroup sums the partials; every sum has a
// fixed order, whatever the hardware schedule.
constexpr int kGradParts = 256;

__device__ void block_sum4(double acc[4], double (*part)[256]) {
    const int t = int(threadIdx.x);
    for (int k = 0; k < 4; ++k) part[k][t] = acc[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s)
            for (int k = 0; k < 4; ++k) part[k][t] += part[k][t + s];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_grad_reduce_part(const double *grad_ray, int64_t n, double *partial) {
    __shared__ double part[4][256];
    const int64_t chunk = (n + kGradParts - 1) / kGradParts;
    const int64_t lo = int64_t(blockIdx.x) * chunk, hi = min(n, lo + chunk);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t i = lo + threadIdx.x; i < hi; i += 256)
        for (int k = 0; k < 4; ++k) acc[k] += grad_ray[4 * i + k];
    block_sum4(acc, part);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) partial[4 * blockIdx.x + k] = part[k][0];
}

__global__ void __launch_bounds__(256) k_grad_reduce_final(const double *partial, double *out) {
    __shared__ double part[4][256];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < kGradParts; b += 256)
        for (int k = 0; k < 4; ++k) acc[k] += partial[4 * b + k];
    block_sum4(acc, part);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) out[k] = part[k][0];
}

hipError_t launch_adj_reverse(const JobDev &job, const AdjDev &ad, hipStream_t st) {
    if (ad.count <= 0) return hipSuccess;
    if (ad.count > ad.seed_lo)
        hipLaunchKernelGGL(k_adj_seed, dim3(unsigned((ad.count - ad.seed_lo + 255) / 256)), dim3(256), 0, st, job,
                           ad);
    const int64_t blocks = (ad.count + 255) / 256 + (9 * ad.njac + 255) / 256;
    hipLaunchKernelGGL(k_adj_reverse, dim3(unsigned(blocks)), dim3(256), 0, st, job, ad);
    if (ad.nlong > 0) hipLaunchKernelGGL(k_adj_chain, dim3(unsigned((ad.nlong + 255) / 256)), dim3(256), 0, st, ad);
    return hipGetLastError();
}

hipError_t launch_adj_init_grad(const JobDev &job, const AdjDev &ad, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_init_grad, dim3(unsigned((job.total + 255) / 256)), dim3(256), 0, st, job, ad);
    return hipGetLastError();
}

hipError_t launch_grad_reduce(const double *grad_ray, int64_t n, double *partial, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_grad_reduce_part, dim3(kGradParts), dim3(256), 0, st, grad_ray, n, partial);
    hipLaunchKernelGGL(k_grad_reduce_final, dim3(1), dim3(256), 0, st, partial, out);
    return hipGetLastError();
}

// ----------------------------------------------------------------- dispatch
template <class F> static hipError_t with_metric(int metric, F &&f) {
    switch (metric) {
    case PRR_FLAT: return f(Flat{});
    case PRR_SCHW_KS_CART: return f(SchwKS{});
    case PRR_SCHW_SPH: return f(SchwSph{});
    case PRR_KERR_BL: return f(Kerr{});
    case PRR_ELLIS: return f(Ellis{});
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_init(int metric, const JobDev &job, hipStream_t st) {
    return with_metric(metric, [&](auto mt) {
        using Mt = decltype(mt);
        hipLaunchKernelGGL((k_init<Mt>), dim3(unsigned((job.total + 255) / 256)), dim3(256), 0, st, job);
        return hipGetLastError();
    });
}

hipError_t launch_shade(int metric, const JobDev &job, hipStream_t st) {
    return with_metric(metric, [&](auto mt) {
        using Mt = decltype(mt);
        hipLaunchKernelGGL((k_shade<Mt>), dim3(unsigned((job.total + 255) / 256)), dim3(256), 0, st, job);
        return hipGetLastError();
    });
}

template <class Mt, int SOLVER> void launch_integrate_s(const JobDev &job, dim3 g, dim3 b, hipStream_t st) {
    if (job.tr.precision == PRR_FP32_BULK) hipLaunchKernelGGL((k_integrate<Mt, SOLVER, true>), g, b, 0, st, job);
    else hipLaunchKernelGGL((k_integrate<Mt, SOLVER, false>), g, b, 0, st, job);
}

template <class Mt> void launch_integrate_t(const JobDev &job, dim3 g, dim3 b, hipStream_t st) {
    switch (job.tr.solver) {
    case PRR_EULER: launch_integrate_s<Mt, PRR_EULER>(job, g, b, st); break;
    case PRR_RK4: launch_integrate_s<Mt, PRR_RK4>(job, g, b, st); break;
    case PRR_DORMAND: launch_integrate_s<Mt, PRR_DORMAND>(job, g, b, st); break;
    default: launch_integrate_s<Mt, PRR_RKF45>(job, g, b, st); break;
    }
}

template <class Mt> static int waves_of(int solver) {
    switch (solver) {
    case PRR_EULER: return WavesOf<Mt, PRR_EULER>::value;
    case PRR_RK4: return WavesOf<Mt, PRR_RK4>::value;
    case PRR_DORMAND: return WavesOf<Mt, PRR_DORMAND>::value;
    default: return WavesOf<Mt, PRR_RKF45>::value;
    }
}

int integrate_waves(int metric, int solver) {
    switch (metric) {
    case PRR_FLAT: return waves_of<NoAdj<Flat>>(solver);
    case PRR_SCHW_KS_CART: return waves_of<NoAdj<SchwKS>>(solver);
    case PRR_SCHW_SPH: return waves_of<NoAdj<SchwSph>>(solver);
    case PRR_KERR_BL: return waves_of<NoAdj<Kerr>>(solver);
    default: return waves_of<NoAdj<Ellis>>(solver);
    }
}

hipError_t launch_integrate(int metric, const JobDev &job, int grid, int block, hipStream_t st) {
    const dim3 g(grid), b(block);
    return with_metric(metric, [&](auto mt) {
        using Mt = decltype(mt);
        if (job.adj.mode == ADJ_NONE) {
            launch_integrate_t<NoAdj<Mt>>(job, g, b, st);
        } else if constexpr (Mt::kind == PRR_KERR_BL) { // the adjoint differentiates Kerr only
            if (job.adj.mode == ADJ_CHECKPOINT) launch_integrate_t<AdjCkpt<Mt>>(job, g, b, st);
            else launch_integrate_t<AdjReplay<Mt>>(job, g, b, st);
        } else {
            return hipErrorInvalidValue;
        }
        return hipGetLastError();
    });
}

hipError_t launch_reset(CountersDev *c, hipStream_t st, bool all) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, st, c, int(all));
    return hipGetLastError();
}

hipError_t launch_ssaa(const PostFrameDev *pfs, int nframes, int64_t max_pixels, hipStream_t st) {
    const dim3 g(unsigned((max_pixels + 255) / 256), unsigned(nframes)), b(256);
    hipLaunchKernelGGL(k_ssaa, g, b, 0, st, pfs);
    return hipGetLastError();
}

hipError_t launch_fxaa(const PostFrameDev *pfs, int nframes, int max_w, int max_h, hipStream_t st) {
    const dim3 g(unsigned((max_w + kFxTile - 1) / kFxTile), unsigned((max_h + kFxTile - 1) / kFxTile),
                 unsigned(nframes)),
        b(kFxThreads);
    hipLaunchKernelGGL(k_fxaa, g, b, 0, st, pfs);
    return hipGetLastError();
}

} // namespace prr
