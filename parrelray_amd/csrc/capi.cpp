// capi.cpp -- the C ABI of include/prr.h: context, device buffers, frame
// set-up and the per-call launch sequence (counter reset -> persistent
// integrator -> SSAA -> FXAA + RGB8), all on one HIP stream.
//
// Replaces the reference's host orchestration (prr.cpp:237-404): global
// Tracer/Spacetime/Camera objects, managed ray buffers, seven synchronous
// launches per frame and exit(99) on any error.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "internal.hpp"
#include "../../include/prr_debug.h"
#include "frame_host.hpp"
#include "kernels.hpp"
#include "context.hpp"
#include "prr_host.hpp"

using namespace prr;

namespace {

thread_local std::string g_error;

} // namespace

namespace prr {
void set_global_error(const std::string &msg) { g_error = msg; }

// Lays the workspace out from `base` (null: offsets only) and returns its size in bytes.
size_t carve_workspace(void *base, const WorkspaceShape &s, Workspace *ws) {
    size_t off = 0;
    auto take = [&](size_t bytes) -> void * {
        void *p = base ? static_cast<char *>(base) + off : nullptr;
        off += (bytes + 255) / 256 * 256;
        return p;
    };
    const size_t R = size_t(std::max<int64_t>(s.rays, 1));
    Workspace w;
    w.frames = static_cast<FrameDev *>(take(sizeof(FrameDev) * size_t(std::max(s.nframes, 1))));
    w.post = static_cast<PostFrameDev *>(take(sizeof(PostFrameDev) * size_t(std::max(s.nframes, 1))));
    w.counters = static_cast<CountersDev *>(take(sizeof(CountersDev)));
    w.color = static_cast<float *>(take(12 * R));
    w.rs_y = static_cast<double *>(take(8 * size_t(s.state_width) * R));
    w.rs_cls = static_cast<int8_t *>(take(R));
    w.rs_hit = static_cast<int8_t *>(take(R));
    w.rs_steps = static_cast<int32_t *>(take(4 * R));
    w.rs_attempts = static_cast<int32_t *>(take(4 * R));
    w.loc = static_cast<LocateCtx *>(take(sizeof(LocateCtx) * size_t(std::max<int64_t>(s.lanes, 1))));
    if (ws) *ws = w;
    return off;
}
} // namespace prr

prr_status prr::fail(prr_ctx *ctx, prr_status st, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    g_error = buf;
    return st;
}

// An exception that reached the C ABI (none may cross it): its status and
// message, on the context when there is one.
prr_status prr::abi_exception(prr_ctx *ctx) noexcept {
    try {
        throw;
    } catch (const std::bad_alloc &) {
        return fail(ctx, PRR_ENOMEM, "out of host memory");
    } catch (const std::exception &e) {
        return fail(ctx, PRR_EINVAL, "%s", e.what());
    } catch (...) {
        return fail(ctx, PRR_EINVAL, "unexpected exception");
    }
}

prr_status prr::ensure(prr_ctx *ctx, DevBuf &b, size_t bytes) {
    if (b.bytes >= bytes && b.p) return PRR_OK;
    if (b.p) {
        hipError_t e = hipFree(b.p);
        if (e != hipSuccess) return fail(ctx, PRR_EHIP, "hipFree failed: %s", hipGetErrorString(e));
        b.p = nullptr;
        b.bytes = 0;
    }
    const size_t alloc = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&b.p, alloc);
    if (e != hipSuccess) return fail(ctx, PRR_ENOMEM, "hipMalloc(%zu) failed: %s", alloc, hipGetErrorString(e));
    b.bytes = alloc;
    return PRR_OK;
}

bool prr::grow_quietly(DevBuf &b, size_t bytes) {
    if (b.bytes >= bytes && b.p) return true;
    if (b.p) {
        if (hipFree(b.p) != hipSuccess) return false;
        b.p = nullptr;
        b.bytes = 0;
    }
    if (hipMalloc(&b.p, bytes) != hipSuccess) {
        b.p = nullptr;
        (void)hipGetLastError(); // the failed allocation leaves no error behind
        return false;
    }
    b.bytes = bytes;
    return true;
}

prr_status prr::ensure_pinned(prr_ctx *ctx, size_t bytes) {
    if (ctx->pinned_bytes >= bytes && ctx->pinned) return PRR_OK;
    if (ctx->pinned) {
        hipError_t e = hipHostFree(ctx->pinned);
        if (e != hipSuccess) return fail(ctx, PRR_EHIP, "hipHostFree failed: %s", hipGetErrorString(e));
        ctx->pinned = nullptr;
        ctx->pinned_bytes = 0;
    }
    hipError_t e = hipHostMalloc(&ctx->pinned, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        ctx->pinned = nullptr;
        return fail(ctx, PRR_ENOMEM, "hipHostMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    ctx->pinned_bytes = bytes;
    return PRR_OK;
}

namespace {

void release(DevBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

prr_status validate_tracer(prr_ctx *ctx, const prr_tracer &t) {
    if (t.solver != PRR_EULER && t.solver != PRR_RKF45 && t.solver != PRR_DORMAND && t.solver != PRR_RK4)
        return fail(ctx, PRR_EINVAL, "unknown solver %d", t.solver);
    if (t.precision != PRR_FP64 && t.precision != PRR_FP32_BULK)
        return fail(ctx, PRR_EINVAL, "unknown precision %d", t.precision);
    if (!(t.h > 0.0) || !std::isfinite(t.h)) return fail(ctx, PRR_EINVAL, "step h must be positive");
    if (t.solver == PRR_RKF45 || t.solver == PRR_DORMAND) {
        if (!(t.hmin > 0.0) || !(t.hmax >= t.hmin) || !std::isfinite(t.hmax))
            return fail(ctx, PRR_EINVAL, "adaptive solvers need 0 < hmin <= hmax < inf");
        if (!(t.atol >= 0.0) || !(t.rtol >= 0.0) || !(t.atol + t.rtol > 0.0))
            return fail(ctx, PRR_EINVAL, "adaptive solvers need atol, rtol >= 0, not both 0");
    }
    if (t.solver != PRR_EULER) {
        // every ray ends within max_attempts loop trips: the persistent kernel's exit bound
        if (t.max_attempts < 0 || t.max_attempts > 100000000LL)
            return fail(ctx, PRR_EINVAL, "max_attempts must be in [0, 1e8]");
    } else {
        // Euler: steps are bounded by max_steps and lambda_max / h
        const double by_lambda = t.lambda_max / t.h;
        if (std::min<double>(double(t.max_steps), by_lambda) > 1e8)
            return fail(ctx, PRR_EINVAL, "Euler needs min(max_steps, lambda_max/h) <= 1e8");
    }
    if (t.max_steps < 0) return fail(ctx, PRR_EINVAL, "max_steps must be >= 0");
    if (!(t.r_escape > 0.0)) return fail(ctx, PRR_EINVAL, "r_escape must be positive");
    return PRR_OK;
}


prr_status validate_frame(prr_ctx *ctx, const prr_frame &f, int idx) {
    if (f.width <= 0 || f.height <= 0 || f.width > 65536 || f.height > 65536)
        return fail(ctx, PRR_EINVAL, "frame %d: bad image size %dx%d", idx, f.width, f.height);
    if (f.ssaa == PRR_SSAA_NONE) {
        if (f.samples != 1)
            return fail(ctx, PRR_EINVAL, "frame %d: SSAA none needs samples = 1 (got %d)", idx, f.samples);
    } else if (f.ssaa == PRR_SSAA_REGULAR || f.ssaa == PRR_SSAA_JITTERED) {
        const int n = int(std::lround(std::sqrt(double(f.samples))));
        if (f.samples < 1 || n * n != f.samples || n > 16)
            return fail(ctx, PRR_EINVAL, "frame %d: %s SSAA needs samples = n*n, n <= 16 (got %d)", idx,
                        f.ssaa == PRR_SSAA_REGULAR ? "regular" : "jittered", f.samples);
    } else if (f.ssaa == PRR_SSAA_RANDOM) {
        if (f.samples < 1 || f.samples > 256)
            return fail(ctx, PRR_EINVAL, "frame %d: random SSAA needs 1 <= samples <= 256 (got %d)", idx, f.samples);
    } else {
        return fail(ctx, PRR_EINVAL, "frame %d: unsupported SSAA distribution %d", idx, f.ssaa);
    }
    if (f.projection != PRR_PROJ_PINHOLE && f.projection != PRR_PROJ_SPHERICAL)
        return fail(ctx, PRR_EINVAL, "frame %d: unknown projection %d", idx, f.projection);
    if (!(f.vfov > 0.0 && f.vfov < M_PI)) return fail(ctx, PRR_EINVAL, "frame %d: vfov must be in (0, pi)", idx);
    if (f.n_spheres < 0 || f.n_spheres > PRR_MAX_SPHERES)
        return fail(ctx, PRR_EINVAL, "frame %d: n_spheres must be in [0, %d]", idx, PRR_MAX_SPHERES);
    for (int k = 0; k < f.n_spheres; ++k)
        if (!(f.spheres[k].radius > 0.0)) return fail(ctx, PRR_EINVAL, "frame %d: sphere %d radius", idx, k);
    if (!(f.mass >= 0.0)) return fail(ctx, PRR_EINVAL, "frame %d: mass must be >= 0", idx);
    if (ctx->metric == PRR_KERR_BL && !(std::fabs(f.spin) <= f.mass))
        return fail(ctx, PRR_EINVAL, "frame %d: Kerr needs |a| <= M", idx);
    const double r = std::sqrt(f.pos[0] * f.pos[0] + f.pos[1] * f.pos[1] + f.pos[2] * f.pos[2]);
    if ((ctx->metric == PRR_SCHW_SPH || ctx->metric == PRR_KERR_BL || ctx->metric == PRR_ELLIS) &&
        !(f.pos[0] * f.pos[0] + f.pos[1] * f.pos[1] > 0.0))
        return fail(ctx, PRR_EINVAL, "frame %d: camera on the polar axis of a spherical chart", idx);
    if (!(r < ctx->tracer.r_escape)) return fail(ctx, PRR_EINVAL, "frame %d: camera outside r_escape", idx);
    if (ctx->metric == PRR_ELLIS) {
        if (!(r > ctx->p[0])) return fail(ctx, PRR_EINVAL, "frame %d: camera inside the wormhole's throat radius", idx);
        if (f.n_spheres != 0) return fail(ctx, PRR_EINVAL, "frame %d: the wormhole scene holds no spheres", idx);
    }
    return PRR_OK;
}


// The metric's parameter that the frame does not carry: the charge of the
// Schwarzschild-coordinate metric, the Ellis throat radius.
double metric_param(const prr_ctx *ctx) { return ctx->metric == PRR_ELLIS ? ctx->p[0] : ctx->q; }
// Escape radius beyond the wormhole's throat: p[1] (the reference's rout2), else r_escape.
// Device memory for the replayed steps of one chunk of adjoint tasks (48 B
// per step, K steps per task); larger replays run in several chunks.
constexpr size_t kRecBudget = size_t(8) << 30;

double far_escape(const prr_ctx *ctx) { return ctx->metric == PRR_ELLIS ? ctx->p[1] : 0.0; }

// Resident workgroups per CU of the persistent grid: the launch shape's, or
// by default as many as the render instantiation's register budget lets
// stay resident (its waves per SIMD x 4 SIMDs).
int blocks_per_cu(const prr_ctx *ctx) {
    if (ctx->launch.blocks_per_cu > 0) return ctx->launch.blocks_per_cu;
    const int waves = integrate_waves(ctx->metric, ctx->tracer.solver);
    return std::max(1, waves * 4 * 64 / std::max(64, ctx->launch.block));
}

// Algorithmic flops of one attempt.  Counting: FMA = 2; add, sub, mul, abs,
// min, max, compare = 1; div, sqrt and each transcendental (sin and cos
// separately) = 1.  F_rhs is counted from the formulas in geodesic.hpp.
// Per state component an RKF45 attempt adds: 6 scalings k = h f, 15 stage
// FMAs (30), the 4th-order update (4 FMAs + 1 add = 9), the error estimate
// (5 nonzero b5 - b4 terms = 9) and the scaled error norm (8); plus 6 for
// the step controller.  The stage part (6 F_rhs + 36 per component) runs in
// FP32 in the FP32-bulk form, the rest always in FP64.  Euler: F_rhs + 2 per
// component.
struct FlopModel {
    double rhs;       // one right-hand side
    int n;            // state components
    double delta_rhs; // one FP32 difference right-hand side (Kerr)
    double prep;      // FP64 difference coefficients (Kerr)
};
FlopModel flop_model(int metric) {
    switch (metric) {
    case PRR_FLAT: return {0.0, 8, 0.0, 0.0};
    case PRR_SCHW_KS_CART: return {49.0, 8, 0.0, 0.0};
    case PRR_SCHW_SPH: return {44.0, 8, 0.0, 0.0};
    default: return {35.0, 6, 43.0, 45.0};
    }
}

// Algorithmic flops per attempt (a division counts as one flop).  RKF45:
// six right-hand sides, the stage sums (15 multiply-adds per component plus
// the h scaling), and the 4th-order / error sums, norm and controller.
// Kerr's FP32-bulk attempt: the FP64 stage 1 and difference coefficients,
// then five FP32 difference stages with their stage sums.
void fill_flops(prr_ctx *ctx, const CountersDev &c) {
    const FlopModel fm = flop_model(ctx->metric);
    const int sv = ctx->tracer.solver;
    const bool rkf = sv != PRR_EULER;
    // stages S, their right-hand sides and stage sums (S (S - 1) / 2 multiply-adds
    // and one scaling per component), then the update / error / norm / controller
    const double S = sv == PRR_DORMAND ? 7.0 : (sv == PRR_RK4 ? 4.0 : 6.0);
    const double f_stage = rkf ? S * fm.rhs + (S * (S - 1.0) + S) * fm.n : fm.rhs;
    const double f_rest = rkf ? (sv == PRR_RK4 ? 9.0 * fm.n : 26.0 * fm.n + 6.0) : 2.0 * fm.n;
    const double a64 = double(c.attempts_fp64), a32 = double(c.attempts) - double(c.attempts_fp64);
    if (rkf && ctx->metric == PRR_KERR_BL) {
        ctx->stats.flops_fp64 = a64 * (f_stage + f_rest) + a32 * (fm.rhs + fm.prep + f_rest + 2.0 * fm.n);
        ctx->stats.flops_fp32 = a32 * ((S - 1.0) * (fm.delta_rhs + 11.6 * fm.n));
        return;
    }
    ctx->stats.flops_fp64 = a64 * (f_stage + f_rest) + a32 * f_rest;
    ctx->stats.flops_fp32 = a32 * f_stage;
}

} // namespace

// ======================================================================== API
extern "C" {

int prr_abi_version(void) { return PRR_ABI_VERSION; }
const char *prr_global_error(void) { return g_error.c_str(); }
const char *prr_last_error(const prr_ctx *ctx) { return ctx ? ctx->err.c_str() : g_error.c_str(); }

prr_status prr_create(prr_ctx **out, int hip_device) try {
    if (!out) return fail(nullptr, PRR_EINVAL, "prr_create: out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(nullptr, PRR_ENODEV, "no HIP device");
    if (hip_device < 0 || hip_device >= n) return fail(nullptr, PRR_EINVAL, "device %d out of range", hip_device);
    hipError_t e = hipSetDevice(hip_device);
    if (e != hipSuccess) return fail(nullptr, PRR_EHIP, "hipSetDevice: %s", hipGetErrorString(e));
    const prr_status ns = noalloc_setup(hip_device);
    if (ns != PRR_OK) return ns;
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, hip_device);
    if (e != hipSuccess) return fail(nullptr, PRR_EHIP, "hipGetDeviceProperties: %s", hipGetErrorString(e));
    prr_ctx *ctx = new prr_ctx();
    ctx->device = hip_device;
    ctx->cus = prop.multiProcessorCount;
    ctx->tracer = prr_tracer{PRR_RKF45, PRR_FP64, 0.01, 1e-4, 0.1, 1e-3, 0.0, 70.0, 0.0, 1e-3, 1000.0,
                             1000000000LL, 1000000LL, 3.0};
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return fail(nullptr, PRR_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    for (auto &ev : ctx->ev) {
        e = hipEventCreate(&ev);
        if (e != hipSuccess) {
            prr_destroy(ctx);
            return fail(nullptr, PRR_EHIP, "hipEventCreate: %s", hipGetErrorString(e));
        }
    }
    *out = ctx;
    return PRR_OK;
} catch (...) {
    return prr::abi_exception(nullptr);
}

namespace {
void free_context(prr_ctx *ctx) {
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &t : ctx->tex) t.reset();
    DevBuf *bufs[] = {&ctx->ws, &ctx->stage_rgbf, &ctx->stage_rgb8, &ctx->cls, &ctx->endpoint, &ctx->steps,
                      &ctx->attempts, &ctx->state8, &ctx->x0u0, &ctx->adj_img, &ctx->adj_dldi,
                      &ctx->adj_off, &ctx->adj_ckpt, &ctx->adj_ckpt2, &ctx->adj_event, &ctx->adj_rays, &ctx->adj_rec,
                      &ctx->adj_task_ray, &ctx->adj_task_seg, &ctx->adj_levelstart, &ctx->adj_nseg,
                      &ctx->adj_over, &ctx->adj_plan, &ctx->adj_lam, &ctx->adj_grad_ray, &ctx->adj_grad,
                      &ctx->adj_jac, &ctx->adj_jacstart, &ctx->adj_recall};
    for (DevBuf *b : bufs) release(*b);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    for (auto &ev : ctx->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->copy_done) (void)hipEventDestroy(ctx->copy_done);
    if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
    delete ctx;
}
} // namespace

// A context with live graphs is released when its last graph is destroyed
// (the graphs replay on its stream and report into its statistics).
void prr_destroy(prr_ctx *ctx) {
    if (!ctx) return;
    if (ctx->live_graphs > 0) {
        ctx->destroyed = true;
        return;
    }
    free_context(ctx);
}

prr_status prr_set_spacetime(prr_ctx *ctx, int32_t metric, double M, double a, double q, const double p[16]) {
    if (!ctx) return fail(nullptr, PRR_EINVAL, "NULL context");
    if (metric < PRR_FLAT || metric > PRR_ELLIS) return fail(ctx, PRR_EINVAL, "unknown metric %d", metric);
    if (metric == PRR_ELLIS && !(p && p[0] > 0.0 && std::isfinite(p[0])))
        return fail(ctx, PRR_EINVAL, "the Ellis wormhole needs a throat radius p[0] > 0");
    if (metric == PRR_ELLIS && !(p[1] >= 0.0 && std::isfinite(p[1])))
        return fail(ctx, PRR_EINVAL, "the wormhole's far-side escape radius p[1] must be >= 0 (0: r_escape)");
    if (!(M >= 0.0)) return fail(ctx, PRR_EINVAL, "mass must be >= 0");
    if (metric == PRR_KERR_BL && !(std::fabs(a) <= M)) return fail(ctx, PRR_EINVAL, "Kerr needs |a| <= M");
    if (q != 0.0 && metric != PRR_SCHW_SPH)
        return fail(ctx, PRR_EINVAL, "a charge needs the Schwarzschild-coordinate metric (Reissner-Nordstrom)");
    if (!(std::fabs(q) <= M) && q != 0.0) return fail(ctx, PRR_EINVAL, "Reissner-Nordstrom needs |q| <= M");
    ctx->metric = metric;
    ctx->M = M;
    ctx->a = a;
    ctx->q = q;
    for (int i = 0; i < 16; ++i) ctx->p[i] = p ? p[i] : 0.0;
    return PRR_OK;
}

prr_status prr_set_tracer(prr_ctx *ctx, const prr_tracer *t) {
    if (!ctx || !t) return fail(ctx, PRR_EINVAL, "NULL argument");
    prr_status st = validate_tracer(ctx, *t);
    if (st != PRR_OK) return st;
    ctx->tracer = *t;
    return PRR_OK;
}

prr_status prr_set_launch(prr_ctx *ctx, const prr_launch *l) {
    if (!ctx || !l) return fail(ctx, PRR_EINVAL, "NULL argument");
    if (l->block < 64 || l->block > 256 || l->block % 64)
        return fail(ctx, PRR_EINVAL, "block must be 64, 128, 192 or 256");
    if (l->blocks_per_cu < 0 || l->blocks_per_cu > 32) return fail(ctx, PRR_EINVAL, "blocks_per_cu in [0, 32]");
    if (l->refill < 1 || l->refill > 64) return fail(ctx, PRR_EINVAL, "refill in [1, 64]");
    if (l->checkpoint < 1) return fail(ctx, PRR_EINVAL, "checkpoint must be >= 1");
    if (l->adjoint_record_bytes < 0) return fail(ctx, PRR_EINVAL, "adjoint_record_bytes must be >= 0");
    ctx->launch = *l;
    return PRR_OK;
}

prr_status prr_get_launch(const prr_ctx *ctx, prr_launch *l) {
    if (!ctx || !l) return fail(nullptr, PRR_EINVAL, "NULL argument");
    *l = ctx->launch;
    return PRR_OK;
}

prr_status prr_set_texture(prr_ctx *ctx, int slot, const uint8_t *rgb8, int w, int h, int32_t filter) try {
    if (!ctx) return fail(nullptr, PRR_EINVAL, "NULL context");
    if (slot < 0 || slot >= PRR_MAX_TEXTURES) return fail(ctx, PRR_EINVAL, "texture slot %d out of range", slot);
    if (filter != PRR_FILTER_SQAVG && filter != PRR_FILTER_BILINEAR)
        return fail(ctx, PRR_EINVAL, "unknown filter %d", filter);
    if (rgb8 && w > 0 && h > 0 && (w < 2 || h < 2)) return fail(ctx, PRR_EINVAL, "texture must be at least 2x2");
    HIPTRY(ctx, hipSetDevice(ctx->device));
    if (ctx->tex[slot]) {
        // work queued on the context's stream may still read the texels; a
        // graph that captured them keeps its own reference
        HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
        ctx->tex[slot].reset();
    }
    if (!rgb8 || w <= 0 || h <= 0) return PRR_OK; // slot cleared
    std::vector<float4> host(size_t(w) * h);
    for (size_t i = 0; i < host.size(); ++i)
        host[i] = make_float4(float(rgb8[3 * i]) / 255.0f, float(rgb8[3 * i + 1]) / 255.0f,
                              float(rgb8[3 * i + 2]) / 255.0f, 1.0f);
    auto t = std::make_shared<TexAlloc>();
    t->device = ctx->device;
    float4 *d = nullptr;
    hipError_t e = hipMalloc(&d, host.size() * sizeof(float4));
    if (e != hipSuccess) return fail(ctx, PRR_ENOMEM, "texture hipMalloc: %s", hipGetErrorString(e));
    t->t.texels = d;
    e = hipMemcpy(d, host.data(), host.size() * sizeof(float4), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(ctx, PRR_EHIP, "texture upload: %s", hipGetErrorString(e));
    t->t.w = w;
    t->t.h = h;
    t->t.filter = filter;
    ctx->tex[slot] = std::move(t);
    return PRR_OK;
} catch (...) {
    return prr::abi_exception(ctx);
}

prr_status prr_set_spheres(prr_ctx *ctx, const prr_sphere *s, int n) try {
    if (!ctx) return fail(nullptr, PRR_EINVAL, "NULL context");
    if (n < 0 || n > PRR_MAX_SPHERES || (n > 0 && !s)) return fail(ctx, PRR_EINVAL, "bad sphere list");
    for (int k = 0; k < n; ++k)
        if (!(s[k].radius > 0.0)) return fail(ctx, PRR_EINVAL, "sphere %d radius must be positive", k);
    ctx->spheres.assign(s, s + n);
    return PRR_OK;
} catch (...) {
    return prr::abi_exception(ctx);
}

prr_status prr_get_stats(const prr_ctx *ctx, prr_stats *out) {
    if (!ctx || !out) return fail(nullptr, PRR_EINVAL, "NULL argument");
    *out = ctx->stats;
    return PRR_OK;
}

} // extern "C"

namespace {

int rs_width(int metric) { return metric == PRR_KERR_BL ? 9 : 8; }

int integrate_grid(const prr_ctx *ctx) { return ctx->cus * blocks_per_cu(ctx); }

} // namespace

// Validates the frames, fills the frame table fd[0..n) and the totals.
prr_status prr::plan_frames(prr_ctx *ctx, const prr_frame *frames, int n, FrameDev *fd, RenderPlan &P) {
    P.n = n;
    P.fd = fd;
    for (int i = 0; i < n; ++i) {
        prr_status st = validate_frame(ctx, frames[i], i);
        if (st != PRR_OK) return st;
        fd[i] = frame_dev(ctx->metric, ctx->tracer, frames[i], P.total, metric_param(ctx), far_escape(ctx));
        const int64_t np = int64_t(frames[i].width) * frames[i].height;
        // each frame's queue starts at a multiple of 64, so that a whole-wave
        // refill never mixes two frames and a frame's waves are the same
        // whether it is rendered alone or in a batch
        fd[i].queue_begin = (P.qtotal + 63) / 64 * 64;
        P.qtotal = fd[i].queue_begin + np * frames[i].samples;
        P.total += np * frames[i].samples;
        P.pixels += np;
        P.max_pix = std::max(P.max_pix, np);
        P.max_w = std::max(P.max_w, frames[i].width);
        P.max_h = std::max(P.max_h, frames[i].height);
        P.any_fxaa = P.any_fxaa || frames[i].fxaa;
        P.any_plain = P.any_plain || !frames[i].fxaa;
        // the frame's ray colours in, the linear image and RGB8 out: k_ssaa
        // for frames without FXAA, k_fxaa (which combines the samples itself)
        // for the others
        (frames[i].fxaa ? P.fxaa_bytes : P.ssaa_bytes) += double(np) * frames[i].samples * 12.0 + double(np) * 15.0;
    }
    P.grid = integrate_grid(ctx);
    P.block = ctx->launch.block;
    return PRR_OK;
}

WorkspaceShape prr::workspace_shape(const prr_ctx *ctx, const RenderPlan &P) {
    WorkspaceShape sh;
    sh.nframes = P.n;
    sh.rays = P.total;
    sh.state_width = rs_width(ctx->metric);
    sh.lanes = int64_t(P.grid) * P.block;
    return sh;
}

// The context's own workspace (eager calls), grown to the plan's shape.
prr_status prr::context_workspace(prr_ctx *ctx, const RenderPlan &P, Workspace &ws) {
    const WorkspaceShape sh = workspace_shape(ctx, P);
    prr_status st = ensure(ctx, ctx->ws, carve_workspace(nullptr, sh, nullptr));
    if (st != PRR_OK) return st;
    carve_workspace(ctx->ws.p, sh, &ws);
    return PRR_OK;
}

// Post-processing descriptors pf[0..n): the frame's ray colours and FXAA
// input in the workspace; outputs to device images in place, to host images
// through the staging buffers (stage_rgbf / stage_rgb8, W*H of every frame).
void prr::plan_post(const prr_frame *frames, const prr_image *images, const Workspace &ws, float *stage_rgbf,
               uint8_t *stage_rgb8, RenderPlan &P, PostFrameDev *pf) {
    P.pf = pf;
    int64_t pix_off = 0;
    for (int i = 0; i < P.n; ++i) {
        PostFrameDev &q = pf[i];
        const int64_t np = int64_t(frames[i].width) * frames[i].height;
        q.ray_color = ws.color + 3 * P.fd[i].ray_begin;
        q.W = frames[i].width;
        q.H = frames[i].height;
        q.s = frames[i].samples;
        q.fxaa = frames[i].fxaa;
        const prr_image *im = images ? &images[i] : nullptr;
        if (im && im->on_device) {
            q.final_rgbf = im->rgbf;
            q.rgb8 = im->rgb8;
        } else if (im) {
            q.final_rgbf = im->rgbf ? stage_rgbf + 3 * pix_off : nullptr;
            q.rgb8 = im->rgb8 ? stage_rgb8 + 3 * pix_off : nullptr;
        } else {
            q.final_rgbf = nullptr;
            q.rgb8 = nullptr;
        }
        pix_off += np;
    }
}

JobDev prr::make_job(const prr_ctx *ctx, const RenderPlan &P, const Workspace &ws) {
    JobDev job{};
    job.frames = ws.frames;
    job.nframes = P.n;
    // FP32-bulk always refills whole waves: its choice of stage precision is
    // made per wave (k_integrate), and whole-wave refills keep it independent
    // of the launch shape
    job.refill = ctx->tracer.precision == PRR_FP32_BULK ? 64 : ctx->launch.refill;
    job.total = P.total;
    job.qtotal = P.qtotal > 0 ? P.qtotal : P.total;
    job.explicit_x0u0 = nullptr;
    job.tr = tracer_dev(ctx->tracer);
    for (int i = 0; i < PRR_MAX_TEXTURES; ++i) job.tex[i] = ctx->tex[i] ? ctx->tex[i]->t : TexDev{};
    job.out.color = ws.color;
    job.rs.y = ws.rs_y;
    job.rs.cls = ws.rs_cls;
    job.rs.hit = ws.rs_hit;
    job.rs.steps = ws.rs_steps;
    job.rs.attempts = ws.rs_attempts;
    job.rs_total = P.total;
    job.loc = ws.loc;
    job.ctr = ws.counters;
    return job;
}

// Enqueue reset -> init -> integrate -> shade -> SSAA -> FXAA, with timing events.
// job.adj.mode picks the integrator binary: the render one, or (the adjoint's
// first pass) the checkpointing one, whose steps the replay must reproduce.
prr_status prr::enqueue_render(prr_ctx *ctx, const RenderPlan &P, const JobDev &job, const Workspace &ws,
                          hipStream_t st, bool events) {
    if (events) HIPTRY(ctx, hipEventRecord(ctx->ev[0], st));
    HIPTRY(ctx, launch_reset(job.ctr, st));
    HIPTRY(ctx, launch_init(ctx->metric, job, st));
    if (events) HIPTRY(ctx, hipEventRecord(ctx->ev[1], st));
    HIPTRY(ctx, launch_integrate(ctx->metric, job, P.grid, P.block, st));
    if (events) HIPTRY(ctx, hipEventRecord(ctx->ev[2], st));
    HIPTRY(ctx, launch_shade(ctx->metric, job, st));
    if (events) HIPTRY(ctx, hipEventRecord(ctx->ev[5], st));
    if (P.post && P.any_plain) HIPTRY(ctx, launch_ssaa(ws.post, P.n, P.max_pix, st));
    if (events) HIPTRY(ctx, hipEventRecord(ctx->ev[3], st));
    if (P.post && P.any_fxaa) HIPTRY(ctx, launch_fxaa(ws.post, P.n, P.max_w, P.max_h, st));
    if (events) HIPTRY(ctx, hipEventRecord(ctx->ev[4], st));
    return PRR_OK;
}

prr_status prr::read_stats(prr_ctx *ctx, const RenderPlan &P, const CountersDev *counters, hipStream_t st, bool events) {
    CountersDev c{};
    HIPTRY(ctx, hipMemcpyAsync(&c, counters, sizeof c, hipMemcpyDeviceToHost, st));
    HIPTRY(ctx, hipStreamSynchronize(st));
    return stats_from(ctx, P, c, events);
}

prr_status prr::stats_from(prr_ctx *ctx, const RenderPlan &P, const CountersDev &c, bool events) {
    prr_stats &s = ctx->stats;
    s = prr_stats{};
    s.rays = int64_t(c.rays);
    s.attempts = int64_t(c.attempts);
    s.accepted = int64_t(c.accepted);
    s.attempts_fp64 = int64_t(c.attempts_fp64);
    s.pixels = P.pixels;
    if (events) {
        float ms = 0.f;
        HIPTRY(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        s.init_ms = ms;
        HIPTRY(ctx, hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]));
        s.integrate_ms = ms;
        HIPTRY(ctx, hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[5]));
        s.shade_ms = ms;
        HIPTRY(ctx, hipEventElapsedTime(&ms, ctx->ev[5], ctx->ev[3]));
        s.ssaa_ms = ms;
        HIPTRY(ctx, hipEventElapsedTime(&ms, ctx->ev[3], ctx->ev[4]));
        s.fxaa_ms = ms;
        HIPTRY(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[4]));
        s.total_ms = ms;
    }
    s.ssaa_bytes = P.ssaa_bytes;
    s.fxaa_bytes = P.fxaa_bytes;
    fill_flops(ctx, c);
    if (c.rays != uint64_t(P.total))
        return fail(ctx, PRR_EHIP, "integrator processed %llu rays, expected %lld", c.rays, (long long)P.total);
    return PRR_OK;
}

extern "C" {

prr_status prr_render(prr_ctx *ctx, const prr_frame *frames, int n, prr_image *images, prr_ray_out *rays,
                      void *stream) try {
    if (!ctx) return fail(nullptr, PRR_EINVAL, "NULL context");
    if (!frames || n <= 0) return fail(ctx, PRR_EINVAL, "no frames");
    HIPTRY(ctx, hipSetDevice(ctx->device));
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    bool host_images = false;
    if (images)
        for (int i = 0; i < n; ++i) host_images = host_images || !images[i].on_device;
    std::vector<FrameDev> fd(static_cast<size_t>(n));
    std::vector<PostFrameDev> pf(static_cast<size_t>(n));
    RenderPlan P;
    prr_status s = plan_frames(ctx, frames, n, fd.data(), P);
    if (s != PRR_OK) return s;
    Workspace ws;
    if ((s = context_workspace(ctx, P, ws)) != PRR_OK) return s;
    if (host_images) {
        if ((s = ensure(ctx, ctx->stage_rgbf, sizeof(float) * 3 * size_t(P.pixels))) != PRR_OK) return s;
        if ((s = ensure(ctx, ctx->stage_rgb8, 3 * size_t(P.pixels))) != PRR_OK) return s;
    }
    plan_post(frames, images, ws, static_cast<float *>(ctx->stage_rgbf.p), static_cast<uint8_t *>(ctx->stage_rgb8.p),
              P, pf.data());
    JobDev job = make_job(ctx, P, ws);
    // optional per-ray diagnostics
    if (rays) {
        const size_t T = size_t(P.total);
        auto bind = [&](void *user, DevBuf &buf, size_t bytes, auto *&dst) -> prr_status {
            using Ptr = std::remove_reference_t<decltype(dst)>;
            if (!user) {
                dst = nullptr;
                return PRR_OK;
            }
            if (rays->on_device) {
                dst = static_cast<Ptr>(user);
                return PRR_OK;
            }
            prr_status e = ensure(ctx, buf, bytes);
            dst = static_cast<Ptr>(buf.p);
            return e;
        };
        if ((s = bind(rays->cls, ctx->cls, T, job.out.cls)) != PRR_OK) return s;
        if ((s = bind(rays->endpoint, ctx->endpoint, 24 * T, job.out.endpoint)) != PRR_OK) return s;
        if ((s = bind(rays->steps, ctx->steps, 4 * T, job.out.steps)) != PRR_OK) return s;
        if ((s = bind(rays->attempts, ctx->attempts, 4 * T, job.out.attempts)) != PRR_OK) return s;
    }
    HIPTRY(ctx, hipMemcpyAsync(ws.frames, fd.data(), sizeof(FrameDev) * n, hipMemcpyHostToDevice, st));
    HIPTRY(ctx, hipMemcpyAsync(ws.post, pf.data(), sizeof(PostFrameDev) * n, hipMemcpyHostToDevice, st));
    if ((s = enqueue_render(ctx, P, job, ws, st, true)) != PRR_OK) return s;
    if ((s = read_stats(ctx, P, ws.counters, st, true)) != PRR_OK) return s;
    // copy results to host buffers
    if (images && host_images) {
        int64_t off = 0;
        for (int i = 0; i < n; ++i) {
            const int64_t np = int64_t(frames[i].width) * frames[i].height;
            if (!images[i].on_device) {
                if (images[i].rgbf)
                    HIPTRY(ctx, hipMemcpyAsync(images[i].rgbf, static_cast<float *>(ctx->stage_rgbf.p) + 3 * off,
                                               12 * np, hipMemcpyDeviceToHost, st));
                if (images[i].rgb8)
                    HIPTRY(ctx, hipMemcpyAsync(images[i].rgb8, static_cast<uint8_t *>(ctx->stage_rgb8.p) + 3 * off,
                                               3 * np, hipMemcpyDeviceToHost, st));
            }
            off += np;
        }
    }
    if (rays) {
        const size_t T = size_t(P.total);
        if (rays->color) {
            HIPTRY(ctx, hipMemcpyAsync(rays->color, ws.color, 12 * T,
                                       rays->on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
        }
        if (!rays->on_device) {
            if (rays->cls) HIPTRY(ctx, hipMemcpyAsync(rays->cls, ctx->cls.p, T, hipMemcpyDeviceToHost, st));
            if (rays->endpoint)
                HIPTRY(ctx, hipMemcpyAsync(rays->endpoint, ctx->endpoint.p, 24 * T, hipMemcpyDeviceToHost, st));
            if (rays->steps) HIPTRY(ctx, hipMemcpyAsync(rays->steps, ctx->steps.p, 4 * T, hipMemcpyDeviceToHost, st));
            if (rays->attempts)
                HIPTRY(ctx, hipMemcpyAsync(rays->attempts, ctx->attempts.p, 4 * T, hipMemcpyDeviceToHost, st));
        }
    }
    HIPTRY(ctx, hipStreamSynchronize(st));
    return PRR_OK;
} catch (...) {
    return prr::abi_exception(ctx);
}

prr_status prr_trace(prr_ctx *ctx, const double *x0u0, int n, double *xend, int32_t *cls, int32_t *steps,
                     int32_t *attempts) try {
    if (!ctx) return fail(nullptr, PRR_EINVAL, "NULL context");
    if (!x0u0 || n <= 0) return fail(ctx, PRR_EINVAL, "no initial states");
    HIPTRY(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    prr_frame f{};
    f.width = f.height = 1;
    f.samples = 1;
    f.mass = ctx->M;
    f.spin = ctx->a;
    f.n_spheres = int(ctx->spheres.size());
    for (int k = 0; k < f.n_spheres; ++k) f.spheres[k] = ctx->spheres[k];
    FrameDev fd = frame_dev(ctx->metric, ctx->tracer, f, 0, metric_param(ctx), far_escape(ctx));
    RenderPlan P;
    P.n = 1;
    P.fd = &fd;
    P.post = false;
    P.total = P.qtotal = n;
    P.grid = integrate_grid(ctx);
    P.block = ctx->launch.block;
    prr_status s;
    Workspace ws;
    if ((s = context_workspace(ctx, P, ws)) != PRR_OK) return s;
    if ((s = ensure(ctx, ctx->x0u0, 64 * size_t(n))) != PRR_OK) return s;
    if ((s = ensure(ctx, ctx->state8, 64 * size_t(n))) != PRR_OK) return s;
    if ((s = ensure(ctx, ctx->steps, 4 * size_t(n))) != PRR_OK) return s;
    if ((s = ensure(ctx, ctx->attempts, 4 * size_t(n))) != PRR_OK) return s;
    if ((s = ensure(ctx, ctx->cls, size_t(n))) != PRR_OK) return s;
    JobDev job = make_job(ctx, P, ws);
    job.explicit_x0u0 = static_cast<const double *>(ctx->x0u0.p);
    job.out = OutDev{};
    job.out.state8 = static_cast<double *>(ctx->state8.p);
    job.out.steps = static_cast<int32_t *>(ctx->steps.p);
    job.out.attempts = static_cast<int32_t *>(ctx->attempts.p);
    job.out.cls = static_cast<int8_t *>(ctx->cls.p);
    HIPTRY(ctx, hipMemcpyAsync(ws.frames, &fd, sizeof(FrameDev), hipMemcpyHostToDevice, st));
    HIPTRY(ctx, hipMemcpyAsync(ctx->x0u0.p, x0u0, 64 * size_t(n), hipMemcpyHostToDevice, st));
    if ((s = enqueue_render(ctx, P, job, ws, st, true)) != PRR_OK) return s;
    std::vector<int8_t> c8(n);
    if (xend) HIPTRY(ctx, hipMemcpyAsync(xend, ctx->state8.p, 64 * size_t(n), hipMemcpyDeviceToHost, st));
    if (steps) HIPTRY(ctx, hipMemcpyAsync(steps, ctx->steps.p, 4 * size_t(n), hipMemcpyDeviceToHost, st));
    if (attempts) HIPTRY(ctx, hipMemcpyAsync(attempts, ctx->attempts.p, 4 * size_t(n), hipMemcpyDeviceToHost, st));
    HIPTRY(ctx, hipMemcpyAsync(c8.data(), ctx->cls.p, size_t(n), hipMemcpyDeviceToHost, st));
    if ((s = read_stats(ctx, P, ws.counters, st, true)) != PRR_OK) return s;
    if (cls)
        for (int i = 0; i < n; ++i) cls[i] = c8[i];
    return PRR_OK;
} catch (...) {
    return prr::abi_exception(ctx);
}

// ---------------------------------------------------------------- graphs
prr_status prr_graph_create(prr_ctx *ctx, const prr_frame *frames, int n, const prr_image *images, prr_graph **out) try {
    if (!ctx || !out) return fail(ctx, PRR_EINVAL, "NULL argument");
    *out = nullptr;
    if (!frames || n <= 0) return fail(ctx, PRR_EINVAL, "no frames");
    if (images)
        for (int i = 0; i < n; ++i)
            if (!images[i].on_device) return fail(ctx, PRR_EINVAL, "graph images must be device buffers");
    HIPTRY(ctx, hipSetDevice(ctx->device));
    std::vector<FrameDev> fd(static_cast<size_t>(n));
    std::vector<PostFrameDev> pf(static_cast<size_t>(n));
    RenderPlan P;
    prr_status s = plan_frames(ctx, frames, n, fd.data(), P);
    if (s != PRR_OK) return s;
    // the graph's own workspace: later eager calls on the context cannot
    // overwrite its tables or move its buffers
    const WorkspaceShape sh = workspace_shape(ctx, P);
    const size_t bytes = carve_workspace(nullptr, sh, nullptr);
    std::unique_ptr<prr_graph, void (*)(prr_graph *)> g(new prr_graph(), prr_graph_destroy);
    hipError_t e = hipMalloc(&g->block, bytes);
    if (e != hipSuccess) return fail(ctx, PRR_ENOMEM, "graph workspace hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    Workspace ws;
    carve_workspace(g->block, sh, &ws);
    plan_post(frames, images, ws, nullptr, nullptr, P, pf.data());
    JobDev job = make_job(ctx, P, ws);
    HIPTRY(ctx, hipMemcpy(ws.frames, fd.data(), sizeof(FrameDev) * n, hipMemcpyHostToDevice));
    HIPTRY(ctx, hipMemcpy(ws.post, pf.data(), sizeof(PostFrameDev) * n, hipMemcpyHostToDevice));
    g->counters = ws.counters;
    g->total_rays = P.total;
    g->pixels = P.pixels;
    g->ssaa_bytes = P.ssaa_bytes;
    g->fxaa_bytes = P.fxaa_bytes;
    for (int i = 0; i < PRR_MAX_TEXTURES; ++i) g->tex[i] = ctx->tex[i]; // the captured texels stay alive
    HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return fail(ctx, PRR_EHIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    s = enqueue_render(ctx, P, job, ws, ctx->stream, false);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(ctx->stream, &graph);
    g->graph = graph;
    if (s != PRR_OK) return s;
    if (e != hipSuccess) return fail(ctx, PRR_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return fail(ctx, PRR_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    g->ctx = ctx;
    ctx->live_graphs++;
    *out = g.release();
    return PRR_OK;
} catch (...) {
    return prr::abi_exception(ctx);
}

prr_status prr_graph_launch(prr_graph *g, void *stream) {
    if (!g || !g->ctx) return fail(nullptr, PRR_EINVAL, "NULL graph");
    prr_ctx *ctx = g->ctx;
    if (ctx->destroyed) return fail(nullptr, PRR_EINVAL, "the graph's context was destroyed");
    HIPTRY(ctx, hipSetDevice(ctx->device));
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    HIPTRY(ctx, hipEventRecord(ctx->ev[0], st));
    HIPTRY(ctx, hipGraphLaunch(g->exec, st));
    HIPTRY(ctx, hipEventRecord(ctx->ev[4], st));
    CountersDev c{};
    HIPTRY(ctx, hipMemcpyAsync(&c, g->counters, sizeof c, hipMemcpyDeviceToHost, st));
    HIPTRY(ctx, hipStreamSynchronize(st));
    prr_stats &s = ctx->stats;
    s = prr_stats{};
    float ms = 0.f;
    HIPTRY(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[4]));
    s.total_ms = ms;
    s.rays = int64_t(c.rays);
    s.attempts = int64_t(c.attempts);
    s.accepted = int64_t(c.accepted);
    s.attempts_fp64 = int64_t(c.attempts_fp64);
    s.pixels = g->pixels;
    s.ssaa_bytes = g->ssaa_bytes;
    s.fxaa_bytes = g->fxaa_bytes;
    fill_flops(ctx, c);
    if (c.rays != uint64_t(g->total_rays))
        return fail(ctx, PRR_EHIP, "graph processed %llu rays, expected %lld", c.rays, (long long)g->total_rays);
    return PRR_OK;
}

void prr_graph_destroy(prr_graph *g) {
    if (!g) return;
    prr_ctx *ctx = g->ctx;
    if (ctx) {
        (void)hipSetDevice(ctx->device);
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    }
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->block) (void)hipFree(g->block);
    for (auto &t : g->tex) t.reset();
    delete g;
    if (ctx && --ctx->live_graphs == 0 && ctx->destroyed) free_context(ctx);
}

// ---------------------------------------------------------------- adjoint
prr_status prr_frame_from_orbit(const prr_frame *tmpl, const prr_orbit *o, prr_frame *out) {
    if (!tmpl || !o || !out) return fail(nullptr, PRR_EINVAL, "NULL argument");
    prr_frame f = *tmpl;
    const double si = std::sin(o->incl), ci = std::cos(o->incl);
    f.pos[0] = o->r_cam * si * std::cos(o->cam_phi);
    f.pos[1] = o->r_cam * si * std::sin(o->cam_phi);
    f.pos[2] = o->r_cam * ci;
    f.alpha = o->cam_phi + M_PI;
    f.beta = M_PI / 2.0 - o->incl;
    f.spin = o->spin;
    if (o->has_sphere) {
        if (f.n_spheres < 1) f.n_spheres = 1;
        f.spheres[0].c[0] = o->sphere_r * std::cos(o->sphere_phi);
        f.spheres[0].c[1] = o->sphere_r * std::sin(o->sphere_phi);
        f.spheres[0].c[2] = 0.0;
        f.spheres[0].radius = o->sphere_radius;
        f.spheres[0].tex = o->sphere_tex;
    }
    *out = f;
    return PRR_OK;
}

prr_status prr_adjoint(prr_ctx *ctx, const prr_frame *tmpl, const prr_orbit *orbit, const float *dL_dimg,
                       int32_t on_device, double grad[4], float *img_out) try {
    if (!ctx) return fail(nullptr, PRR_EINVAL, "NULL context");
    if (!tmpl || !orbit || !dL_dimg || !grad) return fail(ctx, PRR_EINVAL, "NULL argument");
    if (ctx->metric != PRR_KERR_BL) return fail(ctx, PRR_EINVAL, "the adjoint supports the Kerr metric only");
    if (ctx->tracer.solver != PRR_RKF45) return fail(ctx, PRR_EINVAL, "the adjoint needs the RKF45 solver");
    if (tmpl->fxaa) return fail(ctx, PRR_EINVAL, "the adjoint differentiates the SSAA image: set fxaa = 0");
    if (orbit->has_sphere && tmpl->n_spheres > 1)
        return fail(ctx, PRR_EINVAL, "the adjoint's orbit sphere must be the frame's only sphere");
    const int K = ctx->launch.checkpoint;
    if (K <= 0) return fail(ctx, PRR_EINVAL, "checkpoint interval must be positive");
    HIPTRY(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    prr_frame f;
    prr_status s = prr_frame_from_orbit(tmpl, orbit, &f);
    if (s != PRR_OK) return s;
    const int64_t npix = int64_t(f.width) * f.height;
    if ((s = ensure(ctx, ctx->adj_img, 12 * size_t(npix))) != PRR_OK) return s;
    prr_image im{nullptr, static_cast<float *>(ctx->adj_img.p), 1, 0};
    FrameDev fd;
    PostFrameDev pf;
    RenderPlan P;
    if ((s = plan_frames(ctx, &f, 1, &fd, P)) != PRR_OK) return s;
    Workspace ws;
    if ((s = context_workspace(ctx, P, ws)) != PRR_OK) return s;
    plan_post(&f, &im, ws, nullptr, nullptr, P, &pf);
    const int64_t T = P.total;
    if ((s = ensure(ctx, ctx->steps, 4 * size_t(T))) != PRR_OK) return s;
    if ((s = ensure(ctx, ctx->cls, size_t(T))) != PRR_OK) return s;
    if ((s = ensure(ctx, ctx->adj_off, 8 * size_t(T))) != PRR_OK) return s;
    if ((s = ensure(ctx, ctx->adj_rays, 8 * size_t(T))) != PRR_OK) return s;
    if ((s = ensure(ctx, ctx->adj_event, sizeof(EventDev) * size_t(T))) != PRR_OK) return s;
    if ((s = ensure(ctx, ctx->adj_lam, 80 * size_t(T))) != PRR_OK) return s;
    if ((s = ensure(ctx, ctx->adj_grad_ray, 32 * size_t(T))) != PRR_OK) return s;
    if ((s = ensure(ctx, ctx->adj_grad, 32 + 32 * 256)) != PRR_OK) return s; // the gradient, then partial sums
    // first pass: the render, which also stores each ray's first `cap`
    // checkpoints (every K accepted steps) in fixed slots; rays with more
    // segments are checkpointed again by a pass of their own
    const int cap = std::max(8, (512 + K - 1) / K);
    if ((s = ensure(ctx, ctx->adj_ckpt, 64 * size_t(cap) * size_t(T))) != PRR_OK) return s;
    JobDev job = make_job(ctx, P, ws);
    job.out.steps = static_cast<int32_t *>(ctx->steps.p);
    job.out.cls = static_cast<int8_t *>(ctx->cls.p);
    job.adj.mode = ADJ_CHECKPOINT;
    job.adj.K = K;
    job.adj.cap = cap;
    job.adj.ckpt = static_cast<double *>(ctx->adj_ckpt.p);
    // when every ray's first cap * K steps fit the record budget, the first
    // pass also records each accepted step and the terminal event, and the
    // replay is skipped for calls whose rays all fit (no ray beyond cap
    // segments); otherwise the replay records the steps chunk by chunk
    const uint64_t rec_budget = ctx->launch.adjoint_record_bytes > 0 ? uint64_t(ctx->launch.adjoint_record_bytes)
                                                                       : uint64_t(kRecBudget);
    const int64_t rec_cap = int64_t(cap) * K;
    if (48 * uint64_t(rec_cap) * uint64_t(T) <= rec_budget &&
        grow_quietly(ctx->adj_recall, 48 * size_t(rec_cap) * size_t(T))) { // else: replay
        job.adj.rec_all = static_cast<double *>(ctx->adj_recall.p);
        job.adj.rec_cap = rec_cap;
        job.adj.event = static_cast<EventDev *>(ctx->adj_event.p);
    }
    CountersDev *ctr = ws.counters;
    HIPTRY(ctx, hipMemcpyAsync(ws.frames, &fd, sizeof(FrameDev), hipMemcpyHostToDevice, st));
    HIPTRY(ctx, hipMemcpyAsync(ws.post, &pf, sizeof(PostFrameDev), hipMemcpyHostToDevice, st));
    HIPTRY(ctx, hipEventRecord(ctx->ev[0], st));
    if ((s = enqueue_render(ctx, P, job, ws, st, false)) != PRR_OK) return s;
    // dL/dI from the host: copied on a second stream while the first pass
    // runs; the reverse sweep waits for it
    const float *dldi = dL_dimg;
    if (!on_device) {
        if (!ctx->copy_stream) {
            HIPTRY(ctx, hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
            HIPTRY(ctx, hipEventCreateWithFlags(&ctx->copy_done, hipEventDisableTiming));
        }
        if ((s = ensure(ctx, ctx->adj_dldi, 12 * size_t(npix))) != PRR_OK) return s;
        HIPTRY(ctx, hipMemcpyAsync(ctx->adj_dldi.p, dL_dimg, 12 * size_t(npix), hipMemcpyHostToDevice,
                                   ctx->copy_stream));
        HIPTRY(ctx, hipEventRecord(ctx->copy_done, ctx->copy_stream));
        dldi = static_cast<const float *>(ctx->adj_dldi.p);
    }
    // segments per ray, checkpoint offsets and the rays by decreasing segment
    // count (so that the rays with more than j segments are a prefix of the
    // order): on the device when every segment count fits kAdjPlanBins
    // histogram bins, else on the host (a stable counting sort).  A ray keeps
    // its first-pass checkpoint slots i * cap when it has at most cap
    // segments; else its checkpoints go to the overflow area at -1 - off[i].
    const int64_t seg_bound = (ctx->tracer.max_steps - 1) / K + 2; // exceeds every segment count
    const bool plan_dev = seg_bound <= kAdjPlanBins;
    // length classes of the last segment per segment count in the device plan's bins
    const int32_t L = int32_t(std::min<int64_t>(8, std::max<int64_t>(1, kAdjPlanBins / seg_bound)));
    std::vector<unsigned long long> hist; // rays per segment count 0..nseg_max
    std::vector<unsigned long long> histb; // rays per bin of the order (device plan: (count, class))
    int64_t nck2 = 0, swept = 0, n_over = 0;
    int32_t nseg_max = 0;
    if ((s = ensure(ctx, ctx->adj_nseg, 4 * size_t(T))) != PRR_OK) return s;
    if ((s = ensure(ctx, ctx->adj_over, 8 * size_t(T))) != PRR_OK) return s;
    // plan buffer: histogram, counters, then the bin starts and level starts (one copy from the host)
    if ((s = ensure(ctx, ctx->adj_plan, 8 * size_t(3 * kAdjPlanBins + 4))) != PRR_OK) return s;
    // pinned staging: the histogram and plan counters, the first pass's
    // counters, the bin and level starts, then the final counters and gradient
    struct Pinned {
        unsigned long long hist[kAdjPlanBins + 4];
        CountersDev fwd;
        unsigned long long cursor[2 * kAdjPlanBins];
        CountersDev fin;
        double grad[4];
        int64_t jacstart[kAdjPlanBins + 1];
    };
    if ((s = ensure_pinned(ctx, sizeof(Pinned))) != PRR_OK) return s;
    Pinned *pin = static_cast<Pinned *>(ctx->pinned);
    AdjPlanDev plan{};
    plan.steps = static_cast<const int32_t *>(ctx->steps.p);
    plan.T = T;
    plan.K = K;
    plan.cap = cap;
    plan.nbins = kAdjPlanBins;
    plan.L = L;
    plan.nseg = static_cast<int32_t *>(ctx->adj_nseg.p);
    plan.off = static_cast<int64_t *>(ctx->adj_off.p);
    plan.over = static_cast<int64_t *>(ctx->adj_over.p);
    plan.hist = static_cast<unsigned long long *>(ctx->adj_plan.p);
    plan.ctr = plan.hist + kAdjPlanBins;
    plan.cursor = plan.ctr + 4;
    int64_t *levelstart_d = reinterpret_cast<int64_t *>(plan.cursor + kAdjPlanBins);
    plan.order = static_cast<int64_t *>(ctx->adj_rays.p);
    // host buffers of the copies below live until the final synchronisation
    std::vector<int64_t> order_h, off_h, over_h;
    std::vector<int32_t> steps_h;
    if (plan_dev) {
        HIPTRY(ctx, launch_adj_plan_count(plan, st));
        // the histogram and the four plan counters after it, in one copy
        HIPTRY(ctx, hipMemcpyAsync(pin->hist, plan.hist, sizeof pin->hist, hipMemcpyDeviceToHost, st));
    } else {
        steps_h.resize(size_t(T));
        HIPTRY(ctx, hipMemcpyAsync(steps_h.data(), ctx->steps.p, 4 * size_t(T), hipMemcpyDeviceToHost, st));
    }
    HIPTRY(ctx, hipMemcpyAsync(&pin->fwd, ws.counters, sizeof pin->fwd, hipMemcpyDeviceToHost, st));
    HIPTRY(ctx, hipStreamSynchronize(st));
    if ((s = stats_from(ctx, P, pin->fwd, false)) != PRR_OK) return s;
    prr_stats fwd = ctx->stats;
    if (plan_dev) {
        const unsigned long long *pctr = pin->hist + kAdjPlanBins;
        nck2 = int64_t(pctr[0]);
        n_over = int64_t(pctr[1]);
        swept = int64_t(pctr[2]);
        nseg_max = int32_t(pctr[3]);
        histb.assign(pin->hist, pin->hist + size_t(nseg_max + 1) * L);
        hist.assign(size_t(nseg_max) + 1, 0ull);
        for (size_t b = 0; b < histb.size(); ++b) hist[b / L] += histb[b];
    } else {
        std::vector<int32_t> nseg(static_cast<size_t>(T));
        off_h.resize(size_t(T));
        order_h.resize(size_t(T));
        for (int64_t i = 0; i < T; ++i) {
            nseg[i] = steps_h[i] > 0 ? (steps_h[i] - 1) / K + 1 : 0;
            if (nseg[i] <= cap) {
                off_h[i] = i * cap;
            } else {
                off_h[i] = -1 - nck2;
                nck2 += nseg[i];
                over_h.push_back(i);
            }
            swept += steps_h[i];
            nseg_max = std::max(nseg_max, nseg[i]);
        }
        n_over = int64_t(over_h.size());
        hist.assign(size_t(nseg_max) + 1, 0ull);
        for (int64_t i = 0; i < T; ++i) hist[nseg[i]]++;
        histb = hist; // the host plan orders by segment count only
        std::vector<int64_t> start(size_t(nseg_max) + 2, 0);
        for (int32_t n = nseg_max; n >= 0; --n) start[n] = start[n + 1] + (n + 1 <= nseg_max ? int64_t(hist[n + 1]) : 0);
        for (int64_t i = 0; i < T; ++i) order_h[start[nseg[i]]++] = i;
        HIPTRY(ctx, hipMemcpyAsync(plan.off, off_h.data(), 8 * size_t(T), hipMemcpyHostToDevice, st));
        HIPTRY(ctx, hipMemcpyAsync(plan.order, order_h.data(), 8 * size_t(T), hipMemcpyHostToDevice, st));
        if (n_over > 0)
            HIPTRY(ctx, hipMemcpyAsync(plan.over, over_h.data(), 8 * size_t(n_over), hipMemcpyHostToDevice, st));
    }
    // rays with more than j segments (count[j], a prefix of the order) and
    // the first task of each level j (levels from the last); device plan:
    // the start of each bin in the order (decreasing bins first)
    std::vector<int64_t> count(size_t(nseg_max) + 1, 0), levelstart(size_t(nseg_max) + 1, 0);
    for (int32_t j = nseg_max - 1; j >= 0; --j) count[j] = count[j + 1] + int64_t(hist[j + 1]);
    int64_t ntask = 0;
    for (int32_t j = nseg_max - 1; j >= 0; --j) {
        levelstart[j] = ntask;
        ntask += count[j];
    }
    if (nseg_max < kAdjPlanBins) { // bin starts and level starts in one copy
        unsigned long long *cursor = pin->cursor;
        std::fill(cursor, cursor + 2 * kAdjPlanBins, 0ull);
        for (int64_t b = int64_t(histb.size()) - 2; b >= 0; --b) cursor[b] = cursor[b + 1] + histb[size_t(b) + 1];
        for (int32_t j = 0; j <= nseg_max; ++j) cursor[kAdjPlanBins + j] = uint64_t(levelstart[j]);
        HIPTRY(ctx, hipMemcpyAsync(plan.cursor, cursor, sizeof pin->cursor, hipMemcpyHostToDevice, st));
    } else { // host plan with more levels than the plan buffer holds
        if ((s = ensure(ctx, ctx->adj_levelstart, 8 * levelstart.size())) != PRR_OK) return s;
        levelstart_d = static_cast<int64_t *>(ctx->adj_levelstart.p);
        HIPTRY(ctx, hipMemcpyAsync(levelstart_d, levelstart.data(), 8 * levelstart.size(), hipMemcpyHostToDevice, st));
    }
    if (plan_dev) HIPTRY(ctx, launch_adj_plan_scatter(plan, st));
    HIPTRY(ctx, hipEventRecord(ctx->ev[1], st));
    JobDev j2 = job;
    j2.out = OutDev{};
    j2.adj.rec_all = nullptr;
    j2.adj.rec_cap = 0;
    // the first pass's records serve the reverse sweep when every ray fits them
    const bool direct = job.adj.rec_all != nullptr && n_over == 0;
    j2.adj.nsteps = static_cast<const int32_t *>(ctx->steps.p);
    j2.adj.ckpt_off = static_cast<const int64_t *>(ctx->adj_off.p);
    if (n_over > 0) {
        // the rays with more than cap segments, from the start again, storing all their checkpoints
        if ((s = ensure(ctx, ctx->adj_ckpt2, 64 * size_t(nck2))) != PRR_OK) return s;
        j2.adj.ckpt2 = static_cast<double *>(ctx->adj_ckpt2.p);
        j2.adj.rays = plan.over;
        j2.total = n_over;
        HIPTRY(ctx, launch_reset(ctr, st, true));
        JobDev ji = j2;
        ji.total = T;
        HIPTRY(ctx, launch_init(ctx->metric, ji, st));
        const int grid = int(std::min<int64_t>(P.grid, (j2.total + 63) / 64 + 1));
        HIPTRY(ctx, launch_integrate(ctx->metric, j2, grid, P.block, st));
        j2.total = T;
    }
    HIPTRY(ctx, launch_reset(ctr, st, true));
    // pass 3: replay, then the reverse sweep.  A task is one segment j of one
    // ray; level j holds the rays with more than j segments (a prefix of the
    // order), and levels run from the last.  Levels are replayed in chunks
    // whose records fit kRecBudget (all of them at once when they fit), each
    // chunk in one launch, and each chunk's reverse sweep is one launch over
    // its rays: the replays of different segments run concurrently instead
    // of one launch per segment index.
    if (!direct) { // the replay's task list (the reverse sweep reads the first pass's records directly)
        if ((s = ensure(ctx, ctx->adj_task_ray, 8 * size_t(std::max<int64_t>(ntask, 1)))) != PRR_OK) return s;
        if ((s = ensure(ctx, ctx->adj_task_seg, 4 * size_t(std::max<int64_t>(ntask, 1)))) != PRR_OK) return s;
        plan.levelstart = levelstart_d;
        plan.nlevels = nseg_max;
        plan.ntask = ntask;
        plan.task_ray = static_cast<int64_t *>(ctx->adj_task_ray.p);
        plan.task_seg = static_cast<int32_t *>(ctx->adj_task_seg.p);
        HIPTRY(ctx, launch_adj_plan_tasks(plan, st));
    }
    const size_t rec_task = 48 * size_t(K); // bytes of records per task
    const uint64_t budget = ctx->launch.adjoint_record_bytes > 0 ? uint64_t(ctx->launch.adjoint_record_bytes)
                                                                 : uint64_t(kRecBudget);
    const int64_t chunk_cap = std::max<int64_t>(1, int64_t(budget / rec_task));
    std::vector<std::pair<int32_t, int32_t>> chunks; // (jhi, jlo)
    if (direct) { // the first pass recorded every step: one sweep over all levels
        if (nseg_max > 0) chunks.emplace_back(nseg_max - 1, 0);
    } else {
        for (int32_t j = nseg_max - 1; j >= 0;) {
            int32_t lo = j;
            int64_t n = count[j];
            while (lo > 0 && n + count[lo - 1] <= chunk_cap) n += count[--lo];
            chunks.emplace_back(j, lo);
            j = lo - 1;
        }
        int64_t chunk_max = 0;
        for (const auto &c : chunks)
            chunk_max = std::max(chunk_max, levelstart[c.second] + count[c.second] - levelstart[c.first]);
        if ((s = ensure(ctx, ctx->adj_rec, rec_task * size_t(std::max<int64_t>(chunk_max, 1)))) != PRR_OK) return s;
    }
    JobDev j3 = j2;
    j3.adj.mode = ADJ_REPLAY;
    j3.adj.rec = static_cast<double *>(ctx->adj_rec.p);
    j3.adj.event = static_cast<EventDev *>(ctx->adj_event.p);
    AdjDev ad{};
    // long rays (all levels in one chunk): the reverse sweep takes their last
    // segment, their other segments' transposed Jacobians are formed in
    // parallel, and a chain applies them; level j holds the ranks below
    // min(count[j + 1], nlong)
    if (chunks.size() == 1 && nseg_max - 1 > kAdjLongLast && nseg_max < kAdjPlanBins) {
        ad.nlong = count[kAdjLongLast + 1];
        ad.jlevels = nseg_max - 1;
        pin->jacstart[0] = 0;
        for (int32_t j = 0; j < ad.jlevels; ++j)
            pin->jacstart[j + 1] = pin->jacstart[j] + std::min<int64_t>(count[j + 1], ad.nlong);
        ad.njac = pin->jacstart[ad.jlevels];
        if ((s = ensure(ctx, ctx->adj_jac, 81 * 8 * size_t(std::max<int64_t>(ad.njac, 1)))) != PRR_OK) return s;
        if ((s = ensure(ctx, ctx->adj_jacstart, 8 * size_t(ad.jlevels + 1))) != PRR_OK) return s;
        HIPTRY(ctx, hipMemcpyAsync(ctx->adj_jacstart.p, pin->jacstart, 8 * size_t(ad.jlevels + 1),
                                   hipMemcpyHostToDevice, st));
        ad.jacstart = static_cast<const int64_t *>(ctx->adj_jacstart.p);
        ad.jac = static_cast<double *>(ctx->adj_jac.p);
    }
    ad.K = K;
    ad.rays = static_cast<const int64_t *>(ctx->adj_rays.p);
    ad.levelstart = levelstart_d;
    ad.nsteps = j2.adj.nsteps;
    ad.cls = static_cast<const int8_t *>(ctx->cls.p);
    ad.rec = j3.adj.rec;
    ad.event = j3.adj.event;
    ad.lam = static_cast<double *>(ctx->adj_lam.p);
    ad.dLdI = dldi;
    ad.img = static_cast<const float *>(ctx->adj_img.p);
    ad.ray_color = ws.color;
    ad.grad_ray = static_cast<double *>(ctx->adj_grad_ray.p);
    ad.cam_phi = orbit->cam_phi;
    ad.sphere_r = orbit->sphere_r;
    ad.mismatch = &ctr->mismatch;
    if (!on_device) HIPTRY(ctx, hipStreamWaitEvent(st, ctx->copy_done, 0)); // dL/dI is on the device
    for (const auto &c : chunks) {
        const int64_t t0 = levelstart[c.first], t1 = levelstart[c.second] + count[c.second];
        j3.adj.rec_stride = t1 - t0;
        j3.total = t1 - t0;
        if (direct) {
            ad.rec = job.adj.rec_all;
            ad.rec_cap = job.adj.rec_cap;
        } else {
            j3.adj.rays = static_cast<const int64_t *>(ctx->adj_task_ray.p) + t0;
            j3.adj.seg = static_cast<const int32_t *>(ctx->adj_task_seg.p) + t0;
            HIPTRY(ctx, launch_reset(ctr, st, false));
            const int grid = int(std::min<int64_t>(P.grid, (j3.total + 63) / 64 + 1));
            HIPTRY(ctx, launch_integrate(ctx->metric, j3, grid, P.block, st));
        }
        ad.jhi = c.first;
        ad.jlo = c.second;
        ad.t0 = t0;
        ad.rec_stride = t1 - t0;
        ad.count = count[c.second];
        ad.seed_lo = count[c.first + 1]; // rays whose last segment is in this chunk: nseg - 1 in [jlo, jhi]
        HIPTRY(ctx, launch_adj_reverse(j3, ad, st));
    }
    JobDev jg = job;
    jg.adj = AdjRunDev{};
    jg.total = T;
    HIPTRY(ctx, launch_adj_init_grad(jg, ad, st));
    HIPTRY(ctx, launch_grad_reduce(ad.grad_ray, T, static_cast<double *>(ctx->adj_grad.p) + 4,
                                   static_cast<double *>(ctx->adj_grad.p), st));
    HIPTRY(ctx, hipEventRecord(ctx->ev[2], st));
    HIPTRY(ctx, hipMemcpyAsync(&pin->fin, ws.counters, sizeof pin->fin, hipMemcpyDeviceToHost, st));
    HIPTRY(ctx, hipMemcpyAsync(pin->grad, ctx->adj_grad.p, sizeof pin->grad, hipMemcpyDeviceToHost, st));
    if (img_out)
        HIPTRY(ctx, hipMemcpyAsync(img_out, ctx->adj_img.p, 12 * size_t(npix),
                                   on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
    HIPTRY(ctx, hipStreamSynchronize(st));
    const CountersDev c = pin->fin;
    std::copy(pin->grad, pin->grad + 4, grad);
    float ms_all = 0.f, ms_adj = 0.f;
    HIPTRY(ctx, hipEventElapsedTime(&ms_all, ctx->ev[0], ctx->ev[2]));
    HIPTRY(ctx, hipEventElapsedTime(&ms_adj, ctx->ev[1], ctx->ev[2]));
    ctx->stats = fwd;
    ctx->stats.total_ms = ms_all;
    ctx->stats.adjoint_ms = ms_adj;
    ctx->stats.adjoint_steps = swept;
    if (c.mismatch != 0)
        return fail(ctx, PRR_EHIP, "adjoint replay did not reproduce the forward pass (%llu segments or rays)",
                    c.mismatch);
    return PRR_OK;
} catch (...) {
    return prr::abi_exception(ctx);
}


} // extern "C"
