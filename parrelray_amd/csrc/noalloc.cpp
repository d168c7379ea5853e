// noalloc.cpp -- prr_render_noalloc: the render entry point for callers that
// cannot allocate (code under stream capture, tight serving loops).  It has
// no counterpart in the reference.
//
// Its workspace (frame and post tables, counters, ray state, per-ray
// colours, FXAA input, per-lane event-location contexts) is a static device
// array, `scratch`, of PRR_NOALLOC_SCRATCH_BYTES, which every device the
// library runs on holds until the process ends.  In host code the array's
// name is only a stand-in for the symbol: the device address of each
// device's copy is looked up once per device, when a context is created on
// it (prr_noalloc_setup, called by prr_create), and only that address is used
// afterwards.  Calls that overlap on one device share the array and race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "context.hpp"

using namespace prr;

namespace {

constexpr size_t kScratchBytes = PRR_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_addr[kMaxDevices] = {}; // device address of each device's copy of scratch

void *scratch_base(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_addr[device];
}

} // namespace

// With `device` current: look up its copy of the scratch array (once).
prr_status prr::noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices)
        return fail(nullptr, PRR_EINVAL, "device %d: the allocation-free workspace supports %d devices", device,
                    kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_addr[device]) return PRR_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(nullptr, PRR_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratch_addr[device] = ptr;
    return PRR_OK;
}

extern "C" {

// The frame table and post descriptors are on the stack, the workspace is
// this device's scratch array, and the persistent grid shrinks when its
// per-lane contexts would not fit beside the rays (results do not depend on
// the grid).  The counters and queue head are zeroed by a kernel at the start
// of every call; everything else is written before it is read.
prr_status prr_render_noalloc(prr_ctx *ctx, const prr_frame *frames, int n, const prr_image *images,
                              void *stream) {
    if (!ctx) return fail(nullptr, PRR_EINVAL, "NULL context");
    if (!frames || n <= 0) return fail(ctx, PRR_EINVAL, "no frames");
    if (n > PRR_NOALLOC_MAX_FRAMES)
        return fail(ctx, PRR_EBUDGET, "prr_render_noalloc renders at most %d frames per call (got %d)",
                    PRR_NOALLOC_MAX_FRAMES, n);
    if (images)
        for (int i = 0; i < n; ++i)
            if (!images[i].on_device)
                return fail(ctx, PRR_EINVAL, "prr_render_noalloc writes device images only (frame %d)", i);
    HIPTRY(ctx, hipSetDevice(ctx->device));
    void *base = scratch_base(ctx->device);
    if (!base) return fail(ctx, PRR_EHIP, "no scratch array address for device %d", ctx->device);
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    FrameDev fd[PRR_NOALLOC_MAX_FRAMES];
    PostFrameDev pf[PRR_NOALLOC_MAX_FRAMES];
    RenderPlan P;
    prr_status s = plan_frames(ctx, frames, n, fd, P);
    if (s != PRR_OK) return s;
    // the largest persistent grid (at most the context's launch shape) whose
    // per-lane contexts fit beside the per-ray arrays
    WorkspaceShape sh = workspace_shape(ctx, P);
    sh.lanes = 1;
    const size_t fixed = carve_workspace(nullptr, sh, nullptr);
    const size_t per_block = sizeof(LocateCtx) * size_t(P.block);
    int grid = fixed < kScratchBytes ? int(std::min<size_t>(size_t(P.grid), (kScratchBytes - fixed) / per_block)) : 0;
    for (; grid > 0; --grid) {
        sh.lanes = int64_t(grid) * P.block;
        if (carve_workspace(nullptr, sh, nullptr) <= kScratchBytes) break;
    }
    if (grid <= 0)
        return fail(ctx, PRR_EBUDGET,
                    "prr_render_noalloc: %lld rays of %d frames need more than its %zu bytes of scratch "
                    "(PRR_NOALLOC_SCRATCH_BYTES)",
                    (long long)P.total, n, kScratchBytes);
    P.grid = grid;
    Workspace ws;
    carve_workspace(base, sh, &ws);
    plan_post(frames, images, ws, nullptr, nullptr, P, pf);
    const JobDev job = make_job(ctx, P, ws);
    HIPTRY(ctx, hipMemcpyAsync(ws.frames, fd, sizeof(FrameDev) * size_t(n), hipMemcpyHostToDevice, st));
    HIPTRY(ctx, hipMemcpyAsync(ws.post, pf, sizeof(PostFrameDev) * size_t(n), hipMemcpyHostToDevice, st));
    if ((s = enqueue_render(ctx, P, job, ws, st, true)) != PRR_OK) return s;
    return read_stats(ctx, P, ws.counters, st, true); // synchronises: fd and pf outlive their copies
}

} // extern "C"
