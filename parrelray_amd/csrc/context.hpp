// context.hpp -- the library's context, graph and per-call workspace, and
// the render-planning steps shared by the C-ABI entry points (capi.cpp,
// noalloc.cpp).  Host code only; device structures are in internal.hpp.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "internal.hpp"

namespace prr {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// One texture upload.  The context holds a reference per slot and every
// graph that captured the texels holds its own, so that replacing or clearing
// a slot never frees texels a live graph still reads.
struct TexAlloc {
    int device = 0;
    TexDev t{};
    ~TexAlloc() {
        if (t.texels) {
            (void)hipSetDevice(device);
            (void)hipFree(const_cast<float4 *>(t.texels));
        }
    }
};

// Device memory of one render: the frame table, the post-processing table,
// the counters and queue head, per-ray colours, the FXAA input image, the
// ray-state arrays and one event-location context per integrator lane.  All
// arrays are carved from one block at 256-byte aligned offsets, so that a
// context (grown on demand), a graph (its own block, allocated at capture) and
// the allocation-free entry point (a static device array) lay it out alike.
struct Workspace {
    FrameDev *frames = nullptr;
    PostFrameDev *post = nullptr;
    CountersDev *counters = nullptr;
    float *color = nullptr;       // 3 per ray
    double *rs_y = nullptr;       // rs_width(metric) per ray
    int8_t *rs_cls = nullptr, *rs_hit = nullptr;
    int32_t *rs_steps = nullptr, *rs_attempts = nullptr;
    LocateCtx *loc = nullptr;     // one per lane of the integrator grid
};

// Sizes that fix a workspace's layout.
struct WorkspaceShape {
    int nframes = 0;
    int64_t rays = 0;
    int state_width = 0; // doubles of ray state per ray
    int64_t lanes = 0;
};

// Lays the workspace out from `base` (null: offsets only) and returns its size in bytes.
size_t carve_workspace(void *base, const WorkspaceShape &s, Workspace *ws);

// Frames of one call and their totals.  The frame table and the
// post-processing descriptors live in caller storage (fd, pf: n entries).
struct RenderPlan {
    int n = 0;
    FrameDev *fd = nullptr;
    PostFrameDev *pf = nullptr;
    bool post = true; // SSAA / FXAA passes follow (not for prr_trace)
    int64_t total = 0, qtotal = 0, pixels = 0, max_pix = 0;
    int max_w = 0, max_h = 0;
    bool any_fxaa = false, any_plain = false; // some frame with / without FXAA
    double ssaa_bytes = 0, fxaa_bytes = 0;
    int grid = 0, block = 0; // persistent integrator launch
};

} // namespace prr

struct prr_ctx {
    int device = 0;
    int cus = 0;
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr; // the adjoint's host-to-device copy of dL/dI (created on first use)
    hipEvent_t copy_done = nullptr;
    std::string err;
    int32_t metric = PRR_SCHW_KS_CART;
    double M = 1.0, a = 0.0, q = 0.0, p[16] = {};
    prr_tracer tracer{};
    prr_launch launch{256, 0, 64, 64, 0}; // refill 64: a wave takes 64 consecutive rays once all its lanes are idle
    std::shared_ptr<prr::TexAlloc> tex[PRR_MAX_TEXTURES];
    std::vector<prr_sphere> spheres;
    // the workspace of eager calls, grown on demand
    prr::DevBuf ws;
    // host-image staging and per-ray diagnostics
    prr::DevBuf stage_rgbf, stage_rgb8, cls, endpoint, steps, attempts, state8, x0u0;
    // adjoint
    prr::DevBuf adj_img, adj_dldi, adj_off, adj_ckpt, adj_ckpt2, adj_event, adj_rays, adj_rec, adj_lam, adj_grad_ray,
        adj_task_ray, adj_task_seg, adj_levelstart, adj_nseg, adj_over, adj_plan, adj_jac, adj_jacstart, adj_recall,
        adj_grad;
    // pinned host memory for the adjoint's small copies between passes
    // (pageable copies cost tens of microseconds each)
    void *pinned = nullptr;
    size_t pinned_bytes = 0;
    hipEvent_t ev[6] = {};
    prr_stats stats{};
    int live_graphs = 0;    // graphs captured on this context and not yet destroyed
    bool destroyed = false; // prr_destroy was called while graphs were alive
};

// A captured render owns its workspace and holds references to the textures
// it reads: eager calls on the context afterwards neither overwrite its frame
// and post tables nor move or free its buffers.
struct prr_graph {
    prr_ctx *ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void *block = nullptr; // the graph's workspace
    prr::CountersDev *counters = nullptr;
    std::shared_ptr<prr::TexAlloc> tex[PRR_MAX_TEXTURES];
    int64_t total_rays = 0, pixels = 0;
    double ssaa_bytes = 0, fxaa_bytes = 0;
};

namespace prr {

// Records the message (on ctx, and as the global message) and returns st.
prr_status fail(prr_ctx *ctx, prr_status st, const char *fmt, ...);
// Grows b to at least `bytes` (its contents are not kept).
prr_status ensure(prr_ctx *ctx, DevBuf &b, size_t bytes);
// Like ensure, for optional buffers: false (and no error recorded) when the
// memory is not there.
bool grow_quietly(DevBuf &b, size_t bytes);
// Grows the context's pinned host block to at least `bytes` (contents not kept).
prr_status ensure_pinned(prr_ctx *ctx, size_t bytes);

// Validates the frames, fills the frame table fd[0..n) and the totals.
prr_status plan_frames(prr_ctx *ctx, const prr_frame *frames, int n, FrameDev *fd, RenderPlan &P);
WorkspaceShape workspace_shape(const prr_ctx *ctx, const RenderPlan &P);
// The context's own workspace (eager calls), grown to the plan's shape.
prr_status context_workspace(prr_ctx *ctx, const RenderPlan &P, Workspace &ws);
// Post-processing descriptors pf[0..n): the frame's ray colours and FXAA
// input in the workspace; outputs to device images in place, to host images
// through the staging buffers (stage_rgbf / stage_rgb8, W*H of every frame).
void plan_post(const prr_frame *frames, const prr_image *images, const Workspace &ws, float *stage_rgbf,
               uint8_t *stage_rgb8, RenderPlan &P, PostFrameDev *pf);
JobDev make_job(const prr_ctx *ctx, const RenderPlan &P, const Workspace &ws);
// Enqueue reset -> init -> integrate -> shade -> SSAA -> FXAA, with timing events.
prr_status enqueue_render(prr_ctx *ctx, const RenderPlan &P, const JobDev &job, const Workspace &ws, hipStream_t st,
                          bool events);
// Counters of the finished call (synchronises the stream) into ctx->stats.
prr_status read_stats(prr_ctx *ctx, const RenderPlan &P, const CountersDev *counters, hipStream_t st, bool events);
// ctx->stats from counters already copied to the host (the stream synchronised).
prr_status stats_from(prr_ctx *ctx, const RenderPlan &P, const CountersDev &c, bool events);
// With `device` current: the device address of its copy of the allocation-free
// workspace, looked up once (noalloc.cpp; called by prr_create).
prr_status noalloc_setup(int device);

} // namespace prr

#define HIPTRY(ctx, call)                                                                                       \
    do {                                                                                                        \
        hipError_t e_ = (call);                                                                                 \
        if (e_ != hipSuccess) return prr::fail((ctx), PRR_EHIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)
