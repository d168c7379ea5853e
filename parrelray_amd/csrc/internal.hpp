// internal.hpp -- structures shared by the host side (capi.cpp) and the HIP
// kernels (kernels.hip).  Plain data only: they travel as kernel arguments
// or device buffers.
#pragma once

#include <cstdint>

#include "../../include/prr.h"

namespace prr {

// Integrator settings as the kernels see them (Tracer, main/base.h:166-170).
struct TracerDev {
    int32_t solver;
    int32_t precision;
    double h, hmin, hmax;
    double atol, rtol;
    double r_escape;
    double lambda_max;
    double r_fp64;
    double inv_r_escape, inv_r_fp64; // 1 / r_escape, 1 / r_fp64 (Kerr's radius tests in u = 1 / r)
    int64_t max_steps;
    int64_t max_attempts;
};

// Linear RGB texture as float4 texels (alpha unused), row-major.
struct TexDev {
    const float4 *texels;
    int32_t w, h;
    int32_t filter;
    int32_t pad;
};

// Per-frame constants, precomputed on the host from prr_frame
// (Camera::computeVectors, main/Camera.cpp:19-28).
struct FrameDev {
    double pos[3];
    double F[3], U[3], R[3];  // camera forward, up, right
    double alpha, beta, vfov; // radians
    double tan_half;          // tan(vfov / 2)
    double aspect;            // W / H
    int32_t W, H, s, n;       // image, samples per pixel, sqrt(samples) (0 = NONE)
    int32_t projection;
    int32_t n_spheres;
    int32_t ssaa;  // prr_ssaa
    uint32_t seed; // RANDOM / JITTERED sample stream
    double M, a;
    double r_hor; // horizon test radius r_h (1 + eps_rel) + eps_abs (negative: no test)
    double inv_r_hor; // 1 / r_hor (Kerr's horizon test in u = 1 / r)
    double r_esc2; // escape radius beyond a wormhole's throat (l < 0; rout2)
    int64_t ray_begin;
    int64_t queue_begin; // first queue position (frames start at multiples of 64)
    double sph[PRR_MAX_SPHERES][4]; // centre xyz, radius
    int32_t sph_tex[PRR_MAX_SPHERES];
    // the FP32 screen's view of each sphere: centre and radius rounded to
    // float, then |cx| + |cy| + |cz| + R summed in that order in float (one
    // 32-byte record per sphere)
    float sphf[PRR_MAX_SPHERES][8];
    // initRays' per-frame terms for the spherical charts: the camera position
    // in (r, theta, phi), sin and cos theta, x^2 + y^2 and the reciprocals
    // that convert a direction into (dr, dtheta, dphi); for Kerr the metric
    // components at the camera (g_tt, g_tphi, g_rr, g_thth = Sigma, g_phph)
    double cam_r, cam_th, cam_ph, cam_st, cam_ct;
    double cam_ir, cam_x2y2, cam_ix2y2, cam_iu1;
    double g_tt, g_tph, g_rr, g_sig, g_phph;
};

// Terminal RKF45 step of a ray whose escape or sphere hit was located on
// the path: the state at the step's start, its size and the located fraction.
struct EventDev {
    double y_prev[8];
    double h, theta;
    int32_t cls, hit;
};

// Per-ray outputs (device pointers; any may be null).
struct OutDev {
    float *color;     // 3 per ray
    int8_t *cls;
    double *endpoint; // 3 per ray
    int32_t *steps;
    int32_t *attempts;
    double *state8;   // 8 per ray (prr_trace)
};

// Ray state between the passes, structure of arrays over the job's rays:
// y[k] at y + k * total (k < N), then E, L, Q for Kerr.  k_init writes the
// initial state, k_integrate replaces it by the final one, k_shade reads it.
struct RayStateDev {
    double *y;
    int8_t *cls;
    int8_t *hit;
    int32_t *steps;
    int32_t *attempts;
};

// Sphere-hit location in progress for one lane (RKF45): the accepted step's
// end state, the bracket / best states and the search variables.  One per
// lane of the integrator grid, in device memory; only lanes whose step chord
// passes near a sphere touch it (escape location keeps its state in the lane).
struct LocateCtx {
    double yn[8];    // end state of the accepted step
    double yb[8];    // state at the inside end of the bracket
    double ybest[8]; // state at the best hit so far
    double h;        // size of the step
    double th;       // fraction of the step of the pending sub-step
    double ta, t1, tb, Ga, G1, Gb; // minimum search
    double lo, hi, Glo, Ghi;       // entry bracket
    double G0, best_th;
    int32_t phase; // 0 next sphere, 1 first probe, 2 minimum search, 3 bracket
    int32_t k, mask, it, side, best_k;
};

// Adjoint modes of the integrator (AdjRunDev.mode).
enum { ADJ_NONE = 0, ADJ_CHECKPOINT = 1, ADJ_REPLAY = 2 };

// Kerr adjoint passes run by the integrator kernel itself.
struct AdjRunDev {
    int32_t mode;            // ADJ_*
    int32_t K;               // accepted steps between checkpoints
    int32_t pad;
    int32_t cap;             // first pass (ckpt_off null): checkpoint slots per ray
    const int32_t *nsteps;   // accepted steps of every ray (null in the first pass)
    const int64_t *ckpt_off; // first checkpoint of each ray: >= 0 in ckpt, < 0 at -1 - off in ckpt2
    double *ckpt;            // 8 doubles per checkpoint: y[6], h, attempts
    double *ckpt2;           // checkpoints of the rays with more than cap segments
    const int64_t *rays;     // replay, and a checkpoint pass over some rays: ray of each job index
    const int32_t *seg;      // replay: segment index j of each job index (a task is one segment of one ray)
    double *rec;             // replay: K * 6 doubles per task, [slot * 6 + c] * rec_stride + task
    int64_t rec_stride;      // replay: tasks of this launch
    EventDev *event;         // replay (and a recording first pass): terminal event of each ray
    unsigned long long *mismatch; // replay: segments that did not reproduce their checkpoint
    // first pass recording every accepted step (rec_cap > 0): 6 doubles
    // (y[0..4], h) per step at rec_all + (ray * rec_cap + step) * 6, for the
    // first rec_cap steps of each ray
    double *rec_all;
    int64_t rec_cap;
};

// Device-side counters; `queue` is the ray queue head.
struct CountersDev {
    unsigned long long queue;
    unsigned long long attempts;
    unsigned long long accepted;
    unsigned long long attempts_fp64;
    unsigned long long rays;
    unsigned long long mismatch; // adjoint: replayed segments or rays that did not reproduce pass 1
    unsigned long long pad[2];
};

struct JobDev {
    const FrameDev *frames;
    int32_t nframes;
    int32_t refill;
    int64_t total;               // rays over all frames (replay: job indices)
    int64_t qtotal;              // render queue length: total plus the padding between frames
    const double *explicit_x0u0; // prr_trace: 8 initial doubles per ray, frame 0 supplies the rest
    TracerDev tr;
    TexDev tex[PRR_MAX_TEXTURES];
    OutDev out;
    RayStateDev rs;
    int64_t rs_total;            // rays in the ray-state arrays
    int64_t rs_base;             // ray id of the arrays' first entry (0 on the device)
    LocateCtx *loc;              // one per lane of the integrator grid
    AdjRunDev adj;
    CountersDev *ctr;
};

// Kerr adjoint reverse pass (per ray).
struct AdjDev {
    int32_t K;
    int32_t jlo, jhi;        // segment levels of this launch's replayed chunk, jhi >= jlo
    int64_t count;           // rays in this launch (the first `count` of `rays`: those with a segment jlo)
    int64_t seed_lo;         // seeding: ranks seed_lo..count-1 end in this chunk (their last segment is in it)
    const int64_t *rays;     // rays sorted by decreasing segment count
    const int64_t *levelstart; // task index of (level j, rank 0); level j holds the ranks < count of level j
    int64_t t0;              // first task of the chunk (its records start there)
    int64_t rec_stride;      // tasks in the chunk
    const int32_t *nsteps;   // accepted steps of every ray
    const int8_t *cls;       // termination class of every ray
    const double *rec;       // replayed steps of the chunk (AdjRunDev.rec layout), or the first pass's records
    int64_t rec_cap;         // > 0: rec holds the first pass's records (AdjRunDev.rec_all layout)
    const EventDev *event;   // terminal events
    double *lam;             // 10 doubles per ray: lambda (y[0..4], E, L, Q, a), dL/dphi_s
    const float *dLdI;       // dL/dI, W*H*3
    const float *img;        // forward linear image, W*H*3
    const float *ray_color;  // forward ray colours, 3 per ray
    double *grad_ray;        // 4 doubles per ray: dL/d(a, r_cam, incl, phi_s)
    double cam_phi;          // fixed orbit parameters
    double sphere_r;
    unsigned long long *mismatch; // rays whose replay did not end with their class
    // Long rays (ranks < nlong: more than kAdjLongLast segments, all levels in
    // this chunk): the sweep takes only their last segment, while the
    // transposed Jacobians of their other segments are formed in parallel
    // (jac, 81 doubles per (level j, rank) at task jacstart[j] + rank) and
    // chained onto lambda afterwards.
    int64_t nlong;
    int64_t njac;                // Jacobian tasks (times 9 columns: threads)
    int32_t jlevels;             // levels with Jacobian tasks: 0 .. jlevels - 1
    const int64_t *jacstart;     // first task of level j, jlevels + 1 entries
    double *jac;
};
// Rays whose last segment index exceeds this take the Jacobian path.
constexpr int32_t kAdjLongLast = 1;

// Segment-count histogram bins of the device-side adjoint plan; tracers whose
// step cap allows more segments per ray plan on the host.
constexpr int32_t kAdjPlanBins = 4096;

// Device-side planning of the adjoint's replay (after the checkpointing pass).
struct AdjPlanDev {
    const int32_t *steps;        // accepted steps of every ray
    int64_t T;                   // rays
    int32_t K, cap;              // checkpoint interval; first-pass checkpoint slots per ray
    int32_t nbins;               // histogram bins (> every bin below), at most kAdjPlanBins
    int32_t L;                   // length classes of the last segment per segment count (bin n L + class)
    int32_t nlevels;             // tasks: segment levels (the largest segment count)
    int32_t *nseg;               // segments of every ray
    int64_t *off;                // checkpoint offsets (AdjRunDev.ckpt_off)
    int64_t *over;               // the rays with more than cap segments, any order
    unsigned long long *hist;    // rays per bin: segment count, then the last segment's length class
    unsigned long long *ctr;     // [0] overflow checkpoints, [1] overflow rays, [2] accepted steps, [3] max segments
    unsigned long long *cursor;  // order: next position of each bin (start positions from the host)
    int64_t *order;              // rays by decreasing bin (segment count, then last segment's length)
    const int64_t *levelstart;   // tasks: first task of level j (levels descending: level nlevels-1 first)
    int64_t ntask;
    int64_t *task_ray;
    int32_t *task_seg;
};

// Post-processing of one frame (SSAA combine, FXAA, RGB8).
struct PostFrameDev {
    const float *ray_color; // 3 * W*H*s, this frame's rays
    float *final_rgbf;      // final linear image or null
    uint8_t *rgb8;          // final RGB8 or null
    int32_t W, H, s, fxaa;
};

} // namespace prr
