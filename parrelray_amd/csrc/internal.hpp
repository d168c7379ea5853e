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
    double M, a;
    double r_hor; // horizon test radius r_h (1 + eps_rel) + eps_abs (negative: no test)
    int64_t ray_begin;
    double sph[PRR_MAX_SPHERES][4]; // centre xyz, radius
    int32_t sph_tex[PRR_MAX_SPHERES];
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
    EventDev *event;  // 1 per ray (adjoint)
};

// Device-side counters; `queue` is the ray queue head.
struct CountersDev {
    unsigned long long queue;
    unsigned long long attempts;
    unsigned long long accepted;
    unsigned long long attempts_fp64;
    unsigned long long rays;
    unsigned long long mismatch; // adjoint: segments whose replay did not reproduce the checkpoint
    unsigned long long pad[2];
};

struct JobDev {
    const FrameDev *frames;
    int32_t nframes;
    int32_t refill;
    int64_t total;               // rays over all frames
    const double *explicit_x0u0; // prr_trace: 8 initial doubles per ray, frame 0 supplies the rest
    TracerDev tr;
    TexDev tex[PRR_MAX_TEXTURES];
    OutDev out;
    CountersDev *ctr;
};

// Kerr adjoint (prr_adjoint): checkpoints and per-ray gradients.
struct AdjDev {
    int32_t K;               // accepted steps between checkpoints
    int32_t pad;
    const int32_t *steps;    // accepted steps of every ray (first forward pass)
    const int8_t *cls;       // termination class of every ray (first forward pass)
    const int64_t *ckpt_off; // index of each ray's first checkpoint
    double *ckpt;            // 8 doubles per checkpoint: y[6], h, attempts
    const float *dLdI;       // dL/dI, W*H*3
    const float *img;        // forward linear image, W*H*3
    const float *ray_color;  // forward ray colours, 3 per ray
    double *seg;             // per-lane segment buffers: K * 6 doubles per lane, lane-minor
    int64_t seg_lanes;       // lanes of the backward grid
    double *grad_ray;        // 4 doubles per ray: dL/d(a, r_cam, incl, phi_s)
    double cam_phi;          // fixed orbit parameters
    double sphere_r;
    unsigned long long *replay_mismatch; // segments the replay did not reproduce bit for bit
};

// Post-processing of one frame (SSAA combine, FXAA, RGB8).
struct PostFrameDev {
    const float *ray_color; // 3 * W*H*s, this frame's rays
    float *img;             // SSAA result, W*H*3 (scratch when FXAA follows)
    float *final_rgbf;      // final linear image or null
    uint8_t *rgb8;          // final RGB8 or null
    int32_t W, H, s, fxaa;
};

} // namespace prr
