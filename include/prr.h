/*
 * prr.h -- C ABI of the MI355X (gfx950) null-geodesic integrator and its
 * anti-aliasing passes.
 *
 * The reference program (GKylian/ParRelRay) has no library boundary: a metric
 * is a header chosen at compile time (main/metric.h:12) and the host drives
 * seven synchronous <<<>>> launches per frame (prr.cpp:237-341) with Camera,
 * Spacetime and Tracer passed by value.  This header is the drop-in boundary
 * that replaces that path.  Every entry point below names the reference code
 * it stands in for.
 *
 * Conventions
 *  - extern "C", plain pointers and sizes, no exceptions cross the boundary.
 *  - Every call returns a prr_status; prr_last_error() gives the message.
 *    The reference calls exit(99) on any CUDA error (prr.cpp:37-47); this
 *    library never exits.
 *  - Buffers are host memory unless a structure says `on_device = 1`, in
 *    which case they are device pointers on the context's device.
 *  - One context per host thread; calls on one context are serialised.
 *  - Angles are radians; lengths are in units of M (G = c = 1).
 */
#ifndef PRR_H
#define PRR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PRR_ABI_VERSION 3
#define PRR_MAX_SPHERES 8
#define PRR_MAX_TEXTURES 10 /* slot 0 = celestial sphere, slots 1..8 = sphere textures, 9 = second sky */
#define PRR_SKY2_SLOT 9     /* celestial sphere seen through a wormhole's throat (bimage2) */

typedef struct prr_ctx prr_ctx;
typedef struct prr_params prr_params;
typedef struct prr_graph prr_graph;

typedef enum prr_status {
    PRR_OK = 0,
    PRR_EINVAL = 1,  /* invalid argument or configuration */
    PRR_ENOMEM = 2,  /* host or device allocation failed */
    PRR_EHIP = 3,    /* HIP runtime error */
    PRR_EIO = 4,     /* file could not be read or written */
    PRR_EBUDGET = 5, /* a work limit was exceeded */
    PRR_ENODEV = 6   /* no usable GPU */
} prr_status;

/* Metric plug-ins (reference: metrics/flat.cuh, metrics/schwarzschild.cuh,
 * "metrics/schwarzschild - sph.cuh", metrics/kerr.cuh). */
typedef enum prr_metric {
    PRR_FLAT = 0,         /* Minkowski, Cartesian (x, y, z, t) */
    PRR_SCHW_KS_CART = 1, /* Schwarzschild, Kerr-Schild Cartesian (x, y, z, t) */
    PRR_SCHW_SPH = 2,     /* Schwarzschild, Schwarzschild coordinates (r, theta, phi, t) */
    PRR_KERR_BL = 3,      /* Kerr, Boyer-Lindquist (r, theta, phi, t) */
    PRR_ELLIS = 4         /* Ellis wormhole (l, theta, phi, t), throat radius p[0], escape radius
                           * beyond the throat p[1] (0: r_escape; rout2): reaches the reference's
                           * second sky (bimage2, "l < 0 if wormhole", main/base.h:168) */
} prr_metric;

/* Solvers; values follow the reference's SOLVER enum (main/base.h:26-34),
 * which names "rk" and "dormand" without kernels: here RK4 is the classic
 * fixed-step 4th-order Runge-Kutta (step h) and DORMAND the adaptive
 * Dormand-Prince 5(4) pair with the RKF45 controller, propagating the 5th
 * order.  Both use RKF45's located events. */
typedef enum prr_solver { PRR_EULER = 0, PRR_RK4 = 1, PRR_RKF45 = 2, PRR_DORMAND = 5 } prr_solver;

typedef enum prr_precision {
    PRR_FP64 = 0,     /* plain form: FP64 throughout */
    PRR_FP32_BULK = 1 /* RK stages in FP32; state, error estimate and r < r_fp64 in FP64 */
} prr_precision;

/* Sample placement; values follow SSAA_DISTR (main/base.h:46-53). */
/* SSAA sample distributions; values follow the reference's SSAA_DISTR
 * (main/base.h:46-53).  RANDOM and JITTERED draw their offsets from a
 * counter-based stream keyed by prr_frame.seed and the ray index (the
 * reference's curand states, kernels/ssaa.cuh:20-86). */
typedef enum prr_ssaa {
    PRR_SSAA_NONE = 0,
    PRR_SSAA_REGULAR = 1, /* n x n grid at (p + 1) / (n + 1) of the pixel */
    PRR_SSAA_RANDOM = 2,  /* samples uniform over the pixel */
    PRR_SSAA_JITTERED = 4 /* one uniform sample in each cell of an n x n grid */
} prr_ssaa;

typedef enum prr_projection {
    PRR_PROJ_PINHOLE = 0,  /* Camera::rayDir_PinHole (main/Camera.h:30-33) */
    PRR_PROJ_SPHERICAL = 1 /* Camera::rayDir_Spherical (main/Camera.h:39-42) */
} prr_projection;

typedef enum prr_filter {
    PRR_FILTER_SQAVG = 0,   /* RMS of the 2x2 texel block (main/base.h:150-156) */
    PRR_FILTER_BILINEAR = 1 /* bilinear, wrapped in u and clamped in v */
} prr_filter;

/* Per-ray termination class.  The reference's RESULT (main/base.h:37-43)
 * folds escape and horizon into SUCCESS; they are separate here. */
typedef enum prr_class {
    PRR_ESCAPE = 0,   /* r >= r_escape: celestial-sphere colour */
    PRR_HORIZON = 1,  /* r <= r_h (1 + eps_rel) + eps_abs: black */
    PRR_SPHERE = 2,   /* step chord crossed a scene sphere: sphere colour */
    PRR_STEP_CAP = 3, /* step, attempt or affine-parameter cap (INFINITE_GEO): black */
    PRR_NAN = 4       /* non-finite state (MATH_ERROR): black */
} prr_class;

/* Integrator settings (reference Tracer, main/base.h:166-170, filled from
 * the [tracer] block at prr.cpp:184-190). */
typedef struct prr_tracer {
    int32_t solver;    /* prr_solver */
    int32_t precision; /* prr_precision */
    double h;          /* Euler step; RKF45 initial step */
    double hmin, hmax; /* RKF45 step clamp */
    double atol, rtol; /* RKF45 tolerance: |err_i| <= atol + rtol*max(|y_i|,|y4_i|); reference eps = atol, rtol = 0 */
    double r_escape;   /* escape radius (rout1) */
    double horizon_eps_rel, horizon_eps_abs; /* horizon test r <= r_h (1 + rel) + abs */
    double lambda_max;                       /* affine-parameter cap (reference: 1000 RKF, 5000 Euler) */
    int64_t max_steps;                       /* accepted-step cap */
    int64_t max_attempts;                    /* RKF45 attempt cap (reference: 1e6) */
    double r_fp64;                           /* FP32-bulk: FP64 stages when r < r_fp64 (in M) */
} prr_tracer;

typedef struct prr_sphere {
    double c[3];    /* centre, Cartesian (for spherical-type metrics: r sin(th) cos(ph), ...) */
    double radius;
    int32_t tex;    /* texture slot 1..8 (0 = untextured white) */
    int32_t pad;
} prr_sphere;

/* One frame: camera pose and image, spacetime parameters, scene.  The
 * reference keeps these in Camera (main/Camera.h:17-52), Spacetime
 * (main/Spacetime.h:15-47) and the frameN / sphereN blocks read at
 * prr.cpp:245-253 and main/Spacetime.cpp:38-45. */
typedef struct prr_frame {
    double pos[3];      /* camera position, Cartesian */
    double alpha, beta; /* camera azimuth and tilt (radians; Camera::computeVectors) */
    double vfov;        /* vertical field of view (radians) */
    int32_t width, height;
    int32_t ssaa;       /* prr_ssaa */
    int32_t samples;    /* samples per pixel: 1 for NONE, n*n for REGULAR and JITTERED, any for RANDOM */
    int32_t projection; /* prr_projection */
    int32_t fxaa;       /* 1: FXAA after the SSAA combine */
    double mass;        /* M */
    double spin;        /* Kerr a (ignored by the other metrics) */
    int32_t n_spheres;
    uint32_t seed;      /* RANDOM / JITTERED SSAA: key of the sample stream */
    prr_sphere spheres[PRR_MAX_SPHERES];
} prr_frame;

/* Output image of one frame (W*H*3, row-major, row 0 = top). */
typedef struct prr_image {
    uint8_t *rgb8; /* RGB8 = roundf(255 c) (cudamaths/image.cu:31); may be NULL */
    float *rgbf;   /* linear RGB after SSAA (+FXAA); may be NULL */
    int32_t on_device;
    int32_t pad;
} prr_image;

/* Optional per-ray diagnostics, all frames of a call concatenated in frame
 * order; ray index within a frame = (j*W + i)*samples + k
 * (kernels/ssaa.cuh:45).  Any pointer may be NULL. */
typedef struct prr_ray_out {
    int8_t *cls;       /* prr_class */
    double *endpoint;  /* 3 per ray: escape direction, sphere normal at the hit, or final Cartesian position */
    int32_t *steps;    /* accepted steps */
    int32_t *attempts; /* RKF45 attempts (= steps for Euler) */
    float *color;      /* 3 per ray: colour before SSAA */
    int32_t on_device;
    int32_t pad;
} prr_ray_out;

/* Persistent-kernel launch shape (chosen by the autotuner). */
typedef struct prr_launch {
    int32_t block;         /* workgroup size, multiple of 64, <= 1024 */
    int32_t blocks_per_cu; /* resident workgroups per CU in the persistent grid */
    int32_t refill;        /* a wave refills once at least this many lanes are idle (1..64);
                              PRR_FP32_BULK always uses 64 */
    int32_t checkpoint;    /* adjoint: accepted steps between per-ray checkpoints */
    int64_t adjoint_record_bytes; /* adjoint: device memory for recorded steps (48 B per step; 0: 8 GiB).
                                     When every ray's first max(8, 512 / checkpoint) * checkpoint steps
                                     fit it, the first pass records them and nothing is replayed;
                                     otherwise it bounds one chunk of replayed steps, and a smaller
                                     budget replays in more chunks */
} prr_launch;

/* Counters and HIP-event timings of the last prr_render / graph launch. */
typedef struct prr_stats {
    int64_t rays;
    int64_t attempts;       /* RK attempts (accepted + rejected), all rays */
    int64_t accepted;       /* accepted steps */
    int64_t attempts_fp64;  /* attempts whose stages ran in FP64 */
    int64_t pixels;
    double integrate_ms;    /* persistent integrator kernel */
    double ssaa_ms;         /* SSAA combine + RGB8 kernel of the frames without FXAA */
    double fxaa_ms;         /* SSAA combine + FXAA + RGB8 kernel of the frames with FXAA */
    double total_ms;        /* whole call on the stream */
    double flops_fp64;      /* algorithmic FP64 flops of the integration */
    double flops_fp32;      /* algorithmic FP32 flops of the integration */
    double ssaa_bytes;      /* algorithmic HBM bytes of the SSAA kernel (ray colours in, images out) */
    double fxaa_bytes;      /* algorithmic HBM bytes of the FXAA kernel (ray colours in, images out) */
    double adjoint_ms;      /* prr_adjoint: checkpointed forward + backward passes */
    int64_t adjoint_steps;  /* prr_adjoint: accepted steps swept backwards */
    double init_ms;         /* ray set-up kernel */
    double shade_ms;        /* shading kernel */
} prr_stats;

/* Differentiable parameters of an orbiting-camera frame (adjoint). The
 * camera sits at pos = r_cam (sin i cos phc, sin i sin phc, cos i) looking at
 * the origin (alpha = phc + pi, beta = pi/2 - i); the sphere sits on the
 * equator at (sphere_r cos phs, sphere_r sin phs, 0). */
typedef struct prr_orbit {
    double spin, r_cam, incl, sphere_phi; /* gradient is taken w.r.t. these four */
    double cam_phi;
    double sphere_r;
    double sphere_radius;
    int32_t sphere_tex;
    int32_t has_sphere;
} prr_orbit;

/* ---- context --------------------------------------------------------- */

/* Replaces cudaDeviceReset + global Tracer/Spacetime/Camera (prr.cpp:50-52, 348). */
prr_status prr_create(prr_ctx **out, int hip_device);
void prr_destroy(prr_ctx *ctx);
const char *prr_last_error(const prr_ctx *ctx);
/* Message of the last failed call that had no context (prr_create, params). */
const char *prr_global_error(void);
int prr_abi_version(void);

/* Spacetime::setBhMass/Spin/Charge/setBhP (main/Spacetime.h:21-26) plus the
 * compile-time metric choice (main/metric.h:12).  M and a are used by
 * prr_trace; frames carry their own mass and spin.  q: the charge of
 * PRR_SCHW_SPH; p: PRR_ELLIS's throat and far-side escape radii (p[0], p[1]). */
prr_status prr_set_spacetime(prr_ctx *ctx, int32_t metric, double M, double a, double q, const double p[16]);
/* Tracer fields (prr.cpp:184-190). */
prr_status prr_set_tracer(prr_ctx *ctx, const prr_tracer *tracer);
/* tex2d::load (cudamaths/tex2d.cuh:50-80): RGB8 texels, converted to linear
 * float as byte/255 (stbi_ldr_to_hdr_gamma(1), cudamaths/image.cu:15). */
prr_status prr_set_texture(prr_ctx *ctx, int slot, const uint8_t *rgb8, int w, int h, int32_t filter);
prr_status prr_set_launch(prr_ctx *ctx, const prr_launch *launch);
prr_status prr_get_launch(const prr_ctx *ctx, prr_launch *launch);

/* ---- rendering ------------------------------------------------------- */

/* One persistent launch for all n frames: ray generation
 * (kernels/ssaa.cuh:36-48, 92-99), null initialisation
 * (kernels/solver.cuh:26-56), integration (trace_Euler :68-116 /
 * trace_RKF :142-220), colouring (rayToRGB, metrics/schwarzschild.cuh:151-183),
 * then the SSAA combine (kernels/ssaa.cuh:109-131), FXAA (:150-182) and RGB8
 * quantisation (cudamaths/image.cu:31).  `stream` is a hipStream_t or NULL
 * for the context's own stream.  Returns when the results are in `images`
 * (and `rays` when given). */
prr_status prr_render(prr_ctx *ctx, const prr_frame *frames, int n, prr_image *images, prr_ray_out *rays,
                      void *stream);
prr_status prr_get_stats(const prr_ctx *ctx, prr_stats *stats);

/* Allocation-free workspace of prr_render_noalloc: its size on every device
 * the library runs on, and the most frames one call renders. */
#define PRR_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
#define PRR_NOALLOC_MAX_FRAMES 16

/* prr_render without any host or device allocation, for callers that cannot
 * allocate (code under stream capture, tight serving loops).  No counterpart
 * in the reference.  The same kernels and results as prr_render, with the
 * workspace in a static device array of PRR_NOALLOC_SCRATCH_BYTES per device
 * (held until the process ends) instead of buffers the context allocates.
 * Images must be device buffers (on_device = 1); no per-ray outputs.  Up to
 * PRR_NOALLOC_MAX_FRAMES frames whose rays fit the array (about 94 bytes per
 * ray, 12 per FXAA pixel, beside the integrator's per-lane contexts, whose
 * grid shrinks to fit); larger inputs give PRR_EBUDGET.  Calls that overlap
 * on one device share the array and race; the call returns when the images
 * are written. */
prr_status prr_render_noalloc(prr_ctx *ctx, const prr_frame *frames, int n, const prr_image *images,
                              void *stream);

/* Integrator alone from explicit initial states: x0u0 is 8 doubles per ray
 * (x^1, x^2, x^3, t, u^1, u^2, u^3, u^t) in the metric's coordinates, the
 * state the reference's Ray carries (main/base.h:172-176).  xend receives the
 * final 8-vector (for Kerr, t is not integrated and is returned as 0).  Uses
 * the context's spacetime, tracer and spheres.  Host buffers. */
prr_status prr_trace(prr_ctx *ctx, const double *x0u0, int n, double *xend, int32_t *cls, int32_t *steps,
                     int32_t *attempts);
/* Scene spheres used by prr_trace (Spacetime::setupScene, main/Spacetime.cpp:15-29). */
prr_status prr_set_spheres(prr_ctx *ctx, const prr_sphere *spheres, int n);

/* Capture the whole multi-frame render (counter reset, persistent
 * integrator, SSAA, FXAA+RGB8) in a hipGraph; images must be device
 * buffers and stay allocated while the graph lives.  Replay with
 * prr_graph_launch.  The graph owns its frame table and device buffers and
 * keeps the textures it captured: later calls on the context (renders, traces,
 * adjoints, prr_set_launch, prr_set_texture) do not change what it replays.
 * A context destroyed while graphs live is released with its last graph; its
 * graphs then fail to launch. */
prr_status prr_graph_create(prr_ctx *ctx, const prr_frame *frames, int n, const prr_image *images, prr_graph **out);
prr_status prr_graph_launch(prr_graph *g, void *stream);
void prr_graph_destroy(prr_graph *g);

/* ---- differentiation --------------------------------------------------- */

/* Fill a frame from a template and orbit parameters. */
prr_status prr_frame_from_orbit(const prr_frame *tmpl, const prr_orbit *orbit, prr_frame *out);
/* Gradient of L w.r.t. (spin, r_cam, incl, sphere_phi) given dL/dI for the
 * frame's linear-RGB image (W*H*3 floats, after SSAA, no FXAA).  The
 * forward render is recomputed from per-ray checkpoints taken every
 * launch.checkpoint accepted steps.  img_out (optional, W*H*3) receives the
 * forward image.  No counterpart in the reference. */
prr_status prr_adjoint(prr_ctx *ctx, const prr_frame *tmpl, const prr_orbit *orbit, const float *dL_dimg,
                       int32_t on_device, double grad[4], float *img_out);

/* ---- host-side scene description (.params) ----------------------------- */

/* readParams (main/io.h:21-59): parse one .params file into a new store. */
prr_status prr_params_read(prr_params **out, const char *path);
/* readParams into an existing store (the reference reads sim, traj and
 * scene into one map, prr.cpp:207-210). */
prr_status prr_params_read_into(prr_params *p, const char *path);
/* readParams(sim) followed by its [tracer] traj and scene files (init, prr.cpp:203-227). */
prr_status prr_params_load_sim(prr_params **out, const char *sim_path);
void prr_params_free(prr_params *p);
/* headerExists / parExists / par / pari / parf (main/Parameters.cpp:9-43). */
int prr_params_header_exists(const prr_params *p, const char *header);
int prr_params_par_exists(const prr_params *p, const char *header, const char *key);
/* Copies the value (or "" when missing) into buf; returns its length. */
int prr_params_par(const prr_params *p, const char *header, const char *key, char *buf, int buflen);
int prr_params_pari(const prr_params *p, const char *header, const char *key);
double prr_params_parf(const prr_params *p, const char *header, const char *key, double dflt);
/* checkFrame (prr.cpp:139-153): number of consecutive complete frameN blocks. */
int prr_params_frame_count(const prr_params *p);
/* createScene (prr.cpp:159-198): metric name ([spacetime] metric, default
 * the reference's compile-time choice), spacetime and tracer. */
prr_status prr_params_spacetime(const prr_params *p, int32_t *metric, double *M, double *a, double *q,
                                double pvals[16]);
prr_status prr_params_tracer(const prr_params *p, prr_tracer *out);
/* renderframe's per-frame pose and scene update (prr.cpp:245-253,
 * main/Spacetime.cpp:38-45) for frame i; also the texture path of each sphere. */
prr_status prr_params_frame(const prr_params *p, int i, prr_frame *out);

/* ---- texture files ------------------------------------------------------ */

/* stbi_load(path, &w, &h, &n, 3) (cudamaths/image.cu:13-21): a binary PPM
 * (P6), PNG (8/16-bit, any colour type, non-interlaced; alpha dropped) or
 * baseline / extended-sequential Huffman JPEG (grey or YCbCr, any sampling)
 * decoded to RGB8, row-major.  Call with rgb8 = NULL to get *w, *h; then with
 * a w*h*3 buffer and the same *w, *h.  Progressive JPEG is refused
 * (PRR_EINVAL); a missing file gives PRR_EIO.  Messages via prr_global_error. */
prr_status prr_image_read(const char *path, uint8_t *rgb8, int *w, int *h);

/* ---- camera --------------------------------------------------------------- */

/* Camera::computeVectors (main/Camera.cpp:19-28): forward, up and right unit
 * vectors of a camera with azimuth alpha and tilt beta (radians). */
prr_status prr_camera_basis(double alpha, double beta, double forward[3], double up[3], double right[3]);
/* Camera::rayDir_PinHole / rayDir_Spherical (main/Camera.h:30-42), by the
 * frame's projection: the unit direction through screen position (u, v) in
 * [0, 1]^2 (u to the right, v down), as the kernels compute it. */
prr_status prr_camera_dir(const prr_frame *f, double u, double v, double d[3]);

/* ---- procedural inputs -------------------------------------------------- */

/* Seeded RGB8 textures used in place of the reference's image files:
 * a checkerboard of cw x ch cells with per-cell hue jitter, and
 * 4-octave value noise with smoothstep interpolation (persistence 0.5). */
prr_status prr_texture_checker(uint8_t *rgb8, int w, int h, int cw, int ch, uint64_t seed);
prr_status prr_texture_value_noise(uint8_t *rgb8, int w, int h, int octaves, uint64_t seed);

#ifdef __cplusplus
}
#endif

#endif /* PRR_H */
