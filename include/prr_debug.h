/*
 * prr_debug.h -- host-side evaluation of the device metric plug-ins.
 *
 * The reference's metric interface is getGeod / getMetric / initRays per
 * metric header (metrics/schwarzschild.cuh:90,123; kernels/solver.cuh:26-56).
 * The kernels in this library use those formulas through templates that are
 * also compiled for the host; these entry points evaluate them on the CPU so
 * that the formulas can be checked without a GPU.  No device is touched.
 */
#ifndef PRR_DEBUG_H
#define PRR_DEBUG_H

#include "prr.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Integrator state size of a metric: 8 for (x, u) metrics, 6 for Kerr
 * (u = 1/r, du/ds, mu = cos th, dmu/ds, phi, lambda) in Mino time s. */
int prr_debug_state_size(int32_t metric);
/* Right-hand side f(y) of the integrator state in FP64 (fp32 = 0) or FP32
 * (fp32 = 1).  E, L, Q: Kerr's per-ray constants (ignored elsewhere). */
prr_status prr_debug_rhs(int32_t metric, double M, double a, double E, double L, double Q, const double *y,
                         int32_t fp32, double *f);
/* Kerr FP32-bulk stage function: g = f(y0 + d) - f(y0) from the FP64
 * difference coefficients at y0, evaluated in FP32 (y0, d, g: 6 each). */
prr_status prr_debug_kerr_delta_rhs(double M, double a, double E, double L, double Q, const double *y0,
                                    const double *d, double *g);
/* Kerr RKF45 attempt: 4th-order increment d4 and error vector e (6 each),
 * plain FP64 stages (fp32 = 0) or the FP32-bulk stages (fp32 = 1). */
prr_status prr_debug_kerr_stages(double M, double a, double E, double L, double Q, const double *y, double h,
                                 int32_t fp32, double *d4, double *e);
/* Reverse mode of the Kerr right-hand side used by the adjoint kernel:
 * yb[6] += fb . df/dy, cb[4] += fb . df/d(E, L, Q, a). */
prr_status prr_debug_rhs_vjp(double M, double a, double E, double L, double Q, const double *y, const double *fb,
                             double *yb, double *cb);
/* Adjoint pieces (Kerr).  Ray set-up of an orbiting camera at screen
 * position (u, v): z = (u, du/ds, mu, dmu/ds, phi, E, L, Q) and its Jacobian
 * dz/d(a, r_cam, incl) (8 x 3, row-major). */
prr_status prr_debug_orbit_init(double M, double a, double r_cam, double incl, double cam_phi, double vfov,
                                double aspect, double u, double v, double z[8], double dz[24]);
/* One FP64 RKF45 step of fixed size h, and its transpose Jacobian product:
 * lam[9] (adjoint of (y'[0..4], E, L, Q, a)) -> adjoint of (y, E, L, Q, a). */
prr_status prr_debug_kerr_step(double M, double a, double E, double L, double Q, const double y[6], double h,
                               double out[6]);
prr_status prr_debug_kerr_step_vjp(double M, double a, double E, double L, double Q, const double y[6], double h,
                                   double lam[9]);
/* Bilinear sky colour of an escaping Kerr ray and its Jacobian
 * dc/d(y[0..4], E, L, Q, a) (3 x 9, row-major); rgb8 is a w x h texture. */
prr_status prr_debug_escape_color(double M, double a, double E, double L, double Q, const double y[5],
                                  const uint8_t *rgb8, int w, int h, double c[3], double dc[27]);
/* (x, u) -> integrator state y and Kerr's (E, L, Q) (initRays' output in the
 * metric's coordinates). */
prr_status prr_debug_state_from_xu(int32_t metric, double M, double a, const double x[4], const double u[4],
                                   double *y, double ELQ[3]);
/* Integrator state -> (x, u). */
prr_status prr_debug_xu_from_state(int32_t metric, double M, double a, double E, double L, double Q, const double *y,
                                   double x[4], double u[4]);
/* initRays for a camera position and unit direction. */
prr_status prr_debug_init_ray(int32_t metric, double M, double a, const double pos[3], const double d[3], double *y,
                              double ELQ[3]);
/* Direction of ray `local` of a frame (createRays_* + rayDir_*). */
prr_status prr_debug_camera_ray(const prr_frame *f, int64_t local, double d[3]);
/* Ray ids of queue positions q[0..n) of one frame in the persistent
 * integrator (tiles of pixels per wave; a permutation of the frame's ids). */
prr_status prr_debug_queue_rays(const prr_frame *f, const int64_t *q, int64_t n, int64_t *ids);
/* The device integrator's per-ray code run on the host for n rays: rays ids[i]
 * of frame f (x0u0 = NULL) or the initial states x0u0 (8 per ray), with the
 * kernels' tracer semantics.  Untextured (escape and sphere colours are
 * black).  Any output may be NULL: cls, steps, attempts (n each), endpoint
 * (3n), state8 (8n). */
prr_status prr_debug_trace_rays(int32_t metric, const prr_tracer *tracer, const prr_frame *f, const double *x0u0,
                                const int64_t *ids, int n, int32_t *cls, double *endpoint, int32_t *steps,
                                int32_t *attempts, double *state8);
/* Per accepted step of one ray (the same loop as prr_debug_trace_rays): rows of
 * 12 doubles (steps, attempts, step proposal before the accepted attempt, FP64 attempts,
 * state[8]); *n_out = rows written (at most max_n). */
prr_status prr_debug_trace_log(int32_t metric, const prr_tracer *tracer, const prr_frame *f, int64_t id, int max_n,
                               double *log, int *n_out);

#ifdef __cplusplus
}
#endif

#endif /* PRR_DEBUG_H */
