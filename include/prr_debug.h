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

/* Integrator state size of a metric: 8 for (x, u) metrics, 5 for Kerr (r, cos th, ph, p_r, p_mu = -p_th / sin th). */
int prr_debug_state_size(int32_t metric);
/* Right-hand side f(y) of the integrator state in FP64 (fp32 = 0) or FP32 (fp32 = 1). */
prr_status prr_debug_rhs(int32_t metric, double M, double a, double E, double L, const double *y, int32_t fp32,
                         double *f);
/* Reverse mode of the Kerr right-hand side used by the adjoint kernel:
 * yb[5] += fb . df/dy, cb[3] += fb . df/d(E, L, a).  Kerr only. */
prr_status prr_debug_rhs_vjp(int32_t metric, double M, double a, double E, double L, const double *y,
                             const double *fb, double *yb, double *cb);
/* Adjoint pieces (Kerr).  State z = (r, cos th, phi, p_r, p_mu, E, L).
 * Ray set-up of an orbiting camera at screen position (u, v) and its
 * Jacobian dz/d(a, r_cam, incl) (7 x 3, row-major). */
prr_status prr_debug_orbit_init(double M, double a, double r_cam, double incl, double cam_phi, double vfov,
                                double aspect, double u, double v, double z[7], double dz[21]);
/* One FP64 RKF45 step of fixed size h, and its transpose Jacobian product:
 * lam[8] (adjoint of (y', E, L, a)) -> adjoint of (y, E, L, a). */
prr_status prr_debug_kerr_step(double M, double a, double E, double L, const double y[5], double h, double out[5]);
prr_status prr_debug_kerr_step_vjp(double M, double a, double E, double L, const double y[5], double h,
                                   double lam[8]);
/* Bilinear sky colour of an escaping Kerr ray and its Jacobian
 * dc/d(y[5], E, L, a) (3 x 8, row-major); rgb8 is a w x h texture. */
prr_status prr_debug_escape_color(double M, double a, double E, double L, const double y[5], const uint8_t *rgb8,
                                  int w, int h, double c[3], double dc[24]);
/* (x, u) -> integrator state y and (E, L) (initRays' output in the metric's coordinates). */
prr_status prr_debug_state_from_xu(int32_t metric, double M, double a, const double x[4], const double u[4],
                                   double *y, double EL[2]);
/* Integrator state -> (x, u). */
prr_status prr_debug_xu_from_state(int32_t metric, double M, double a, double E, double L, const double *y,
                                   double x[4], double u[4]);
/* initRays for a camera position and unit direction. */
prr_status prr_debug_init_ray(int32_t metric, double M, double a, const double pos[3], const double d[3], double *y,
                              double EL[2]);
/* Direction of ray `local` of a frame (createRays_* + rayDir_*). */
prr_status prr_debug_camera_ray(const prr_frame *f, int64_t local, double d[3]);

#ifdef __cplusplus
}
#endif

#endif /* PRR_DEBUG_H */
