// frame_host.hpp -- host-side construction of the kernels' per-frame and
// tracer constants (shared by the C ABI and the host evaluation hooks).
#pragma once

#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace prr {

// Horizon radius r_h of the metric (0 for flat space).
double horizon_radius(int metric, double M, double a);
// prr_tracer -> TracerDev.
TracerDev tracer_dev(const prr_tracer &t);
// prr_frame -> FrameDev (Camera::computeVectors, main/Camera.cpp:19-28, and
// the horizon test radius r_h (1 + eps_rel) + eps_abs); begin = first ray id.
// param: the metric's second parameter when it is not the frame's spin: the
// charge of the Schwarzschild-coordinate metric, the Ellis throat radius.
// r_esc2: the escape radius beyond the wormhole's throat (<= 0: r_escape).
FrameDev frame_dev(int metric, const prr_tracer &tr, const prr_frame &f, int64_t begin, double param = 0.0,
                   double r_esc2 = 0.0);

} // namespace prr
