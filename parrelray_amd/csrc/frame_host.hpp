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
// charge: the spacetime's q, used by the Schwarzschild-coordinate metric (Reissner-Nordstrom).
FrameDev frame_dev(int metric, const prr_tracer &tr, const prr_frame &f, int64_t begin, double charge = 0.0);

} // namespace prr
