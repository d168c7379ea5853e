// kernels.hpp -- host-side launchers of the gfx950 kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace prr {

// Persistent integrator over job.total rays; grid/block in workgroups/threads.
hipError_t launch_integrate(int metric, const JobDev &job, int grid, int block, hipStream_t st);
// Zero the device counters and the ray-queue head.
hipError_t launch_reset(CountersDev *c, hipStream_t st);
// SSAA combine for nframes frames (pfs is a device array).
hipError_t launch_ssaa(const PostFrameDev *pfs, int nframes, int64_t max_pixels, hipStream_t st);
// FXAA + RGB8 for the frames whose fxaa flag is set.
hipError_t launch_fxaa(const PostFrameDev *pfs, int nframes, int max_w, int max_h, hipStream_t st);
// Kerr adjoint: pass 2 (checkpointed forward + seeds) or pass 3 (replay + reverse sweep).
hipError_t launch_adjoint(int pass, int precision, const JobDev &job, const AdjDev &ad, int grid, int block,
                          hipStream_t st);
// Deterministic sum of n per-ray gradients (4 doubles each) into out[4].
hipError_t launch_grad_reduce(const double *grad_ray, int64_t n, double *out, hipStream_t st);

} // namespace prr
