// kernels.hpp -- host-side launchers of the gfx950 kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace prr {

// Ray set-up (k_init) and shading (k_shade), one thread per ray of the job.
hipError_t launch_init(int metric, const JobDev &job, hipStream_t st);
hipError_t launch_shade(int metric, const JobDev &job, hipStream_t st);
// Persistent integrator over job.total rays (or replay tasks); grid/block in workgroups/threads.
// job.adj.mode picks the instantiation: renders (ADJ_NONE) carry no adjoint
// code; the Kerr adjoint's checkpointing passes and its replay each carry
// their own path (other metrics have no adjoint: hipErrorInvalidValue).
hipError_t launch_integrate(int metric, const JobDev &job, int grid, int block, hipStream_t st);
// Waves per SIMD that the render instantiation of (metric, solver) is compiled for.
int integrate_waves(int metric, int solver);
// Zero the device counters and the ray-queue head (all: also the adjoint's mismatch count).
hipError_t launch_reset(CountersDev *c, hipStream_t st, bool all = true);
// SSAA combine for nframes frames (pfs is a device array).
hipError_t launch_ssaa(const PostFrameDev *pfs, int nframes, int64_t max_pixels, hipStream_t st);
// FXAA + RGB8 for the frames whose fxaa flag is set.
hipError_t launch_fxaa(const PostFrameDev *pfs, int nframes, int max_w, int max_h, hipStream_t st);
// Kerr adjoint: seed at the terminal events and reverse sweep over one chunk of
// segment levels; the set-up Jacobian.
hipError_t launch_adj_reverse(const JobDev &job, const AdjDev &ad, hipStream_t st);
hipError_t launch_adj_init_grad(const JobDev &job, const AdjDev &ad, hipStream_t st);
// Adjoint planning: segment counts, checkpoint offsets, the overflow rays and
// the segment-count histogram (clear + count); the order by decreasing
// segment count (scatter, from the host's start positions in cursor); the
// replay tasks of every level (tasks).
hipError_t launch_adj_plan_count(const AdjPlanDev &p, hipStream_t st);
hipError_t launch_adj_plan_scatter(const AdjPlanDev &p, hipStream_t st);
hipError_t launch_adj_plan_tasks(const AdjPlanDev &p, hipStream_t st);
// Deterministic sum of n per-ray gradients (4 doubles each) into out[4];
// partial: 4 * 256 doubles of scratch.
hipError_t launch_grad_reduce(const double *grad_ray, int64_t n, double *partial, double *out, hipStream_t st);

} // namespace prr
