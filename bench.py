"""Benchmark: Kerr RKF45 1080p rays/s (FP32-bulk batched and plain FP64), one MI355X.

The C1 workload: Kerr a = 0.95 (Boyer-Lindquist), adaptive RKF45 at
rtol = atol = 1e-8, 1920x1080 with 2x2 regular SSAA (8,294,400 primary rays),
a textured sphere at r = 8, then the SSAA combine and FXAA + RGB8; inputs
(textures, frame table) are resident in HBM and the images are written to
device buffers.  A step is one batch of BATCH such frames rendered by one
persistent launch replayed from a hipGraph, in the FP32-bulk form; `value` is
its primary rays per second.  The plain form (FP64 throughout, one launch
per frame) is timed in the same run.

roofline: the dominant kernel is the persistent integrator, bound by vector
ALU throughput (no dense contraction: no MFMA).  Its algorithmic FLOP count
comes from the kernel's own attempt counters times the per-attempt model of
capi.cpp (FP32 stage flops and FP64 flops kept apart); `achieved` divides by
the kernel's duration measured with HIP events on its own stream.  frac =
(F32/P32 + F64/P64) / t against MI355X's published vector peaks
(157.3 TFLOP/s FP32, 78.6 TFLOP/s FP64); frac_nonpacked uses 78.6 TFLOP/s for
FP32 as well, the rate of the non-packed FP32 instructions the kernels are compiled to.

cpu_baseline: the CPU FP64 oracle (oracle/, a restatement of the reference
solver) on a stratified row subset of the same frame, all host cores, OpenMP;
beside it the same solver in FP32, the reference's own precision
(fp32_reference_like, on a seeded sample of those rays).
The same leg times the oracle on the whole C0 frame (the reference's
CPU-runnable case, beside the GPU's C0 leg) and, as the checker only, runs
the oracle at tol 1e-12 on seeded subsets of the rays the GPU legs rendered:
the `accuracy` block (per config and form: class agreement with the rays
within 1e-6 rad of a class boundary excused, and the endpoint-angle error).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

P32, P64 = 157.3e12, 78.6e12
P32_NONPACKED = 78.6e12
BATCH = 4
# HBM bytes per launch on one C1 frame (profiles/k_integrate_kerr_c1_1080p.md: FETCH_SIZE x 2 + WRITE_SIZE,
# calibrated on a 1 GiB copy): read + written
TRAFFIC_C1 = {"fp64": (884.5 + 741.9) * 1e6, "fp32": (875.7 + 744.8) * 1e6}
METRIC = "Kerr RKF45 1080p rays/s + adjoint rays/s, FP32-bulk vs FP64 at ≤1e-6 rad"


def subset(n, k, seed):
    """A seeded uniform subset of k of n ray ids, sorted."""
    import numpy as np
    return np.sort(np.random.default_rng(seed).choice(n, size=min(k, n), replace=False))


def gpu_rays(ctx, frames, ids, torch, base=0):
    """Per-ray class and endpoint of rays base + ids of one more (untimed) render of frames."""
    total = sum(f.nrays for f in frames)
    rays = dict(cls=torch.empty(total, dtype=torch.int8, device="cuda"),
                endpoint=torch.empty((total, 3), dtype=torch.float64, device="cuda"))
    ctx.render_device(frames, rays=rays)
    idx = torch.from_numpy(base + ids).cuda()
    out = {k: v.index_select(0, idx).cpu().numpy() for k, v in rays.items()}
    del rays
    torch.cuda.synchronize()
    return out


def run_form(ctx, frame, precision, steps, warmup, out8, outf, torch):
    from parrelray_amd import _abi as A
    from parrelray_amd import configs as CF
    ctx.set_tracer(CF.tracer_kerr(precision=precision))
    for _ in range(warmup):
        ctx.render_device([frame], rgb8=[out8], rgbf=[outf])
    torch.cuda.synchronize()
    per = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.render_device([frame], rgb8=[out8], rgbf=[outf])
        per.append(ctx.stats())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / steps
    integ = sum(s["integrate_ms"] for s in per) / steps
    st = per[-1]
    f32, f64 = st["flops_fp32"], st["flops_fp64"]
    t_int = integ / 1e3
    return dict(ms_per_step=ms, integrate_ms=integ, ssaa_ms=sum(s["ssaa_ms"] for s in per) / steps,
                init_ms=sum(s["init_ms"] for s in per) / steps, shade_ms=sum(s["shade_ms"] for s in per) / steps,
                fxaa_ms=sum(s["fxaa_ms"] for s in per) / steps, rays=st["rays"], attempts=st["attempts"],
                accepted=st["accepted"], attempts_fp64=st["attempts_fp64"], flops_fp32=f32, flops_fp64=f64,
                rays_per_s=st["rays"] / (ms / 1e3), attempts_per_s=st["attempts"] / t_int,
                accepted_per_s=st["accepted"] / t_int,
                achieved_tflops=(f32 + f64) / t_int / 1e12,
                roofline_frac=(f32 / P32 + f64 / P64) / t_int,
                ssaa_gbs=gbs(st["ssaa_bytes"], sum(s["ssaa_ms"] for s in per) / steps),
                fxaa_gbs=gbs(st["fxaa_bytes"], sum(s["fxaa_ms"] for s in per) / steps))


def gbs(nbytes, ms):
    """GB/s of a pass that moved nbytes in ms (None when the pass did not run)."""
    return nbytes / (ms / 1e3) / 1e9 if nbytes > 0 and ms > 0 else None


def run_adjoint(api, A, CF, reps, torch=None, ids=None, warmup=2):
    """C3: 512x512 Kerr a=0.7, FP32-bulk RKF45 tol 1e-8; gradient of the L2 loss against the target
    render w.r.t. (a, r_cam, incl, sphere phi).  Rays/s counts primary rays through the whole
    prr_adjoint call (forward render, checkpointed forward, replay and reverse sweep).  The FP32-bulk
    form's stage precision is autotuned for the adjoint call (the faster of FP32 and FP64 stages),
    as the render legs' is."""
    import numpy as np
    from parrelray_amd import autotune
    ctx = api.Context(0)
    ctx.set_spacetime(A.PRR_KERR_BL, 1.0, 0.7)
    ctx.set_tracer(CF.tracer_kerr(precision=A.PRR_FP32_BULK))
    ctx.set_texture(0, CF.sky_texture(), A.PRR_FILTER_BILINEAR)
    ctx.set_texture(1, CF.sphere_texture(), A.PRR_FILTER_BILINEAR)
    tmpl = CF.frame_c3()
    o0, ot = CF.c3_orbit(), CF.c3_orbit(target=True)
    T = ctx.render(api.frame_from_orbit(tmpl, ot))[0]["rgbf"]
    I = ctx.render(api.frame_from_orbit(tmpl, o0))[0]["rgbf"]
    dLdI = (2.0 * (I - T) / I.size).astype(np.float32)
    stages = autotune.tune_adjoint_precision(ctx, tmpl, o0, dLdI, CF.tracer_kerr(precision=A.PRR_FP32_BULK))
    for _ in range(max(1, warmup)):
        g = ctx.adjoint(tmpl, o0, dLdI)
    per = []
    for _ in range(reps):
        t0 = time.perf_counter()
        g = ctx.adjoint(tmpl, o0, dLdI)
        per.append(time.perf_counter() - t0)
    sec = sum(per) / reps
    st = ctx.stats()
    out = dict(rays=tmpl.nrays, seconds=sec, rays_per_s=tmpl.nrays / sec, call_ms=[round(x * 1e3, 3) for x in per],
               stage_precision="fp64" if stages["precision"] == A.PRR_FP64 else "fp32", stage_tuning=stages,
               median_call_ms=sorted(per)[len(per) // 2] * 1e3,
               adjoint_ms=st["adjoint_ms"],
               total_ms=st["total_ms"], steps_swept=st["adjoint_steps"],
               steps_per_s=st["adjoint_steps"] / (st["adjoint_ms"] / 1e3),
               grad={"a": g[0], "r_cam": g[1], "incl": g[2], "sphere_phi": g[3]},
               loss=float(((I - T) ** 2).mean()))
    fd_path = os.path.join(ROOT, "tests", "golden", "c3_fd_gradient.npz")
    if os.path.exists(fd_path):
        # against the committed FP64 finite-difference reference (oracle at tol 1e-12 with FP64
        # colours, central differences at h = 1e-6, 5e-7, 2.5e-7 of the same contraction w . I)
        z = np.load(fd_path)
        nkept = int(z["size"]) ** 2 - z["excluded_class"].size - z["excluded_fd"].size
        w = z["s16"].astype(np.float32) * np.float32(2.0 / (3 * nkept))
        gw = ctx.adjoint(tmpl, o0, w)
        fd = z["fd"][-1]
        names = ("a", "r_cam", "incl", "sphere_phi")
        out["vs_fd"] = dict(
            reference="FP64 oracle, tol 1e-12, FP64 colours; central differences, h = 1.25e-7; pixels whose "
                      "class changes or whose own difference quotients do not converge left out of both sides",
            adjoint={n: float(v) for n, v in zip(names, gw)}, fd={n: float(v) for n, v in zip(names, fd)},
            rel_err={n: float(abs(a - b) / abs(b)) for n, a, b in zip(names, gw, fd)},
            fd_spread_rel={n: float(np.abs(z["fd"][:, k] - fd[k]).max() / abs(fd[k])) for k, n in enumerate(names)})
        # the same contraction with the integration at tol 1e-10: the workload's tol 1e-8 moves the
        # gradient of the integrated map by about 1e-4 of r_cam; with that error taken out the
        # adjoint meets the finite differences of the converged map to a few 1e-6
        tuned = A.PRR_FP64 if stages["precision"] == A.PRR_FP64 else A.PRR_FP32_BULK
        ctx.set_tracer(CF.tracer_kerr(precision=tuned, tol=1e-10))
        g10 = ctx.adjoint(tmpl, o0, w)
        out["vs_fd"]["tol_1e-10"] = dict(adjoint={n: float(v) for n, v in zip(names, g10)},
                                         rel_err={n: float(abs(a - b) / abs(b)) for n, a, b in zip(names, g10, fd)})
        ctx.set_tracer(CF.tracer_kerr(precision=tuned))
    if ids is not None:
        out["_rays"] = gpu_rays(ctx, [api.frame_from_orbit(tmpl, o0)], ids, torch)
    return out


def run_c1_batched(api, A, CF, ctx, steps, warmup, torch, precision=None, tune_precision=True):
    """BATCH C1 frames in one persistent launch from a hipGraph (FP32-bulk unless precision says
    otherwise), with the autotuned launch shape and, for FP32-bulk, the autotuned stage precision
    (tune_precision: the faster of FP32 and FP64 stages; False keeps the FP32 stages)."""
    from parrelray_amd import autotune
    frames = [CF.frame_c1() for _ in range(BATCH)]
    tracer = CF.tracer_kerr(precision=A.PRR_FP32_BULK if precision is None else precision)
    ctx.set_tracer(tracer)
    shape = autotune.tune_launch(ctx, frames[:1], A.PRR_KERR_BL, tracer, use_cache=False,
                                 tune_precision=tune_precision)
    rgb8 = [torch.empty((f.height, f.width, 3), dtype=torch.uint8, device="cuda") for f in frames]
    g = api.Graph(ctx, frames, rgb8=rgb8)
    for _ in range(warmup):
        g.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.launch()
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / steps
    st = ctx.stats()
    g.close()
    ctx.set_launch()  # back to the default shape for the other legs
    return dict(frames=BATCH, rays=st["rays"], ms_per_step=sec * 1e3, rays_per_s=st["rays"] / sec, launch=shape,
                ms_per_frame=sec * 1e3 / BATCH, attempts=st["attempts"], attempts_fp64=st["attempts_fp64"],
                graph_ms=st["total_ms"],
                stage_precision="fp64" if shape.get("precision", tracer.precision) == A.PRR_FP64 else "fp32")


def run_c4(api, A, CF, torch, steps, ids):
    """C4: Kerr a = 0.998, camera r = 6M edge-on, 2048x2048 with 4x4 SSAA (67,108,864 rays) + FXAA,
    tol 1e-10, one frame per step: plain FP64, the FP32-bulk form with FP32 stages, and the FP32-bulk
    form with the autotuned launch shape and stage precision; plus the per-ray results of a subset
    in the first two."""
    from parrelray_amd import autotune
    f = CF.frame_c4()
    out = {"rays_per_frame": f.nrays}
    img = torch.empty((f.height, f.width, 3), dtype=torch.uint8, device="cuda")
    for name, prec in (("fp64", A.PRR_FP64), ("fp32_bulk", A.PRR_FP32_BULK), ("fp32_bulk_tuned", A.PRR_FP32_BULK)):
        ctx = api.Context(0)
        ctx.set_spacetime(A.PRR_KERR_BL, 1.0, 0.998)
        ctx.set_tracer(CF.tracer_c4(precision=prec))
        ctx.set_texture(0, CF.sky_texture())
        if name == "fp32_bulk_tuned":
            # the FP32-bulk form with the autotuned launch shape and stage
            # precision, tuned on a 512 x 512 frame of the same scene
            out["tuned_launch"] = autotune.tune_launch(ctx, [CF.frame_c4(512, 512)], A.PRR_KERR_BL,
                                                       CF.tracer_c4(precision=prec), use_cache=False,
                                                       budget_s=20.0)
        ctx.render_device([f], rgb8=[img])  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        per = []
        for _ in range(steps):
            ctx.render_device([f], rgb8=[img])
            per.append(ctx.stats())
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        st = per[-1]
        integ = sum(p["integrate_ms"] for p in per) / steps
        out[name] = dict(ms_per_frame=ms, rays_per_s=f.nrays / (ms / 1e3), integrate_ms=integ,
                         attempts=st["attempts"], attempts_fp64=st["attempts_fp64"],
                         attempts_per_s=st["attempts"] / (integ / 1e3),
                         roofline_frac=(st["flops_fp32"] / P32 + st["flops_fp64"] / P64) / (integ / 1e3))
        if name != "fp32_bulk_tuned":
            out[name]["_rays"] = gpu_rays(ctx, [f], ids, torch)
        ctx.close()
    out["speedup_fp32_bulk_vs_fp64"] = out["fp64"]["ms_per_frame"] / out["fp32_bulk"]["ms_per_frame"]
    out["speedup_fp32_bulk_tuned_vs_fp64"] = out["fp64"]["ms_per_frame"] / out["fp32_bulk_tuned"]["ms_per_frame"]
    return out


def run_c0(api, A, CF, torch, steps):
    """C0, the reference's CPU-runnable case: Schwarzschild (Schwarzschild coordinates), fixed-step
    Euler h = 0.01 M in FP64, 256x256, camera r = 30M at 85 deg, 10,000-step cap; per-ray results of
    all its rays for the checker."""
    import numpy as np
    f = CF.frame_c0()
    ctx = api.Context(0)
    ctx.set_spacetime(A.PRR_SCHW_SPH, 1.0, 0.0)
    ctx.set_tracer(CF.tracer_c0())
    ctx.set_texture(0, CF.sky_texture())
    img = torch.empty((f.height, f.width, 3), dtype=torch.uint8, device="cuda")
    ctx.render_device([f], rgb8=[img])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.render_device([f], rgb8=[img])
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    st = ctx.stats()
    out = dict(rays=f.nrays, ms_per_frame=ms, rays_per_s=f.nrays / (ms / 1e3), integrate_ms=st["integrate_ms"],
               euler_steps=st["accepted"], steps_per_s=st["accepted"] / (st["integrate_ms"] / 1e3))
    out["_rays"] = gpu_rays(ctx, [f], np.arange(f.nrays), torch)
    ctx.close()
    return out


def run_copy(torch, nbytes=1 << 30):
    """A plain device-to-device copy of nbytes: the calibration kernel for the
    FETCH_SIZE / WRITE_SIZE counters (it reads and writes exactly nbytes)."""
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    del src, dst
    return dict(bytes_read=nbytes, bytes_written=nbytes, ms=ms, gbs=2 * nbytes / ms / 1e6)


def run_batch(api, A, CF, reps, torch):
    """C2: 64 frames (389,520,384 rays) in one persistent launch, replayed from a hipGraph, with the
    launch shape (workgroup size, workgroups per CU, refill threshold) from the autotuner."""
    from parrelray_amd import autotune
    frames = CF.frames_c2()
    ctx = api.Context(0)
    ctx.set_spacetime(A.PRR_KERR_BL, 1.0, 0.0)
    ctx.set_tracer(ctx_tracer(A, CF))
    ctx.set_texture(0, CF.sky_texture())
    ctx.set_texture(1, CF.sphere_texture(), A.PRR_FILTER_BILINEAR)
    # launch shape from the autotuner, on a quarter-resolution copy of the batch
    shape = autotune.tune_launch(ctx, CF.frames_c2(scale=0.25), A.PRR_KERR_BL, ctx_tracer(A, CF), use_cache=False)
    rgb8 = [torch.empty((f.height, f.width, 3), dtype=torch.uint8, device="cuda") for f in frames]
    g = api.Graph(ctx, frames, rgb8=rgb8)
    g.launch()  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        g.launch()
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / reps
    st = ctx.stats()
    g.close()
    return dict(frames=len(frames), rays=st["rays"], seconds=sec, rays_per_s=st["rays"] / sec,
                attempts=st["attempts"], attempts_fp64=st["attempts_fp64"], graph_ms=st["total_ms"], launch=shape)


def ctx_tracer(A, CF):
    """The C2 tracer: Kerr RKF45 tol 1e-8, FP32-bulk."""
    return CF.tracer_kerr(precision=A.PRR_FP32_BULK)


def cpu_model():
    """The host CPU's model name (first processor in /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def angle(a, b):
    import numpy as np
    a = a / np.linalg.norm(a, axis=-1, keepdims=True)
    b = b / np.linalg.norm(b, axis=-1, keepdims=True)
    return np.arctan2(np.linalg.norm(np.cross(a, b), axis=-1), np.clip(np.sum(a * b, axis=-1), -1.0, 1.0))


def check_rays(S, frame, ids, orc, gpu, cores):
    """The GPU's per-ray class and endpoint against the oracle's (orc) on rays ids of frame: class
    mismatches, those not within 1e-6 rad of a class boundary (none of the oracle's four +-1e-6 rad
    perturbations of the ray changes its class), and the endpoint angle of agreeing escape / sphere rays."""
    import numpy as np

    from parrelray_amd import _abi as A
    same = gpu["cls"] == orc["cls"]
    bad = np.nonzero(~same)[0]
    excused = np.zeros(bad.size, bool)
    if bad.size:
        for axis in (1, 2, 3, 4):
            r = S.render_rays(frame, ids[bad], delta=1e-6, axis=axis, nthreads=cores)
            excused |= r["cls"] != orc["cls"][bad]
    m = same & ((orc["cls"] == A.PRR_ESCAPE) | (orc["cls"] == A.PRR_SPHERE))
    ang = angle(gpu["endpoint"][m], orc["endpoint"][m])
    st = {}
    if ang.size:
        st = dict(max=float(ang.max()), mean=float(ang.mean()), median=float(np.median(ang)),
                  p99=float(np.quantile(ang, 0.99)), frac_le_1e6=float((ang <= 1e-6).mean()), compared=int(ang.size))
    return dict(rays=int(ids.size), class_mismatch=int(bad.size), class_mismatch_unexcused=int((~excused).sum()),
                endpoint_angle_rad=st)


def cpu_leg(frame, c0=None, checks=()):
    """The CPU leg: the FP64 oracle timed on every 16th row of the C1 frame (cpu_baseline) and on the
    whole C0 frame, then, as the checker only, the oracle at tol 1e-12 on the rays the GPU legs kept.
    checks: (name, frame, oracle tracer, ray ids, {form: gpu rays}) per config."""
    import numpy as np

    import oracle as O
    from parrelray_amd import _abi as A
    from parrelray_amd import configs as CF
    cores = len(os.sched_getaffinity(0))
    base = cpu_baseline(frame, cores=cores)
    accuracy = {}
    if c0 is not None:
        f0 = CF.frame_c0()
        S0 = O.Scene(A.PRR_SCHW_SPH, 1.0, 0.0, CF.tracer_c0())
        S0.set_texture(0, CF.sky_texture())
        ids = np.arange(f0.nrays)
        r = S0.render_rays(f0, ids, nthreads=cores)
        c0["cpu"] = dict(rays_per_s=f0.nrays / r["seconds"], seconds=r["seconds"], cores=cores, kind="port",
                         sample="the whole C0 frame (65,536 rays), FP64 Euler h = 0.01")
        c0["speedup_gpu_vs_cpu"] = c0["rays_per_s"] / c0["cpu"]["rays_per_s"]
        S0.set_fp32(True)
        r32 = S0.render_rays(f0, ids, nthreads=cores)
        S0.set_fp32(False)
        c0["cpu_fp32_reference_like"] = dict(
            rays_per_s=f0.nrays / r32["seconds"], seconds=r32["seconds"], cores=cores, kind="port",
            class_agreement_with_fp64=float((r32["cls"] == r["cls"]).mean()),
            sample="the whole C0 frame, Euler h = 0.01 with the state and right-hand side in FP32")
        accuracy["C0"] = {"fp64": check_rays(S0, f0, ids, r, c0.pop("_rays"), cores)}
    for name, f, tracer, ids, forms in checks:
        S = O.Scene(A.PRR_KERR_BL, 1.0, 0.0, tracer)
        t0 = time.perf_counter()
        orc = S.render_rays(f, ids, nthreads=cores)
        accuracy[name] = {"oracle": f"FP64 RKF45 tol {tracer.atol:g}, {ids.size} seeded rays of {f.nrays}",
                          "oracle_seconds": time.perf_counter() - t0}
        for form, g in forms.items():
            accuracy[name][form] = check_rays(S, f, ids, orc, g, cores)
    return base, accuracy


def cpu_baseline(frame, row_stride=16, cores=None, fp32_rays=96):
    """Oracle (FP64, Christoffel form) on every row_stride-th image row, all cores; its FP32 variant
    (the reference's precision) on fp32_rays of them."""
    import numpy as np

    import oracle as O
    from parrelray_amd import _abi as A
    from parrelray_amd import configs as CF
    S = O.Scene(A.PRR_KERR_BL, 1.0, 0.95, CF.tracer_kerr())
    S.set_texture(0, CF.sky_texture())
    S.set_texture(1, CF.sphere_texture(), A.PRR_FILTER_BILINEAR)
    per_row = frame.width * frame.samples
    rows = np.arange(0, frame.height, row_stride)
    ids = (rows[:, None] * per_row + np.arange(per_row)[None, :]).ravel()
    cores = cores or len(os.sched_getaffinity(0))
    r = S.render_rays(frame, ids, nthreads=cores)
    base = dict(value=ids.size / r["seconds"], unit="rays/s", cores=cores, kind="port", model=cpu_model(),
                sample=f"C1 rows j % {row_stride} == 0 ({ids.size} rays of {frame.nrays}), FP64 RKF45 tol 1e-8, "
                       f"{r['seconds']:.1f} s")
    # The reference's own precision: the same solver with the state, the stages, the error norm and
    # the controller in FP32, on a seeded sample of the same rows.  An FP32 state cannot meet
    # rtol = atol = 1e-8 (FP32 rounding of the state is ~6e-8 of it), so most rays spin at the smallest
    # steps until the attempt cap: the rate is reported with the share of rays that end there.
    sub = np.sort(np.random.default_rng(32).choice(ids, size=min(fp32_rays, ids.size), replace=False))
    r64 = {k: v[np.searchsorted(ids, sub)] for k, v in r.items() if k != "seconds"}
    S.set_fp32(True)
    r32 = S.render_rays(frame, sub, nthreads=cores)
    cap = r32["attempts"] >= S.tracer.max_attempts
    base["fp32_reference_like"] = dict(
        value=sub.size / r32["seconds"], unit="rays/s", cores=cores, seconds=r32["seconds"],
        sample=f"{sub.size} seeded rays of the same rows, RKF45 tol 1e-8 in FP32",
        rays_at_attempt_cap=float(cap.mean()), mean_attempts=float(r32["attempts"].mean()),
        mean_attempts_fp64=float(r64["attempts"].mean()),
        class_agreement_with_fp64=float((r32["cls"] == r64["cls"]).mean()))
    return base


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU leg (baseline, C0 on the CPU, accuracy)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the C3 adjoint leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the C2 64-frame batch leg")
    ap.add_argument("--legs", default="fp64,fp32,adjoint,c2,batched,batched32,batched64,c4,c0,cpu",
                    help="comma-separated legs to run: fp64, fp32, adjoint, c2, batched (FP32-bulk form, autotuned "
                         "stage precision), batched32 (the same batch with FP32 stages), batched64 (the same "
                         "batch in plain FP64), c4, c0, cpu, copy "
                         "(copy: a 1 GiB device-to-device copy, the calibration kernel for HBM byte counters)")
    args = ap.parse_args()
    legs = set(args.legs.split(","))
    for flag, leg in ((args.no_cpu, "cpu"), (args.no_adjoint, "adjoint"), (args.no_batch, "c2")):
        if flag:
            legs.discard(leg)
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the integrator has no multi-GPU path "
              f"(--gpus {args.gpus} requested).", file=sys.stderr)
        sys.exit(2)
    import torch

    from parrelray_amd import _abi as A
    from parrelray_amd import api
    from parrelray_amd import configs as CF

    def progress(**kw):
        print(json.dumps({"progress": kw}), file=sys.stderr, flush=True)

    copy = None
    if "copy" in legs:
        copy = run_copy(torch)
        progress(copy=copy)
    frame = CF.frame_c1()
    ids_c1 = subset(frame.nrays, 4000, 101)
    ids_c3 = subset(CF.frame_c3().nrays, 2000, 103)
    ids_c4 = subset(CF.frame_c4().nrays, 4000, 104)
    ctx = api.Context(0)
    ctx.set_spacetime(A.PRR_KERR_BL, 1.0, 0.95)
    ctx.set_texture(0, CF.sky_texture())
    ctx.set_texture(1, CF.sphere_texture(), A.PRR_FILTER_BILINEAR)
    out8 = torch.empty((frame.height, frame.width, 3), dtype=torch.uint8, device="cuda")
    outf = torch.empty((frame.height, frame.width, 3), dtype=torch.float32, device="cuda")
    fp64 = fp32 = adj = batch = bat = bat32 = bat64 = c4 = c0 = None
    c1_rays = {}
    if "fp64" in legs:
        fp64 = run_form(ctx, frame, A.PRR_FP64, args.steps, args.warmup, out8, outf, torch)
        c1_rays["fp64"] = gpu_rays(ctx, [frame], ids_c1, torch)
        progress(fp64_ms_per_step=fp64["ms_per_step"])
    if "fp32" in legs:
        fp32 = run_form(ctx, frame, A.PRR_FP32_BULK, args.steps, args.warmup, out8, outf, torch)
        c1_rays["fp32_bulk"] = gpu_rays(ctx, [frame], ids_c1, torch)
        progress(fp32_bulk_ms_per_step=fp32["ms_per_step"])
    if "adjoint" in legs:
        adj = run_adjoint(api, A, CF, max(1, args.steps // 2), torch, ids_c3, warmup=args.warmup)
        progress(adjoint={k: v for k, v in adj.items() if not k.startswith("_")})
    if "c4" in legs:
        c4 = run_c4(api, A, CF, torch, 2, ids_c4)
        progress(c4={k: v for k, v in c4.items() if k.startswith("speedup")})
    if "c0" in legs:
        c0 = run_c0(api, A, CF, torch, args.steps)
        progress(c0={k: v for k, v in c0.items() if not k.startswith("_")})
    if "c2" in legs:
        batch = run_batch(api, A, CF, 2, torch)
        progress(batch=batch)
    if "batched" in legs:
        bat = run_c1_batched(api, A, CF, ctx, args.steps, args.warmup, torch)
        progress(c1_batched_ms_per_frame=bat["ms_per_frame"])
    if "batched32" in legs:
        bat32 = run_c1_batched(api, A, CF, ctx, args.steps, args.warmup, torch, tune_precision=False)
        progress(c1_batched32_ms_per_frame=bat32["ms_per_frame"])
    if "batched64" in legs:
        bat64 = run_c1_batched(api, A, CF, ctx, args.steps, args.warmup, torch, precision=A.PRR_FP64)
        progress(c1_batched64_ms_per_frame=bat64["ms_per_frame"])
    cpu = accuracy = None
    if "cpu" in legs:
        checks = []
        if c1_rays:
            checks.append(("C1", frame, CF.tracer_kerr(tol=1e-12), ids_c1, c1_rays))
        if adj and "_rays" in adj:
            checks.append(("C3", api.frame_from_orbit(CF.frame_c3(), CF.c3_orbit()), CF.tracer_kerr(tol=1e-12),
                           ids_c3, {"fp32_bulk": adj.pop("_rays")}))
        if c4:
            checks.append(("C4", CF.frame_c4(), CF.tracer_kerr(tol=1e-12, max_steps=400000), ids_c4,
                           {k: c4[k].pop("_rays") for k in ("fp64", "fp32_bulk")}))
        cpu, accuracy = cpu_leg(frame, c0, checks)
        progress(accuracy=accuracy)
    for leg in (adj, c0):
        if leg:
            leg.pop("_rays", None)
    if c4:
        for k in ("fp64", "fp32_bulk"):
            c4[k].pop("_rays", None)
    roof = None
    # the dominant kernel: the integrator instantiation the headline runs (the FP32-bulk form's
    # autotuned stage precision), timed one C1 frame per launch by its own leg
    stages = bat["stage_precision"] if bat else ("fp32" if fp32 else "fp64")
    leg = fp64 if stages == "fp64" else fp32
    if leg:
        t_int = leg["integrate_ms"] / 1e3
        f32, f64 = leg["flops_fp32"], leg["flops_fp64"]
        t_ideal = f32 / P32 + f64 / P64
        t_ideal_np = f32 / P32_NONPACKED + f64 / P64
        roof = {"bound": "valu",
                "kernel": f"k_integrate<Kerr,RKF45,{'FP64' if stages == 'fp64' else 'FP32-bulk'} stages>, "
                          "one C1 frame per launch",
                "achieved": (f32 + f64) / t_int / 1e12, "peak": (f32 + f64) / t_ideal / 1e12,
                "unit": "TFLOP/s", "frac": t_ideal / t_int, "frac_nonpacked": t_ideal_np / t_int,
                # HBM bytes per launch of this kernel on this frame: FETCH_SIZE x 2 + WRITE_SIZE
                # (calibrated on a 1 GiB copy), from the rocprofv3 --pmc passes of
                # profiles/k_integrate_kerr_c1_1080p.md; a constant of the kernel and workload
                "traffic": TRAFFIC_C1[stages], "traffic_unit": "bytes",
                "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, profiles/k_integrate_kerr_c1_1080p.md",
                "flops_fp32_per_launch": f32, "flops_fp64_per_launch": f64,
                "launch_ms": leg["integrate_ms"]}
    head = bat or fp32 or fp64
    line = {
        "metric": METRIC,
        "value": head["rays_per_s"] if head else None,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"] if head else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-bulk/fp64",
        "data": "synthetic (seeded checker sky 512x256, value-noise sphere texture 1024x512)",
        "config": {"workload": (f"C1 x{BATCH} per step: " if bat else "C1, one frame per step: ")
                   + "Kerr a=0.95 BL, RKF45 rtol=atol=1e-8, 1920x1080, 2x2 SSAA + FXAA, camera r=20M i=80deg, "
                   "sphere R=1M at r=8M; " + (f"FP32-bulk form, one persistent launch per batch from a hipGraph, "
                                              f"autotuned launch shape and stage precision "
                                              f"({bat['stage_precision'].upper()} stages chosen)"
                                              if bat else ("FP32-bulk" if fp32 else "plain FP64")),
                   "rays_per_frame": frame.nrays, "frames_per_step": BATCH if bat else 1},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if fp64:
        line["fp64_plain"] = {k: fp64[k] for k in ("rays_per_s", "ms_per_step", "integrate_ms", "attempts",
                                                   "attempts_per_s", "accepted_per_s", "achieved_tflops",
                                                   "roofline_frac")}
    if fp32:
        line["fp32_bulk_single"] = {k: fp32[k] for k in ("rays_per_s", "ms_per_step", "integrate_ms", "init_ms",
                                                         "shade_ms", "ssaa_ms", "fxaa_ms", "attempts", "accepted",
                                                         "attempts_fp64", "attempts_per_s", "accepted_per_s",
                                                         "achieved_tflops",
                                                         "roofline_frac", "ssaa_gbs", "fxaa_gbs")}
    if bat:
        line["fp32_bulk_batched"] = bat
        if fp64:
            line["speedup_batched_fp32_bulk_vs_plain_fp64"] = bat["rays_per_s"] / fp64["rays_per_s"]
            line["target_3x_fp32_bulk_vs_plain_fp64"] = "missed" if bat["rays_per_s"] < 3 * fp64["rays_per_s"] \
                else "met"
    if bat32:
        line["fp32_stages_batched"] = bat32
        if fp64:
            line["speedup_batched_fp32_stages_vs_plain_fp64"] = bat32["rays_per_s"] / fp64["rays_per_s"]
    if bat64:
        line["fp64_batched"] = bat64
        if bat:
            line["speedup_batched_fp32_bulk_vs_batched_fp64"] = bat["rays_per_s"] / bat64["rays_per_s"]
    if adj:
        line["adjoint_rays_per_s"] = adj["rays_per_s"]
        line["adjoint"] = adj
    if c4:
        line["c4"] = c4
    if c0:
        line["c0"] = c0
    if batch:
        line["batch_c2"] = batch
    if accuracy:
        line["accuracy"] = accuracy
    if copy:
        line["copy_calibration"] = copy
    print(json.dumps(line))


if __name__ == "__main__":
    main()
