"""Autotuner for the persistent integrator's launch shape.

Searches workgroup size, resident workgroups per CU and the wave refill
threshold on a short calibration workload (a scaled-down copy of the batch
to be rendered), timing the integrator kernel with the library's own HIP
events, and the adjoint's checkpoint interval on a small adjoint problem.
Results are cached per (device, metric, solver, precision) in a JSON file.

Every launch shape yields bit-identical images (the per-ray work does not
depend on which lanes share a wave), so tuning only changes speed; the
tests check that invariance.

For an FP32-bulk tracer it also times the plain FP64 stages at the tuned
shape and keeps the faster of the two (tune_precision; for the adjoint,
tune_adjoint_precision): both meet the FP32-bulk form's accuracy contract,
and on devices whose non-packed FP32 instructions issue at the FP64 rate
(MI355X) the FP64 stages need fewer instructions than the FP32 difference
stages with their conversions.
"""
import json
import os
import time

from . import _abi as A

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "autotune.json")

BLOCKS = (64, 128, 256)
BLOCKS_PER_CU = (1, 2, 3, 4, 6, 8)
REFILLS = (8, 16, 32, 48, 56, 64)
CHECKPOINTS = (16, 32, 64, 128)


def _key(device_name, metric, solver, precision):
    return f"{device_name}|m{metric}|s{solver}|p{precision}"


def _load():
    try:
        with open(CACHE) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _save(d):
    os.makedirs(os.path.dirname(CACHE), exist_ok=True)
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(d, fh, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


def _time_integrate(ctx, frames, reps=2):
    best = float("inf")
    for _ in range(reps):
        ctx.render(frames, rgbf=False)
        best = min(best, ctx.stats()["integrate_ms"])
    return best


def tune_launch(ctx, frames, metric, tracer, device_name="gpu0", budget_s=60.0, use_cache=True, tune_precision=True):
    """Pick (block, blocks_per_cu, refill) for rendering `frames` and, for an FP32-bulk tracer, the
    stage precision (tune_precision); sets them on ctx (launch shape, tracer) and returns them."""
    cache = _load()
    key = _key(device_name, metric, tracer.solver, tracer.precision) + ("" if tune_precision else "|stages")
    if use_cache and key in cache:
        best = cache[key]
        ctx.set_launch(best["block"], best["blocks_per_cu"], best["refill"], best.get("checkpoint", 64))
        if best.get("precision", tracer.precision) != tracer.precision:
            t = A.Tracer.from_buffer_copy(tracer)
            t.precision = best["precision"]
            ctx.set_tracer(t)
        return best
    start = time.perf_counter()
    results = []
    # coarse: occupancy-like dimension first at the default refill (whole-wave
    # refills, which keep neighbouring rays together), then refill at the best shape
    for block in BLOCKS:
        for bpc in BLOCKS_PER_CU:
            waves = block // 64 * bpc
            if waves > 32:
                continue
            ctx.set_launch(block, bpc, 64, 64)
            results.append((_time_integrate(ctx, frames), block, bpc, 64))
            if time.perf_counter() - start > budget_s / 2:
                break
    results.sort()
    _, block, bpc, _ = results[0]
    # FP32-bulk always refills whole waves (its stage precision is chosen per
    # wave), so the refill threshold is not a dimension there
    for refill in (REFILLS if tracer.precision != A.PRR_FP32_BULK else ()):
        ctx.set_launch(block, bpc, refill, 64)
        results.append((_time_integrate(ctx, frames), block, bpc, refill))
        if time.perf_counter() - start > budget_s:
            break
    results.sort()
    ms, block, bpc, refill = results[0]
    best = {"block": block, "blocks_per_cu": bpc, "refill": refill, "checkpoint": 64, "integrate_ms": ms,
            "trials": len(results), "precision": int(tracer.precision)}
    if tune_precision and tracer.precision == A.PRR_FP32_BULK:
        # the FP64 stages at the same shape, refilling as the FP32-bulk form does (whole waves)
        t64 = A.Tracer.from_buffer_copy(tracer)
        t64.precision = A.PRR_FP64
        ctx.set_tracer(t64)
        ctx.set_launch(block, bpc, 64, 64)
        ms64 = _time_integrate(ctx, frames)
        best["integrate_ms_fp32_stages"], best["integrate_ms_fp64_stages"] = ms, ms64
        best["trials"] += 1
        if ms64 < ms:
            best["precision"], best["integrate_ms"], best["refill"] = int(A.PRR_FP64), ms64, 64
        else:
            ctx.set_tracer(tracer)
    ctx.set_launch(block, bpc, best["refill"], 64)
    cache[key] = best
    _save(cache)
    return best


def tune_checkpoint(ctx, tmpl, orbit, dLdI, device_name="gpu0", use_cache=True):
    """Pick the adjoint's checkpoint interval (accepted steps between checkpoints)."""
    cache = _load()
    key = _key(device_name, A.PRR_KERR_BL, A.PRR_RKF45, -1) + "|adjoint"
    cur = ctx.get_launch()
    if use_cache and key in cache:
        k = cache[key]["checkpoint"]
    else:
        timings = []
        for k in CHECKPOINTS:
            ctx.set_launch(cur["block"], cur["blocks_per_cu"], cur["refill"], k)
            ctx.adjoint(tmpl, orbit, dLdI)
            timings.append((ctx.stats()["adjoint_ms"], k))
        timings.sort()
        k = timings[0][1]
        cache[key] = {"checkpoint": k, "adjoint_ms": timings[0][0], "trials": len(timings)}
        _save(cache)
    ctx.set_launch(cur["block"], cur["blocks_per_cu"], cur["refill"], k)
    return k


def tune_adjoint_precision(ctx, tmpl, orbit, dLdI, tracer, reps=3):
    """For an FP32-bulk tracer: time prr_adjoint (the whole call, HIP events) with the FP32 and with
    the FP64 stages and keep the faster on ctx; both meet the FP32-bulk form's accuracy contract
    (the adjoint's passes all follow the tracer's stage precision, so the replay still repeats
    the first pass bit for bit).  Returns {"precision", "ms_fp32_stages", "ms_fp64_stages"}."""
    if tracer.precision != A.PRR_FP32_BULK:
        return {"precision": int(tracer.precision)}
    t64 = A.Tracer.from_buffer_copy(tracer)
    t64.precision = A.PRR_FP64
    ms = {}
    for name, t in (("ms_fp32_stages", tracer), ("ms_fp64_stages", t64)):
        ctx.set_tracer(t)
        ctx.adjoint(tmpl, orbit, dLdI)  # warm-up (buffers, code objects)
        best = float("inf")
        for _ in range(reps):
            ctx.adjoint(tmpl, orbit, dLdI)
            best = min(best, ctx.stats()["total_ms"])
        ms[name] = best
    fp64 = ms["ms_fp64_stages"] < ms["ms_fp32_stages"]
    ctx.set_tracer(t64 if fp64 else tracer)
    return dict(precision=int(A.PRR_FP64 if fp64 else A.PRR_FP32_BULK), **ms)

