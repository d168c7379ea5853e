"""Precision report (needs a GPU; not collected by pytest).

For each Kerr workload, renders the frame on the GPU in the plain FP64 form
and in the FP32-bulk form, then integrates a seeded stratified subset of the
same rays with the CPU FP64 oracle at tol 1e-12 and reports, per form:
termination-class agreement (rays within 1e-6 rad of a class boundary
excused by the oracle's own perturbations), and the endpoint-angle error
(max, mean, percentiles) of rays whose class agrees.  Also reports the
full-frame comparison of each form against the GPU's own FP64 tol-1e-12
render.

    python tests/precision_report.py [--out precision.json] [--rays 4000] [--configs c1,c4,c4full,c2]

c4 is a 512x512 2x2-SSAA crop of the precision-stress frame; c4full is the
whole 2048x2048 4x4-SSAA frame (67,108,864 rays).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

from parity import angle, boundary_excused, gpu_context, oracle_scene  # noqa: E402
from parrelray_amd import _abi as A  # noqa: E402
from parrelray_amd import configs as CF  # noqa: E402


def stats(a):
    if a.size == 0:
        return {}
    return {"max": float(a.max()), "mean": float(a.mean()), "median": float(np.median(a)),
            "p99": float(np.quantile(a, 0.99)), "p999": float(np.quantile(a, 0.999)),
            "frac_le_1e-6": float((a <= 1e-6).mean())}


def compare(gpu, ref, ids=None):
    g_cls, g_end = (gpu["cls"], gpu["endpoint"]) if ids is None else (gpu["cls"][ids], gpu["endpoint"][ids])
    same = g_cls == ref["cls"]
    m = same & ((ref["cls"] == A.PRR_ESCAPE) | (ref["cls"] == A.PRR_SPHERE))
    return same, angle(g_end[m], ref["endpoint"][m])


def run(name, frame, tol, nrays, seed, log):
    rng = np.random.default_rng(seed)
    out = {"frame": f"{frame.width}x{frame.height}x{frame.samples}", "tol": tol}
    ids = np.sort(rng.choice(frame.nrays, size=min(nrays, frame.nrays), replace=False))
    S = oracle_scene(A.PRR_KERR_BL, CF.tracer_kerr(tol=1e-12, max_steps=400000))
    t0 = time.time()
    orc = S.render_rays(frame, ids)
    out["oracle_seconds"] = time.time() - t0
    log(f"{name}: oracle {len(ids)} rays in {out['oracle_seconds']:.1f} s")
    ref_gpu = gpu_context(A.PRR_KERR_BL, CF.tracer_kerr(tol=1e-12, max_steps=400000)).render(frame, rays=True)[0]
    for pname, prec in (("fp64", A.PRR_FP64), ("fp32_bulk", A.PRR_FP32_BULK)):
        ctx = gpu_context(A.PRR_KERR_BL, CF.tracer_kerr(precision=prec, tol=tol, max_steps=400000))
        g = ctx.render(frame, rays=True)[0]
        st = ctx.stats()
        same, ang = compare(g, orc, ids)
        bad = np.nonzero(~same)[0]
        exc = boundary_excused(S, frame, ids[bad], orc["cls"][bad])
        same_f, ang_f = compare(g, ref_gpu)
        out[pname] = {"vs_oracle_subset": {"rays": int(len(ids)), "class_mismatch": int(bad.size),
                                           "class_mismatch_unexcused": int((~exc).sum()),
                                           "endpoint_angle": stats(ang)},
                      "vs_gpu_fp64_tol1e-12_full": {"rays": int(frame.nrays),
                                                    "class_agreement": float(same_f.mean()),
                                                    "endpoint_angle": stats(ang_f)},
                      "integrate_ms": st["integrate_ms"], "attempts": st["attempts"],
                      "attempts_fp64": st["attempts_fp64"]}
        log(f"{name} {pname}: {json.dumps(out[pname])}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="precision.json")
    ap.add_argument("--rays", type=int, default=4000)
    ap.add_argument("--configs", default="c1,c4")
    args = ap.parse_args()

    def log(msg):
        print(msg, flush=True)

    report = {}
    cfgs = args.configs.split(",")
    if "c1" in cfgs:
        report["C1"] = run("C1", CF.frame_c1(), 1e-8, args.rays, 101, log)
    if "c4" in cfgs:
        report["C4"] = run("C4", CF.frame_c4(512, 512, ssaa=2), 1e-10, args.rays, 104, log)
    if "c4full" in cfgs:
        report["C4_full"] = run("C4 full", CF.frame_c4(), 1e-10, args.rays, 104, log)
    if "c2" in cfgs:
        report["C2_frame10"] = run("C2 frame 10", CF.frames_c2()[10], 1e-8, args.rays, 102, log)
    os.makedirs(os.path.dirname(args.out) or ".", exist_ok=True)
    with open(args.out, "w") as fh:
        json.dump(report, fh, indent=1)
    log(f"wrote {args.out}")


if __name__ == "__main__":
    main()
