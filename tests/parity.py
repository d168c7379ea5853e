"""Shared helpers of the GPU-vs-oracle parity tests.

Parity bar (per-pixel / per-ray), against the CPU FP64 oracle on the same
seeded inputs:
  * termination class identical, except rays within 1e-6 rad of a class
    boundary: a ray whose class differs is excused only if one of the oracle's
    four +-1e-6 rad perturbations of its direction changes the oracle's class;
  * endpoint direction within the angular tolerance stated by each test;
  * RGB8 within 1 LSB on >= 99.9% of pixels.
"""
import numpy as np

import oracle as O
from parrelray_amd import _abi as A
from parrelray_amd import api
from parrelray_amd import configs as CF

BOUNDARY_DELTA = 1e-6


def oracle_scene(metric, tracer, sphere_filter=A.PRR_FILTER_BILINEAR, sky_filter=A.PRR_FILTER_SQAVG):
    S = O.Scene(metric, 1.0, 0.0, tracer)
    S.set_texture(0, CF.sky_texture(), sky_filter)
    S.set_texture(1, CF.sphere_texture(), sphere_filter)
    return S


def gpu_context(metric, tracer, sphere_filter=A.PRR_FILTER_BILINEAR, sky_filter=A.PRR_FILTER_SQAVG):
    ctx = api.Context(0)
    ctx.set_spacetime(metric, 1.0, 0.0)
    ctx.set_tracer(tracer)
    ctx.set_texture(0, CF.sky_texture(), sky_filter)
    ctx.set_texture(1, CF.sphere_texture(), sphere_filter)
    return ctx


def boundary_excused(S, frame, ray_ids, oracle_cls):
    """For rays whose classes disagree: True where a +-1e-6 rad perturbation flips the oracle's class."""
    ray_ids = np.asarray(ray_ids, np.int64)
    excused = np.zeros(ray_ids.size, bool)
    if ray_ids.size == 0:
        return excused
    for axis in (1, 2, 3, 4):
        r = S.render_rays(frame, ray_ids, delta=BOUNDARY_DELTA, axis=axis)
        excused |= r["cls"] != oracle_cls
    return excused


def angle(a, b):
    a = a / np.linalg.norm(a, axis=-1, keepdims=True)
    b = b / np.linalg.norm(b, axis=-1, keepdims=True)
    c = np.clip(np.sum(a * b, axis=-1), -1.0, 1.0)
    s = np.linalg.norm(np.cross(a, b), axis=-1)
    return np.arctan2(s, c)


def compare_rays(S, frame, gpu, orc, ids=None):
    """Class agreement (with boundary excuses) and endpoint angles of agreeing escape/sphere rays."""
    ids = np.arange(frame.nrays) if ids is None else np.asarray(ids)
    gcls, ocls = gpu["cls"][ids], orc["cls"]
    bad = np.nonzero(gcls != ocls)[0]
    exc = boundary_excused(S, frame, ids[bad], ocls[bad])
    same = gcls == ocls
    dir_mask = same & ((ocls == A.PRR_ESCAPE) | (ocls == A.PRR_SPHERE))
    ang = angle(gpu["endpoint"][ids][dir_mask], orc["endpoint"][dir_mask])
    return dict(n=ids.size, mismatched=int(bad.size), excused=int(exc.sum()), unexcused=int((~exc).sum()),
                unexcused_ids=ids[bad][~exc], angles=ang, same=same)


def rgb8_agreement(a, b):
    d = np.abs(a.astype(np.int16) - b.astype(np.int16)).max(axis=-1)
    return float((d <= 1).mean()), int(d.max())
