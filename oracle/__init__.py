"""TEST INFRASTRUCTURE, NOT PRODUCT CODE.

ctypes binding of the CPU FP64 oracle (oracle/liboracle.so, built from
oracle/prr_oracle.hpp by oracle/Makefile).  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module; it is the checker, never
the thing measured or shipped.
"""
import ctypes as C
import os

import numpy as np

from parrelray_amd import _abi as A

_HERE = os.path.dirname(os.path.abspath(__file__))
# PRR_ORACLE_LIB: another build of the same source (the sanitizer build of tests/run_sanitized.sh)
_LIB = os.environ.get("PRR_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError(f"oracle library missing: {_LIB} (run `make -C oracle`)")
        L = C.CDLL(_LIB)
        L.orc_scene_new.restype = C.c_void_p
        L.orc_scene_new.argtypes = [C.c_int, C.c_double, C.c_double, C.POINTER(A.Tracer)]
        L.orc_scene_free.argtypes = [C.c_void_p]
        L.orc_scene_set_far_escape.argtypes = [C.c_void_p, C.c_double]
        L.orc_scene_set_fp32.argtypes = [C.c_void_p, C.c_int]
        L.orc_scene_set_tracer.argtypes = [C.c_void_p, C.POINTER(A.Tracer)]
        L.orc_scene_set_texture.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.orc_scene_set_spheres.argtypes = [C.c_void_p, C.POINTER(A.Sphere), C.c_int]
        L.orc_render_rays.argtypes = [C.c_void_p, C.POINTER(A.Frame), C.c_void_p, C.c_int64, C.c_double, C.c_int,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_int, _dp, C.c_void_p]
        L.orc_trace.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_int]
        L.orc_render_frame.argtypes = [C.c_void_p, C.POINTER(A.Frame), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.orc_ssaa.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.orc_fxaa.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.orc_rgb8.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
        L.orc_metric.argtypes = [C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_void_p]
        L.orc_dmetric_kerr.argtypes = [C.c_double, C.c_double, C.c_void_p, C.c_void_p]
        L.orc_christoffel_kerr.argtypes = [C.c_double, C.c_double, C.c_void_p, C.c_void_p]
        L.orc_geod.argtypes = [C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_kerr_constants.argtypes = [C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_camera_dir.argtypes = [C.POINTER(A.Frame), C.c_double, C.c_double, C.c_void_p]
        L.orc_camera_basis.argtypes = [C.c_double, C.c_double, C.c_void_p]
        L.orc_sample_uv.argtypes = [C.POINTER(A.Frame), C.c_int64, C.c_void_p]
        L.orc_init_ray.argtypes = [C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_sky_lookup.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_void_p]
        L.orc_bilinear.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_void_p]
        L.orc_chord_hit.argtypes = [C.POINTER(A.Sphere), C.c_int, C.c_void_p, C.c_void_p, _dp]
        L.orc_chord_hit.restype = C.c_int
        L.orc_tableau.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_tableau.restype = C.c_double
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None:
        assert a.size == n
    return a


class Scene:
    """Oracle scene: metric, spacetime, tracer, textures, spheres."""

    def __init__(self, metric, M=1.0, a=0.0, tracer=None):
        self.metric, self.M, self.a = metric, M, a
        self.tracer = tracer if tracer is not None else A.make_tracer()
        self.h = lib().orc_scene_new(metric, M, a, C.byref(self.tracer))
        self._keep = []

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_scene_free(self.h)
            self.h = None

    def set_far_escape(self, r):
        """Escape radius beyond a wormhole's throat (rout2); 0: the tracer's r_escape."""
        lib().orc_scene_set_far_escape(self.h, r)

    def set_fp32(self, on=True):
        """Integrate in FP32 (state, stages, error norm, controller), the
        reference's own precision; the event tests take an FP64 copy of the
        state.  For the CPU baseline's reference-like timing only."""
        lib().orc_scene_set_fp32(self.h, int(bool(on)))

    def set_tracer(self, tracer):
        self.tracer = tracer
        lib().orc_scene_set_tracer(self.h, C.byref(tracer))

    def set_texture(self, slot, rgb8, filt=A.PRR_FILTER_SQAVG):
        rgb8 = np.ascontiguousarray(rgb8, dtype=np.uint8)
        h, w, _ = rgb8.shape
        lib().orc_scene_set_texture(self.h, slot, _ptr(rgb8), w, h, filt)

    def set_spheres(self, spheres):
        arr = (A.Sphere * max(1, len(spheres)))()
        for i, (c, r, tex) in enumerate(spheres):
            arr[i].c[:] = list(c)
            arr[i].radius = r
            arr[i].tex = tex
        lib().orc_scene_set_spheres(self.h, arr, len(spheres))

    def render_rays(self, frame, ray_ids=None, delta=0.0, axis=0, nthreads=0, want_state=False, want_color64=False):
        """Per-ray results: dict of cls, endpoint, steps, attempts, color (+ state8, + color64: the
        colours through the FP64 colour path), seconds."""
        if ray_ids is None:
            n = frame.nrays
            ids = None
        else:
            ids = np.ascontiguousarray(ray_ids, dtype=np.int64)
            n = ids.size
        out = dict(cls=np.empty(n, np.int8), endpoint=np.empty((n, 3)), steps=np.empty(n, np.int32),
                   attempts=np.empty(n, np.int32), color=np.empty((n, 3), np.float32))
        if want_state:
            out["state8"] = np.empty((n, 8))
        if want_color64:
            out["color64"] = np.empty((n, 3))
        sec = C.c_double(0.0)
        lib().orc_render_rays(self.h, C.byref(frame), _ptr(ids), n, delta, axis, _ptr(out["cls"]),
                              _ptr(out["endpoint"]), _ptr(out["steps"]), _ptr(out["attempts"]),
                              _ptr(out["color"]), _ptr(out.get("state8")), nthreads, C.byref(sec),
                              _ptr(out.get("color64")))
        out["seconds"] = sec.value
        return out

    def render_frame(self, frame, nthreads=0):
        """(ray colours, final linear image HxWx3, rgb8 HxWx3)."""
        W, H = frame.width, frame.height
        col = np.empty((frame.nrays, 3), np.float32)
        img = np.empty((H, W, 3), np.float32)
        rgb8 = np.empty((H, W, 3), np.uint8)
        lib().orc_render_frame(self.h, C.byref(frame), _ptr(col), _ptr(img), _ptr(rgb8), nthreads)
        return col, img, rgb8

    def trace(self, x0u0, nthreads=0):
        x0u0 = _f64(x0u0).reshape(-1, 8)
        n = x0u0.shape[0]
        xend = np.empty((n, 8))
        cls = np.empty(n, np.int32)
        steps = np.empty(n, np.int32)
        att = np.empty(n, np.int32)
        lib().orc_trace(self.h, _ptr(x0u0), n, _ptr(xend), _ptr(cls), _ptr(steps), _ptr(att), nthreads)
        return dict(xend=xend, cls=cls, steps=steps, attempts=att)

    def sky_lookup(self, theta, phi):
        out = np.empty(3, np.float32)
        lib().orc_sky_lookup(self.h, theta, phi, _ptr(out))
        return out

    def bilinear(self, slot, u, v):
        out = np.empty(3, np.float32)
        lib().orc_bilinear(self.h, slot, u, v, _ptr(out))
        return out


def metric(kind, M, a, x):
    x = _f64(x, 4)
    g = np.empty(16)
    lib().orc_metric(kind, M, a, _ptr(x), _ptr(g))
    return g.reshape(4, 4)


def dmetric_kerr(M, a, x):
    x = _f64(x, 4)
    dg = np.empty(64)
    lib().orc_dmetric_kerr(M, a, _ptr(x), _ptr(dg))
    return dg.reshape(4, 4, 4)


def christoffel_kerr(M, a, x):
    x = _f64(x, 4)
    G = np.empty(64)
    lib().orc_christoffel_kerr(M, a, _ptr(x), _ptr(G))
    return G.reshape(4, 4, 4)


def geod(kind, M, a, x, u):
    x, u = _f64(x, 4), _f64(u, 4)
    du = np.empty(4)
    lib().orc_geod(kind, M, a, _ptr(x), _ptr(u), _ptr(du))
    return du


def kerr_constants(M, a, x, u):
    x, u = _f64(x, 4), _f64(u, 4)
    out = np.empty(3)
    lib().orc_kerr_constants(M, a, _ptr(x), _ptr(u), _ptr(out))
    return out


def camera_dir(frame, u, v):
    d = np.empty(3)
    lib().orc_camera_dir(C.byref(frame), u, v, _ptr(d))
    return d


def camera_basis(alpha, beta):
    out = np.empty(9)
    lib().orc_camera_basis(alpha, beta, _ptr(out))
    return out.reshape(3, 3)  # rows: forward, up, right


def sample_uv(frame, ray):
    uv = np.empty(2)
    lib().orc_sample_uv(C.byref(frame), ray, _ptr(uv))
    return uv


def init_ray(kind, M, a, pos, d):
    pos, d = _f64(pos, 3), _f64(d, 3)
    x, u = np.empty(4), np.empty(4)
    lib().orc_init_ray(kind, M, a, _ptr(pos), _ptr(d), _ptr(x), _ptr(u))
    return x, u


def ssaa(ray_rgb, W, H, s):
    ray_rgb = np.ascontiguousarray(ray_rgb, dtype=np.float32)
    img = np.empty((H, W, 3), np.float32)
    lib().orc_ssaa(_ptr(ray_rgb), W, H, s, _ptr(img))
    return img


def fxaa(img):
    img = np.ascontiguousarray(img, dtype=np.float32)
    H, W, _ = img.shape
    out = np.empty_like(img)
    lib().orc_fxaa(_ptr(img), W, H, _ptr(out))
    return out


def rgb8(img):
    img = np.ascontiguousarray(img, dtype=np.float32)
    out = np.empty(img.shape, np.uint8)
    lib().orc_rgb8(_ptr(img), img.size, _ptr(out))
    return out


def chord_hit(spheres, P0, P1):
    arr = (A.Sphere * max(1, len(spheres)))()
    for i, (c, r) in enumerate(spheres):
        arr[i].c[:] = list(c)
        arr[i].radius = r
    t = C.c_double(0.0)
    P0, P1 = _f64(P0, 3), _f64(P1, 3)
    k = lib().orc_chord_hit(arr, len(spheres), _ptr(P0), _ptr(P1), C.byref(t))
    return k, t.value


def tableau(solver):
    """(A, b propagated, bh embedded) of a Runge-Kutta solver of the oracle."""
    L = lib()
    S = int(L.orc_tableau(solver, 3, 0, 0))
    Amat = np.array([[L.orc_tableau(solver, 0, i, j) if j < i else 0.0 for j in range(S)] for i in range(S)])
    b = np.array([L.orc_tableau(solver, 1, i, 0) for i in range(S)])
    bh = np.array([L.orc_tableau(solver, 2, i, 0) for i in range(S)])
    return Amat, b, bh


def fehlberg():
    """Fehlberg's tableau as (A, b4, b5)."""
    Amat, b4, b5 = tableau(A.PRR_RKF45)
    return Amat[:, :5], b4, b5
