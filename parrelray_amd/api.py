"""Python host API over the C ABI (include/prr.h).

Mirrors the reference's host objects: the `.params` store (Parameters,
main/Parameters.{h,cpp}), the spacetime / tracer / camera set-up of
createScene (prr.cpp:159-198) and one render call per batch of frames
(renderframe, prr.cpp:237-341).  All computation runs in libprr.so's HIP
kernels; numpy arrays carry host inputs and outputs, torch tensors (when
given) stay on the GPU.
"""
import ctypes as C
import os

import numpy as np

from . import _abi as A
from ._lib import check, lib


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


# ---------------------------------------------------------------- textures
def checker_texture(w=512, h=256, cw=16, ch=8, seed=1):
    out = np.empty((h, w, 3), np.uint8)
    check(lib().prr_texture_checker(_p(out), w, h, cw, ch, seed))
    return out


def noise_texture(w=1024, h=512, octaves=4, seed=2):
    out = np.empty((h, w, 3), np.uint8)
    check(lib().prr_texture_value_noise(_p(out), w, h, octaves, seed))
    return out


def read_image(path):
    """RGB8 array (H, W, 3) of a binary PPM, PNG or baseline JPEG file (prr_image_read)."""
    w, h = C.c_int(0), C.c_int(0)
    check(lib().prr_image_read(os.fsencode(path), None, C.byref(w), C.byref(h)))
    out = np.empty((h.value, w.value, 3), np.uint8)
    check(lib().prr_image_read(os.fsencode(path), _p(out), C.byref(w), C.byref(h)))
    return out


# ---------------------------------------------------------------- .params
class Params:
    """A parsed .params store (readParams, main/io.h:21-59)."""

    def __init__(self, path=None, sim=None):
        h = C.c_void_p()
        if sim is not None:
            check(lib().prr_params_load_sim(C.byref(h), sim.encode()))
        else:
            check(lib().prr_params_read(C.byref(h), path.encode()))
        self.h = h
        self._free = lib().prr_params_free  # kept: module globals may be gone at interpreter exit

    def __del__(self):
        if getattr(self, "h", None):
            self._free(self.h)
            self.h = None

    def read_into(self, path):
        check(lib().prr_params_read_into(self.h, path.encode()))

    def header_exists(self, header):
        return bool(lib().prr_params_header_exists(self.h, header.encode()))

    def par_exists(self, header, key):
        return bool(lib().prr_params_par_exists(self.h, header.encode(), key.encode()))

    def par(self, header, key):
        n = lib().prr_params_par(self.h, header.encode(), key.encode(), None, 0)
        buf = C.create_string_buffer(n + 1)
        lib().prr_params_par(self.h, header.encode(), key.encode(), buf, n + 1)
        return buf.value.decode()

    def pari(self, header, key):
        return lib().prr_params_pari(self.h, header.encode(), key.encode())

    def parf(self, header, key, default=0.0):
        return lib().prr_params_parf(self.h, header.encode(), key.encode(), default)

    def frame_count(self):
        return lib().prr_params_frame_count(self.h)

    def spacetime(self):
        m = C.c_int32()
        M, a, q = C.c_double(), C.c_double(), C.c_double()
        pv = (C.c_double * 16)()
        check(lib().prr_params_spacetime(self.h, C.byref(m), C.byref(M), C.byref(a), C.byref(q), pv))
        return dict(metric=m.value, M=M.value, a=a.value, q=q.value, p=list(pv))

    def tracer(self):
        t = A.Tracer()
        check(lib().prr_params_tracer(self.h, C.byref(t)))
        return t

    def frame(self, i):
        f = A.Frame()
        check(lib().prr_params_frame(self.h, i, C.byref(f)))
        return f


# ---------------------------------------------------------------- context
class Context:
    """One prr_ctx: device buffers, spacetime, tracer, textures and launch shape."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().prr_create(C.byref(h), device))
        self.h = h
        self._destroy = lib().prr_destroy  # kept: module globals may be gone at interpreter exit
        self.metric = A.PRR_SCHW_KS_CART

    def close(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, st):
        check(st, self.h)

    def set_spacetime(self, metric, M=1.0, a=0.0, q=0.0, p=None):
        pv = (C.c_double * 16)(*(p or [0.0] * 16))
        self._check(lib().prr_set_spacetime(self.h, metric, M, a, q, pv))
        self.metric = metric

    def set_tracer(self, tracer):
        self._check(lib().prr_set_tracer(self.h, C.byref(tracer)))

    def set_texture(self, slot, rgb8, filt=A.PRR_FILTER_SQAVG):
        if rgb8 is None:
            self._check(lib().prr_set_texture(self.h, slot, None, 0, 0, filt))
            return
        rgb8 = np.ascontiguousarray(rgb8, dtype=np.uint8)
        h, w, _ = rgb8.shape
        self._check(lib().prr_set_texture(self.h, slot, _p(rgb8), w, h, filt))

    def set_spheres(self, spheres):
        arr = (A.Sphere * max(1, len(spheres)))()
        for i, (c, r, tex) in enumerate(spheres):
            arr[i].c[:] = list(c)
            arr[i].radius = r
            arr[i].tex = tex
        self._check(lib().prr_set_spheres(self.h, arr, len(spheres)))

    def set_launch(self, block=256, blocks_per_cu=0, refill=64, checkpoint=64, adjoint_record_bytes=0):
        l = A.Launch(block, blocks_per_cu, refill, checkpoint, adjoint_record_bytes)
        self._check(lib().prr_set_launch(self.h, C.byref(l)))

    def get_launch(self):
        l = A.Launch()
        self._check(lib().prr_get_launch(self.h, C.byref(l)))
        return dict(block=l.block, blocks_per_cu=l.blocks_per_cu, refill=l.refill, checkpoint=l.checkpoint,
                    adjoint_record_bytes=l.adjoint_record_bytes)

    def stats(self):
        s = A.Stats()
        self._check(lib().prr_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    def render(self, frames, rays=False, rgbf=True):
        """Render frames into host arrays.

        Returns a list of dicts per frame with 'rgb8' (HxWx3 uint8), 'rgbf'
        (HxWx3 float32) and, with rays=True, per-ray 'cls', 'endpoint',
        'steps', 'attempts' and 'color'.
        """
        if isinstance(frames, A.Frame):
            frames = [frames]
        n = len(frames)
        farr = (A.Frame * n)(*frames)
        outs, imgs = [], (A.Image * n)()
        for i, f in enumerate(frames):
            o = dict(rgb8=np.empty((f.height, f.width, 3), np.uint8))
            if rgbf:
                o["rgbf"] = np.empty((f.height, f.width, 3), np.float32)
            imgs[i].rgb8 = o["rgb8"].ctypes.data
            imgs[i].rgbf = o["rgbf"].ctypes.data if rgbf else None
            imgs[i].on_device = 0
            outs.append(o)
        ro = None
        if rays:
            total = sum(f.nrays for f in frames)
            rr = dict(cls=np.empty(total, np.int8), endpoint=np.empty((total, 3)), steps=np.empty(total, np.int32),
                      attempts=np.empty(total, np.int32), color=np.empty((total, 3), np.float32))
            ro = A.RayOut(rr["cls"].ctypes.data, rr["endpoint"].ctypes.data, rr["steps"].ctypes.data,
                          rr["attempts"].ctypes.data, rr["color"].ctypes.data, 0, 0)
        self._check(lib().prr_render(self.h, farr, n, imgs, C.byref(ro) if ro is not None else None, None))
        if rays:
            off = 0
            for o, f in zip(outs, frames):
                k = f.nrays
                for key, arr in rr.items():
                    o[key] = arr[off:off + k]
                off += k
        return outs

    def render_device(self, frames, rgb8=None, rgbf=None, rays=None, stream=None):
        """Render into device buffers (lists of torch tensors, one per frame, or None).

        rays: optional dict of device tensors over all rays of the call (frame
        order) that receive per-ray outputs: 'cls' (int8), 'endpoint'
        (float64, n x 3), 'steps', 'attempts' (int32), 'color' (float32, n x 3).
        """
        if isinstance(frames, A.Frame):
            frames = [frames]
        n = len(frames)
        farr = (A.Frame * n)(*frames)
        imgs = (A.Image * n)()
        for i in range(n):
            imgs[i].rgb8 = rgb8[i].data_ptr() if rgb8 is not None else None
            imgs[i].rgbf = rgbf[i].data_ptr() if rgbf is not None else None
            imgs[i].on_device = 1
        ro = None
        if rays is not None:
            ptr = lambda k: rays[k].data_ptr() if rays.get(k) is not None else None  # noqa: E731
            ro = A.RayOut(ptr("cls"), ptr("endpoint"), ptr("steps"), ptr("attempts"), ptr("color"), 1, 0)
        self._check(lib().prr_render(self.h, farr, n, imgs, C.byref(ro) if ro is not None else None,
                                     C.c_void_p(stream) if stream else None))

    def render_noalloc(self, frames, rgb8=None, rgbf=None, stream=None):
        """prr_render_noalloc: render into device buffers (lists of torch tensors) without allocating."""
        if isinstance(frames, A.Frame):
            frames = [frames]
        n = len(frames)
        farr = (A.Frame * n)(*frames)
        imgs = (A.Image * n)()
        for i in range(n):
            imgs[i].rgb8 = rgb8[i].data_ptr() if rgb8 is not None else None
            imgs[i].rgbf = rgbf[i].data_ptr() if rgbf is not None else None
            imgs[i].on_device = 1
        self._check(lib().prr_render_noalloc(self.h, farr, n, imgs, C.c_void_p(stream) if stream else None))

    def adjoint(self, tmpl, orbit, dL_dimg, img_out=None):
        """Gradient of L w.r.t. (spin, r_cam, incl, sphere_phi) given dL/dI of the frame's image.

        dL_dimg: HxWx3 float32 numpy array or CUDA torch tensor.  img_out (same
        kind, optional) receives the forward image.  Returns a numpy array of 4.
        """
        grad = (C.c_double * 4)()
        on_dev = not isinstance(dL_dimg, np.ndarray)
        if on_dev:
            ptr = C.c_void_p(dL_dimg.data_ptr())
            out = C.c_void_p(img_out.data_ptr()) if img_out is not None else None
        else:
            dL_dimg = np.ascontiguousarray(dL_dimg, np.float32)
            ptr = C.c_void_p(dL_dimg.ctypes.data)
            out = C.c_void_p(img_out.ctypes.data) if img_out is not None else None
        self._check(lib().prr_adjoint(self.h, C.byref(tmpl), C.byref(orbit), ptr, int(on_dev), grad, out))
        return np.array(list(grad))

    def trace(self, x0u0):
        """Integrate explicit initial states (n x 8): returns xend, cls, steps, attempts."""
        x = np.ascontiguousarray(x0u0, dtype=np.float64).reshape(-1, 8)
        n = x.shape[0]
        xend = np.empty((n, 8))
        cls = np.empty(n, np.int32)
        steps = np.empty(n, np.int32)
        att = np.empty(n, np.int32)
        self._check(lib().prr_trace(self.h, _p(x), n, _p(xend), _p(cls), _p(steps), _p(att)))
        return dict(xend=xend, cls=cls, steps=steps, attempts=att)


class Graph:
    """A captured multi-frame render (prr_graph_create / prr_graph_launch).

    The graph writes into the given device tensors on every replay; it keeps
    references to them (and to its context) for as long as it lives."""

    def __init__(self, ctx, frames, rgb8=None, rgbf=None):
        n = len(frames)
        self.ctx = ctx
        self._images = (list(rgb8) if rgb8 is not None else [], list(rgbf) if rgbf is not None else [])
        self._frames = (A.Frame * n)(*frames)
        imgs = (A.Image * n)()
        for i in range(n):
            imgs[i].rgb8 = rgb8[i].data_ptr() if rgb8 is not None else None
            imgs[i].rgbf = rgbf[i].data_ptr() if rgbf is not None else None
            imgs[i].on_device = 1
        h = C.c_void_p()
        ctx._check(lib().prr_graph_create(ctx.h, self._frames, n, imgs, C.byref(h)))
        self.h = h
        self._destroy = lib().prr_graph_destroy  # kept: module globals may be gone at interpreter exit

    def launch(self, stream=None):
        self.ctx._check(lib().prr_graph_launch(self.h, C.c_void_p(stream) if stream else None))

    def close(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def frame_from_orbit(tmpl, orbit):
    out = A.Frame()
    check(lib().prr_frame_from_orbit(C.byref(tmpl), C.byref(orbit), C.byref(out)))
    return out


# ---------------------------------------------------------------- host evaluation of the device formulas
def state_size(metric):
    return lib().prr_debug_state_size(metric)


def _f64(a, n=None):
    a = np.ascontiguousarray(a, np.float64)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def rhs(metric, y, M=1.0, a=0.0, E=0.0, L=0.0, Q=0.0, fp32=False):
    """Right-hand side f(y) of the integrator state (host evaluation of the device formula)."""
    y = _f64(y, state_size(metric))
    f = np.empty_like(y)
    check(lib().prr_debug_rhs(metric, M, a, E, L, Q, _p(y), int(fp32), _p(f)))
    return f


def kerr_delta_rhs(y0, d, a, E, L, Q, M=1.0):
    """FP32-bulk stage function g = f(y0 + d) - f(y0) (Kerr)."""
    y0, d = _f64(y0, 6), _f64(d, 6)
    g = np.empty(6)
    check(lib().prr_debug_kerr_delta_rhs(M, a, E, L, Q, _p(y0), _p(d), _p(g)))
    return g


def kerr_stages(y, h, a, E, L, Q, fp32=False, M=1.0):
    """One Kerr RKF45 attempt: (d4, e), FP64 or FP32-bulk stages."""
    y = _f64(y, 6)
    d4, e = np.empty(6), np.empty(6)
    check(lib().prr_debug_kerr_stages(M, a, E, L, Q, _p(y), h, int(fp32), _p(d4), _p(e)))
    return d4, e


def rhs_vjp(y, fb, a, E, L, Q, M=1.0):
    """Kerr: (fb . df/dy [6], fb . df/d(E, L, Q, a) [4])."""
    y, fb = _f64(y, 6), _f64(fb, 6)
    yb, cb = np.zeros(6), np.zeros(4)
    check(lib().prr_debug_rhs_vjp(M, a, E, L, Q, _p(y), _p(fb), _p(yb), _p(cb)))
    return yb, cb


def orbit_init(a, r_cam, incl, cam_phi, vfov, aspect, u, v, M=1.0):
    """z = (u, du/ds, mu, dmu/ds, phi, E, L, Q) of an orbiting camera's ray and dz/d(a, r_cam, incl)."""
    z, dz = np.empty(8), np.empty(24)
    check(lib().prr_debug_orbit_init(M, a, r_cam, incl, cam_phi, vfov, aspect, u, v, _p(z), _p(dz)))
    return z, dz.reshape(8, 3)


def kerr_step(y, h, a, E, L, Q, M=1.0):
    y = _f64(y, 6)
    out = np.empty(6)
    check(lib().prr_debug_kerr_step(M, a, E, L, Q, _p(y), h, _p(out)))
    return out


def kerr_step_vjp(y, h, lam, a, E, L, Q, M=1.0):
    """lam[9] = adjoint of (y'[0..4], E, L, Q, a) -> adjoint of (y[0..4], E, L, Q, a)."""
    y = _f64(y, 6)
    lam = _f64(lam, 9).copy()
    check(lib().prr_debug_kerr_step_vjp(M, a, E, L, Q, _p(y), h, _p(lam)))
    return lam


def escape_color(y, a, E, L, Q, rgb8, M=1.0):
    """Bilinear sky colour of an escaping Kerr state y[0..4] and d(colour)/d(y[0..4], E, L, Q, a)."""
    y = _f64(y[:5], 5)
    rgb8 = np.ascontiguousarray(rgb8, np.uint8)
    c, dc = np.empty(3), np.empty(27)
    check(lib().prr_debug_escape_color(M, a, E, L, Q, _p(y), _p(rgb8), rgb8.shape[1], rgb8.shape[0], _p(c), _p(dc)))
    return c, dc.reshape(3, 9)


def state_from_xu(metric, x, u, M=1.0, a=0.0):
    """(x, u) -> (integrator state, (E, L, Q))."""
    x, u = _f64(x, 4), _f64(u, 4)
    y = np.empty(state_size(metric))
    ELQ = np.empty(3)
    check(lib().prr_debug_state_from_xu(metric, M, a, _p(x), _p(u), _p(y), _p(ELQ)))
    return y, ELQ


def xu_from_state(metric, y, M=1.0, a=0.0, E=0.0, L=0.0, Q=0.0):
    y = _f64(y, state_size(metric))
    x, u = np.empty(4), np.empty(4)
    check(lib().prr_debug_xu_from_state(metric, M, a, E, L, Q, _p(y), _p(x), _p(u)))
    return x, u


def init_ray(metric, pos, d, M=1.0, a=0.0):
    """initRays for a camera position and direction: (state, (E, L, Q))."""
    pos, d = _f64(pos, 3), _f64(d, 3)
    y = np.empty(state_size(metric))
    ELQ = np.empty(3)
    check(lib().prr_debug_init_ray(metric, M, a, _p(pos), _p(d), _p(y), _p(ELQ)))
    return y, ELQ


def trace_rays_host(metric, tracer, frame, ids=None, x0u0=None):
    """The device integrator's per-ray code run on the host (untextured)."""
    if x0u0 is not None:
        x0u0 = _f64(x0u0)
        n = x0u0.size // 8
        idp = None
    else:
        ids = np.ascontiguousarray(ids, np.int64)
        n = ids.size
        idp = _p(ids)
    cls = np.empty(n, np.int32)
    steps = np.empty(n, np.int32)
    att = np.empty(n, np.int32)
    ep = np.empty((n, 3))
    st8 = np.empty((n, 8))
    check(lib().prr_debug_trace_rays(metric, C.byref(tracer), C.byref(frame),
                                     _p(x0u0) if x0u0 is not None else None, idp, n, _p(cls), _p(ep), _p(steps),
                                     _p(att), _p(st8)))
    return dict(cls=cls, endpoint=ep, steps=steps, attempts=att, xend=st8)


def trace_log(metric, tracer, frame, ray, max_n=100000):
    """Per accepted step of one ray on the host: (steps, attempts, next h, FP64 attempts, state[8]) rows."""
    log = np.empty((max_n, 12))
    n = C.c_int(0)
    check(lib().prr_debug_trace_log(metric, C.byref(tracer), C.byref(frame), int(ray), max_n, _p(log), C.byref(n)))
    return log[:n.value]


def queue_rays(frame, q):
    """Ray ids of the integrator's queue positions q within one frame."""
    q = np.ascontiguousarray(q, dtype=np.int64)
    ids = np.empty_like(q)
    check(lib().prr_debug_queue_rays(C.byref(frame), _p(q), q.size, _p(ids)))
    return ids


def camera_ray(frame, local):
    d = np.empty(3)
    check(lib().prr_debug_camera_ray(C.byref(frame), local, _p(d)))
    return d
