"""prr_render_noalloc: the allocation-free render entry point, called through
the C ABI, against prr_render on the same frames (same kernels: bit-identical
images).  Two calls in a row, the second with a smaller input, so that the
second call runs on the scratch array the first one left behind; inputs
beyond the array's size are refused."""
import ctypes as C

import numpy as np
import pytest

from parrelray_amd import _abi as A
from parrelray_amd import api
from parrelray_amd import configs as CF
from parrelray_amd._lib import PrrError, lib

pytestmark = pytest.mark.gpu


def context(metric, tracer, spin=0.0):
    ctx = api.Context(0)
    ctx.set_spacetime(metric, 1.0, spin)
    ctx.set_tracer(tracer)
    ctx.set_texture(0, CF.sky_texture())
    ctx.set_texture(1, CF.sphere_texture(), A.PRR_FILTER_BILINEAR)
    return ctx


def device_images(torch, frames):
    rgb8 = [torch.full((f.height, f.width, 3), 7, dtype=torch.uint8, device="cuda") for f in frames]
    rgbf = [torch.full((f.height, f.width, 3), -1.0, dtype=torch.float32, device="cuda") for f in frames]
    return rgb8, rgbf


def render_both(torch, ctx, frames):
    """(prr_render images, prr_render_noalloc images, stats of each) as numpy arrays."""
    r8, rf = device_images(torch, frames)
    ctx.render_device(frames, rgb8=r8, rgbf=rf)
    st_ref = ctx.stats()
    n8, nf = device_images(torch, frames)
    farr = (A.Frame * len(frames))(*frames)
    imgs = (A.Image * len(frames))()
    for i in range(len(frames)):
        imgs[i].rgb8, imgs[i].rgbf, imgs[i].on_device = n8[i].data_ptr(), nf[i].data_ptr(), 1
    print("prr_render_noalloc:", [(f.width, f.height, f.samples) for f in frames], flush=True)
    ctx._check(lib().prr_render_noalloc(ctx.h, farr, len(frames), imgs, None))
    st = ctx.stats()
    torch.cuda.synchronize()
    to_np = lambda ts: [t.cpu().numpy() for t in ts]  # noqa: E731
    return (to_np(r8), to_np(rf)), (to_np(n8), to_np(nf)), st_ref, st


def test_noalloc_matches_render_twice_second_smaller():
    torch = pytest.importorskip("torch")
    # C0 (Schwarzschild coordinates, Euler, 256x256), then a smaller Kerr RKF45
    # batch of two frames with 2x2 SSAA, FXAA and the textured sphere
    ctx = context(A.PRR_SCHW_SPH, CF.tracer_c0())
    first = [CF.frame_c0()]
    ref, got, st_ref, st = render_both(torch, ctx, first)
    assert st["rays"] == st_ref["rays"] == first[0].nrays
    for a, b in zip(ref[0] + ref[1], got[0] + got[1]):
        assert np.array_equal(a, b)
    ctx2 = context(A.PRR_KERR_BL, CF.tracer_kerr(precision=A.PRR_FP32_BULK), spin=0.95)
    second = [CF.frame_c1(96, 54), CF.frame_c1(40, 30)]
    assert sum(f.nrays for f in second) < first[0].nrays
    ref, got, st_ref, st = render_both(torch, ctx2, second)
    assert st["rays"] == st_ref["rays"] == sum(f.nrays for f in second)
    assert st["attempts"] == st_ref["attempts"]
    for a, b in zip(ref[0] + ref[1], got[0] + got[1]):
        assert np.array_equal(a, b)
    assert got[1][0].min() >= 0.0  # every pixel written


def test_noalloc_refuses_inputs_beyond_its_scratch():
    torch = pytest.importorskip("torch")
    ctx = context(A.PRR_KERR_BL, CF.tracer_kerr(), spin=0.95)
    big = CF.frame_c1()  # 8.3 M rays: about 780 MB of workspace
    img = torch.empty((big.height, big.width, 3), dtype=torch.uint8, device="cuda")
    im = (A.Image * 1)()
    im[0].rgb8, im[0].on_device = img.data_ptr(), 1
    st = lib().prr_render_noalloc(ctx.h, (A.Frame * 1)(big), 1, im, None)
    assert st == A.PRR_EBUDGET
    assert b"scratch" in lib().prr_last_error(ctx.h)
    host = (A.Image * 1)()
    host[0].rgb8 = C.cast(C.create_string_buffer(3 * 16), C.c_void_p).value
    host[0].on_device = 0
    with pytest.raises(PrrError, match="device images"):
        ctx._check(lib().prr_render_noalloc(ctx.h, (A.Frame * 1)(CF.frame_c1(4, 4)), 1, host, None))
    # the context still renders afterwards
    small = CF.frame_c1(32, 18)
    out = torch.empty((18, 32, 3), dtype=torch.uint8, device="cuda")
    ctx.render_noalloc([small], rgb8=[out])
    assert ctx.stats()["rays"] == small.nrays
