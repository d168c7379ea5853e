"""Batched multi-frame rendering in one persistent launch, captured in a
hipGraph, and the launch-shape autotuner."""
import numpy as np
import pytest

import oracle as O
from parrelray_amd import _abi as A
from parrelray_amd import api
from parrelray_amd import configs as CF

pytestmark = pytest.mark.gpu


def context():
    ctx = api.Context(0)
    ctx.set_spacetime(A.PRR_KERR_BL, 1.0, 0.0)
    ctx.set_tracer(CF.tracer_kerr(precision=A.PRR_FP32_BULK))
    ctx.set_texture(0, CF.sky_texture())
    ctx.set_texture(1, CF.sphere_texture(), A.PRR_FILTER_BILINEAR)
    return ctx


def test_graph_replays_match_eager_render():
    torch = pytest.importorskip("torch")
    frames = CF.frames_c2(nframes=12, scale=0.06)
    ctx = context()
    eager = ctx.render(frames)
    rgb8 = [torch.zeros((f.height, f.width, 3), dtype=torch.uint8, device="cuda") for f in frames]
    rgbf = [torch.zeros((f.height, f.width, 3), dtype=torch.float32, device="cuda") for f in frames]
    g = api.Graph(ctx, frames, rgb8=rgb8, rgbf=rgbf)
    for rep in range(3):
        for t in rgbf:  # eager work on the same buffers between replays
            t.fill_(-1.0)
        torch.cuda.synchronize()
        g.launch()
        st = ctx.stats()
        assert st["rays"] == sum(f.nrays for f in frames)
        for e, b8, bf in zip(eager, rgb8, rgbf):
            assert np.array_equal(b8.cpu().numpy(), e["rgb8"]), rep
            assert np.array_equal(bf.cpu().numpy(), e["rgbf"]), rep
    g.close()


def test_graph_replays_survive_later_calls_on_its_context():
    """A graph owns its workspace and keeps the textures it captured: between
    replays, a larger eager batch (which grows the context's buffers), a
    trace, a larger launch shape and a new texture in a captured slot leave
    its replays equal to the render eagerly made before the capture."""
    torch = pytest.importorskip("torch")
    frames = CF.frames_c2(nframes=6, scale=0.05)
    ctx = context()
    eager = ctx.render(frames)
    rgb8 = [torch.zeros((f.height, f.width, 3), dtype=torch.uint8, device="cuda") for f in frames]
    g = api.Graph(ctx, frames, rgb8=rgb8)
    del rgb8  # the graph keeps its output tensors alive
    bigger = CF.frames_c2(nframes=8, scale=0.12)
    x, u = O.init_ray(A.PRR_KERR_BL, 1.0, 0.0, [20.0, 3.0, 1.0], [-0.9, 0.1, 0.0])
    x0 = np.tile(np.concatenate([x, u]), (100, 1))
    other_sky = api.checker_texture(512, 256, 8, 4, 99)
    steps = [lambda: ctx.render(bigger),
             lambda: ctx.trace(x0),
             lambda: ctx.set_launch(block=256, blocks_per_cu=4, refill=64),
             lambda: ctx.set_texture(0, other_sky)]
    for rep, step in enumerate([None] + steps):
        if step is not None:
            step()
        g.launch()
        st = ctx.stats()
        assert st["rays"] == sum(f.nrays for f in frames), rep
        for e, b8 in zip(eager, g._images[0]):
            assert np.array_equal(b8.cpu().numpy(), e["rgb8"]), rep
    # the context itself now renders with the new sky
    now = ctx.render(frames[:1])[0]["rgb8"]
    assert not np.array_equal(now, eager[0]["rgb8"])
    g.close()


def test_graph_outlives_a_destroyed_context():
    """prr_destroy with a live graph defers the release to the graph's
    destruction; the graph then refuses to launch."""
    torch = pytest.importorskip("torch")
    from parrelray_amd._lib import PrrError, lib
    frames = CF.frames_c2(nframes=2, scale=0.04)
    ctx = context()
    rgb8 = [torch.zeros((f.height, f.width, 3), dtype=torch.uint8, device="cuda") for f in frames]
    g = api.Graph(ctx, frames, rgb8=rgb8)
    g.launch()
    lib().prr_destroy(ctx.h)
    ctx.h = None
    with pytest.raises(PrrError, match="destroyed"):
        g.launch()
    g.close()


def test_autotuner_picks_a_shape_and_results_do_not_change(tmp_path, monkeypatch):
    from parrelray_amd import autotune
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "tune.json"))
    frames = CF.frames_c2(nframes=5, scale=0.05)
    ctx = context()
    ref = ctx.render(frames)
    best = autotune.tune_launch(ctx, frames, A.PRR_KERR_BL, CF.tracer_kerr(precision=A.PRR_FP32_BULK),
                                budget_s=20.0, use_cache=False, tune_precision=False)
    assert best["block"] in autotune.BLOCKS and 1 <= best["refill"] <= 64
    out = ctx.render(frames)
    for a, b in zip(ref, out):
        assert np.array_equal(a["rgbf"], b["rgbf"])
    again = autotune.tune_launch(ctx, frames, A.PRR_KERR_BL, CF.tracer_kerr(precision=A.PRR_FP32_BULK),
                                 tune_precision=False)
    assert again["block"] == best["block"] and again["refill"] == best["refill"]


def test_autotuner_stage_precision_choice(tmp_path, monkeypatch):
    """For an FP32-bulk tracer the autotuner times the FP64 stages too and
    keeps the faster; the context then renders exactly what a render in the
    chosen precision gives, and the cached choice is applied again."""
    from parrelray_amd import autotune
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "tune.json"))
    frames = CF.frames_c2(nframes=5, scale=0.05)
    ctx = context()
    best = autotune.tune_launch(ctx, frames, A.PRR_KERR_BL, CF.tracer_kerr(precision=A.PRR_FP32_BULK),
                                budget_s=20.0, use_cache=False)
    assert best["precision"] in (A.PRR_FP32_BULK, A.PRR_FP64)
    assert min(best["integrate_ms_fp32_stages"], best["integrate_ms_fp64_stages"]) == best["integrate_ms"]
    out = ctx.render(frames)
    ref_ctx = context()
    ref_ctx.set_tracer(CF.tracer_kerr(precision=best["precision"]))
    ref = ref_ctx.render(frames)
    for a, b in zip(ref, out):
        assert np.array_equal(a["rgbf"], b["rgbf"])
    ctx2 = context()
    again = autotune.tune_launch(ctx2, frames, A.PRR_KERR_BL, CF.tracer_kerr(precision=A.PRR_FP32_BULK))
    assert again["precision"] == best["precision"]
    for a, b in zip(ref, ctx2.render(frames)):
        assert np.array_equal(a["rgbf"], b["rgbf"])


def test_invalid_inputs_are_rejected_without_touching_the_gpu_state():
    from parrelray_amd._lib import PrrError
    ctx = context()
    f = CF.frame_c1(32, 18)
    bad = A.Frame.from_buffer_copy(f)
    bad.samples = 5  # not a square for regular SSAA
    with pytest.raises(PrrError):
        ctx.render(bad)
    bad = A.Frame.from_buffer_copy(f)
    bad.spin = 1.5  # |a| > M
    with pytest.raises(PrrError):
        ctx.render(bad)
    with pytest.raises(PrrError):
        ctx.set_tracer(A.make_tracer(hmin=0.0))
    for w, h in ((0, 18), (32, 0), (70000, 2)):  # empty and oversized frames
        bad = A.Frame.from_buffer_copy(f)
        bad.width, bad.height = w, h
        with pytest.raises(PrrError):
            ctx.render(bad)
    with pytest.raises(PrrError):  # no frames at all
        ctx.render([])
    # the context still works
    out = ctx.render(f)[0]
    assert out["rgb8"].shape == (18, 32, 3)


def test_wormhole_inputs_are_checked():
    """Ellis wormhole: throat radius p[0] > 0, far escape radius p[1] >= 0,
    the camera outside the throat's areal radius, no spheres."""
    from parrelray_amd._lib import PrrError
    ctx = api.Context(0)
    for p in ([0.0] * 16, [-1.0] + [0.0] * 15, [2.0, -5.0] + [0.0] * 14):
        with pytest.raises(PrrError):
            ctx.set_spacetime(A.PRR_ELLIS, 0.0, 0.0, 0.0, p=p)
    with pytest.raises(PrrError):  # a charge belongs to the Schwarzschild-coordinate metric
        ctx.set_spacetime(A.PRR_ELLIS, 0.0, 0.0, 0.5, p=[2.0] + [0.0] * 15)
    ctx.set_spacetime(A.PRR_ELLIS, 0.0, 0.0, 0.0, p=[2.0] + [0.0] * 15)
    ctx.set_tracer(CF.tracer_kerr())
    inside = A.make_frame([1.0, 1.0, 0.5], 3.9, 0.0, 1.0, 8, 8, mass=0.0)
    with pytest.raises(PrrError, match="throat"):
        ctx.render(inside)
    sph = A.make_frame([20.0, 1.0, 0.5], 3.1, 0.0, 1.0, 8, 8, mass=0.0, spheres=[([5.0, 0.0, 0.0], 1.0, 1)])
    with pytest.raises(PrrError, match="spheres"):
        ctx.render(sph)
    ok = A.make_frame([20.0, 1.0, 0.5], 3.1, 0.0, 1.0, 8, 8, mass=0.0)
    assert ctx.render(ok)[0]["rgb8"].shape == (8, 8, 3)
    # the adjoint differentiates Kerr only
    with pytest.raises(PrrError, match="Kerr"):
        ctx.adjoint(CF.frame_c3(8, 8), CF.c3_orbit(), np.zeros((8, 8, 3), np.float32))
