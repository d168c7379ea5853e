"""The `prr -t sim.params` tool end to end on the GPU (prr.cpp:80-135 cmdinput,
:139-341 renderframe, :345-404 main): it reads the sim / traj / scene files,
renders every complete frameN block and writes <out>image_<i>.png, which
must hold the same RGB8 image as the library renders through the C ABI for
the frame that the same files describe.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from parrelray_amd import _abi as A
from parrelray_amd import api

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PRR = os.path.join(ROOT, "parrelray_amd", "lib", "prr")
CONFIGS = os.path.join(ROOT, "configs")


def read_png(path):
    from PIL import Image
    with Image.open(path) as im:
        assert im.mode == "RGB"
        return np.asarray(im).copy()


def run_prr(cwd, sim):
    print("running", PRR, "-t", sim, "in", cwd, flush=True)
    r = subprocess.run([PRR, "-t", sim], cwd=cwd, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


def library_render(tmp_path, monkeypatch, i):
    """The same frame through the C ABI, set up from the same .params files
    (relative paths in them resolve against the working directory, as in the
    reference's main)."""
    monkeypatch.chdir(tmp_path)
    p = api.Params(sim="c0.params")
    st, tr = p.spacetime(), p.tracer()
    ctx = api.Context(0)
    ctx.set_spacetime(st["metric"], st["M"], st["a"], st["q"])
    ctx.set_tracer(tr)
    ctx.set_texture(0, api.checker_texture(512, 256, 16, 8, 1592590337), A.PRR_FILTER_SQAVG)
    f = p.frame(i)
    return ctx.render(f, rays=False)[0]["rgb8"]


def prepare(tmp_path, traj=None, sky=None):
    """configs/c0.params at 128 x 128 with a 40,000-step cap (the C0 cap of
    10,000 Euler steps of 0.01 M stops rays aimed at the hole before they get
    back out to r = 100 M, which leaves that image black)."""
    text = open(os.path.join(CONFIGS, "c0.params")).read()
    for old, new in (("<max-steps> = <10000>", "<max-steps> = <40000>"), ("<width> = <256>", "<width> = <128>"),
                     ("<height> = <256>", "<height> = <128>")):
        assert old in text
        text = text.replace(old, new)
    if sky:
        text = text.replace("checker:512:256:16:8:1592590337", sky)
    (tmp_path / "c0.params").write_text(text)
    if traj:
        (tmp_path / "c0.traj").write_text(traj)
    else:
        shutil.copy(os.path.join(CONFIGS, "c0.traj"), tmp_path / "c0.traj")


def test_c0_params_render_to_png_like_the_library(tmp_path, monkeypatch):
    prepare(tmp_path)
    out = run_prr(tmp_path, "c0.params")
    assert "frame 1/1" in out
    img = read_png(tmp_path / "c0-image_0.png")
    assert img.shape == (128, 128, 3)
    assert np.array_equal(img, library_render(tmp_path, monkeypatch, 0))
    # a black hole in the middle, sky around it
    assert img[64, 64].max() == 0 and img[3, 3].max() > 0


def test_trajectory_frames_are_rendered_in_order(tmp_path, monkeypatch):
    # three frames moving the camera outwards along the same line of sight
    blocks = []
    for i, r in enumerate((20.0, 30.0, 45.0)):
        blocks.append(f"_frame{i}_\n<x> = <{r * 0.9961946980917455}>\n<y> = <0.0>\n"
                      f"<z> = <{r * 0.08715574274765817}>\n<a> = <180>\n<b> = <5>\n")
    prepare(tmp_path, traj="three frames\n\n" + "\n".join(blocks))
    out = run_prr(tmp_path, "c0.params")
    assert "frame 3/3" in out
    imgs = [read_png(tmp_path / f"c0-image_{i}.png") for i in range(3)]
    # the shadow shrinks as the camera recedes
    dark = [int((im.max(axis=2) == 0).sum()) for im in imgs]
    assert dark[0] > dark[1] > dark[2] > 0, dark
    assert np.array_equal(imgs[1], library_render(tmp_path, monkeypatch, 1))


def test_sky_texture_from_a_png_file_renders_like_the_procedural_one(tmp_path, monkeypatch):
    from PIL import Image
    Image.fromarray(api.checker_texture(512, 256, 16, 8, 1592590337)).save(tmp_path / "sky.png")
    prepare(tmp_path, sky="sky.png")
    run_prr(tmp_path, "c0.params")
    img = read_png(tmp_path / "c0-image_0.png")
    assert np.array_equal(img, library_render(tmp_path, monkeypatch, 0))
    assert img[3, 3].max() > 0


def test_wormhole_with_a_second_sky(tmp_path, monkeypatch):
    """metric = wormhole with throat radius p0 = 8: the C0 camera at r = 30
    (its centre ray passes the centre at impact parameter ~5) looks into the
    throat, whose disc shows bimage2 (the sky beyond it)."""
    prepare(tmp_path)
    text = (tmp_path / "c0.params").read_text()
    text = text.replace("<metric> = <schwarzschild-sph>", "<metric> = <wormhole>")
    text = text.replace("<bimage1> = <checker:512:256:16:8:1592590337>",
                        "<bimage1> = <checker:512:256:16:8:1592590337>\n<bimage2> = <checker:256:128:4:2:77>")
    text = text.replace("<spin> = <0.0>", "<spin> = <0.0>\n<p0> = <8.0>")
    text = text.replace("<rout1> = <100.0>", "<rout1> = <100.0>\n<rout2> = <70.0>")
    (tmp_path / "c0.params").write_text(text)
    run_prr(tmp_path, "c0.params")
    img = read_png(tmp_path / "c0-image_0.png")
    monkeypatch.chdir(tmp_path)
    p = api.Params(sim="c0.params")
    st, tr = p.spacetime(), p.tracer()
    assert st["metric"] == A.PRR_ELLIS and st["p"][0] == 8.0 and st["p"][1] == 70.0
    ctx = api.Context(0)
    ctx.set_spacetime(st["metric"], st["M"], st["a"], st["q"], p=st["p"])
    ctx.set_tracer(tr)
    ctx.set_texture(0, api.checker_texture(512, 256, 16, 8, 1592590337), A.PRR_FILTER_SQAVG)
    ctx.set_texture(A.PRR_SKY2_SLOT, api.checker_texture(256, 128, 4, 2, 77), A.PRR_FILTER_SQAVG)
    f = p.frame(0)
    assert np.array_equal(img, ctx.render(f, rays=False)[0]["rgb8"])
    # the throat's disc is lit by the second sky; without it, it is black
    assert img[64, 64].max() > 0
    ctx.set_texture(A.PRR_SKY2_SLOT, None)
    assert ctx.render(f, rays=False)[0]["rgb8"][64, 64].max() == 0
