"""The .params grammar and scene set-up (host C++ in libprr.so; no GPU).

Fixtures in tests/golden/params follow the reference's grammar
(main/io.h:21-59) including its quirks: any '_' makes a line a header (even
inside a value), only the first <...> on each side counts, a line without
'=' still yields a key/value pair, and lookups of missing keys return
defaults (main/Parameters.cpp:20-43).
"""
import math
import os
import shutil

import numpy as np

import pytest

from parrelray_amd import _abi as A
from parrelray_amd import api
from parrelray_amd._lib import PrrError

HERE = os.path.join(os.path.dirname(__file__), "golden", "params")


@pytest.fixture()
def sim(tmp_path, monkeypatch):
    # relative paths, as in the reference's input/sim.params: an absolute
    # temporary path may contain '_', which would make the line a header
    for name in ("traj.params", "scene.params"):
        shutil.copy(os.path.join(HERE, name), tmp_path / name)
    text = open(os.path.join(HERE, "sim.params")).read()
    text = text.replace("<TRAJ>", "<traj.params>").replace("<SCENE>", "<scene.params>")
    p = tmp_path / "sim.params"
    p.write_text(text)
    monkeypatch.chdir(tmp_path)
    return str(p)


def test_underscore_in_a_path_value_makes_a_header(tmp_path):
    p = tmp_path / "q.params"
    p.write_text("_tracer_\n<traj> = <dir_name/traj.params>\n<h> = <0.5>\n")
    P = api.Params(str(p))
    assert not P.par_exists("tracer", "traj")
    assert P.parf("name/traj.params>", "h") == 0.5  # block = text between the first and last '_'


def test_grammar_and_quirks(sim):
    P = api.Params(sim)
    assert P.par("tracer", "solver") == "rkf"  # trailing comment ignored
    assert P.parf("tracer", "hmin") == 1e-4
    assert P.par("tracer", "bimage2") == ""
    assert P.par_exists("tracer", "bimage2")
    assert P.par("comments", "comment") == "A parameter file for the tests"
    # '_' inside a value turns the line into a header: "ffmpeg command" is lost
    assert not P.par_exists("comments", "ffmpeg command")
    # a line without '=' still produces key and value from its first <...>
    assert P.par("camera", "stray") == "stray"
    # defaults of missing keys
    assert P.par("nope", "x") == "" and P.pari("nope", "x") == 0 and P.parf("nope", "x") == 0.0
    assert P.parf("nope", "x", 3.5) == 3.5
    assert not P.header_exists("nope")
    assert P.pari("camera", "width") == 32


def test_load_sim_reads_traj_and_scene(sim):
    P = api.Params(sim=sim)
    assert P.frame_count() == 3  # frame3 missing: frame4 is never reached (checkFrame, prr.cpp:139-153)
    st = P.spacetime()
    assert st["metric"] == A.PRR_SCHW_KS_CART  # the reference's compile-time default
    assert st["M"] == 1.5 and st["a"] == 0.25 and st["p"][3] == 2.5
    t = P.tracer()
    assert t.solver == A.PRR_RKF45 and t.atol == 1e-3 and t.rtol == 0.0
    assert t.hmin == 1e-4 and t.h == 0.01 and t.hmax == 0.1 and t.r_escape == 70.0
    assert t.lambda_max == 1000.0 and t.max_attempts == 10 ** 6
    assert t.horizon_eps_abs == 1e-3 and t.horizon_eps_rel == 0.0


def test_frames_sticky_mass_and_sphere_updates(sim):
    P = api.Params(sim=sim)
    f0, f1, f2 = P.frame(0), P.frame(1), P.frame(2)
    assert list(f0.pos) == [30.0, 0.0, 2.0]
    assert abs(f0.alpha - math.pi) < 1e-15 and abs(f0.beta - math.radians(5)) < 1e-15
    assert abs(f0.vfov - math.radians(60)) < 1e-15
    assert (f0.width, f0.height, f0.ssaa, f0.samples) == (32, 24, A.PRR_SSAA_REGULAR, 4)
    assert f0.projection == A.PRR_PROJ_SPHERICAL  # Cartesian metrics use rayDir_Spherical (kernels/solver.cuh:39)
    assert (f0.mass, f1.mass, f2.mass) == (1.5, 2.0, 2.0)  # per-frame mass is sticky (prr.cpp:253)
    assert f0.n_spheres == 2
    assert list(f0.spheres[0].c) == [-10.0, 0.0, 0.0] and f0.spheres[0].radius == 2.0
    assert list(f1.spheres[0].c) == [-9.0, 0.0, 0.0] and f1.spheres[0].radius == 2.0
    assert list(f2.spheres[0].c) == [-9.0, 0.0, 0.0] and f2.spheres[0].radius == 3.0
    assert f0.spheres[1].tex == 2
    with pytest.raises(PrrError):
        P.frame(3)


def test_solver_names(tmp_path):
    def tracer_for(name):
        p = tmp_path / f"{name}.params"
        p.write_text(f"_tracer_\n<solver> = <{name}>\n<h> = <0.01>\n")
        return api.Params(str(p)).tracer()
    assert tracer_for("euler").solver == A.PRR_EULER
    assert tracer_for("verlet").solver == A.PRR_EULER  # unknown -> euler (main/Parameters.h:18-22)
    assert tracer_for("rkf").solver == A.PRR_RKF45
    assert tracer_for("dormand").solver == A.PRR_DORMAND  # declared by the reference without a kernel
    assert tracer_for("rk").solver == A.PRR_RK4
    for name in ("adams", "gragg", "taylor"):  # declared there, not implemented there or here
        with pytest.raises(PrrError):
            tracer_for(name)


def test_ssaa_names(tmp_path):
    def frame_for(name, samples):
        p = tmp_path / f"{name}.params"
        p.write_text("_frame0_\n<x> = <20>\n<y> = <0>\n<z> = <1>\n<a> = <180>\n<b> = <3>\n"
                     f"_camera_\n<vfov> = <50>\n<width> = <8>\n<height> = <4>\n<SSAA> = <{name}>\n"
                     f"<SSAAsamples> = <{samples}>\n<SSAAseed> = <77>\n")
        return api.Params(str(p)).frame(0)
    assert frame_for("random", 5).ssaa == A.PRR_SSAA_RANDOM
    f = frame_for("jittered", 9)
    assert f.ssaa == A.PRR_SSAA_JITTERED and f.samples == 9 and f.seed == 77
    assert frame_for("regular", 4).ssaa == A.PRR_SSAA_REGULAR
    with pytest.raises(PrrError):
        frame_for("poisson", 4)


def test_new_keys_live_in_their_own_block(tmp_path):
    p = tmp_path / "k.params"
    p.write_text("_prr_\n<metric> = <kerr>\n<precision> = <fp32-bulk>\n<rtol> = <1e-8>\n<projection> = <pinhole>\n"
                 "_spacetime_\n<hasBlackHole> = <1>\n<mass> = <1>\n<spin> = <0.9>\n"
                 "_frame0_\n<x> = <20>\n<y> = <0>\n<z> = <1>\n<a> = <180>\n<b> = <3>\n"
                 "_camera_\n<vfov> = <50>\n<width> = <8>\n<height> = <4>\n<SSAA> = <none>\n")
    P = api.Params(str(p))
    assert P.spacetime()["metric"] == A.PRR_KERR_BL
    t = P.tracer()
    assert t.precision == A.PRR_FP32_BULK and t.rtol == 1e-8 and t.solver == A.PRR_EULER
    f = P.frame(0)
    assert f.projection == A.PRR_PROJ_PINHOLE and f.samples == 1 and f.spin == 0.9


def test_wormhole_reads_throat_and_far_escape_radius(tmp_path):
    """metric = wormhole: p0 is the throat radius; p1, the escape radius
    beyond the throat, defaults to the tracer block's rout2 (prr.cpp:190)."""
    def spacetime(extra_tracer, extra_st):
        p = tmp_path / "w.params"
        p.write_text("_tracer_\n<rout1> = <60>\n" + extra_tracer + "_prr_\n<metric> = <wormhole>\n"
                     "_spacetime_\n<hasBlackHole> = <0>\n<p0> = <2.5>\n" + extra_st)
        return api.Params(str(p)).spacetime()
    st = spacetime("<rout2> = <45>\n", "")
    assert st["metric"] == A.PRR_ELLIS and st["p"][0] == 2.5 and st["p"][1] == 45.0
    assert spacetime("<rout2> = <45>\n", "<p1> = <30>\n")["p"][1] == 30.0
    assert spacetime("", "")["p"][1] == 0.0  # 0: the near side's r_escape


def test_missing_file_is_an_io_error():
    with pytest.raises(PrrError) as e:
        api.Params("/nonexistent/sim.params")
    assert e.value.status == A.PRR_EIO


def test_procedural_textures_are_seeded_and_deterministic():
    a = api.checker_texture(512, 256, 16, 8, 1)
    b = api.checker_texture(512, 256, 16, 8, 1)
    c = api.checker_texture(512, 256, 16, 8, 2)
    assert (a == b).all() and not (a == c).all()
    # cells are constant 32x32 blocks
    assert (a[:32, :32] == a[0, 0]).all() and not (a[0, 0] == a[0, 32]).all()
    n = api.noise_texture(1024, 512, 4, 3)
    assert n.shape == (512, 1024, 3) and 0 < n.mean() < 255 and n.std() > 5
    # value noise wraps in x (the sphere texture's phi seam): columns 0 and W-1 are close
    assert abs(n[:, 0].astype(int) - n[:, -1].astype(int)).max() <= 12


CONFIGS = os.path.join(os.path.dirname(os.path.dirname(__file__)), "configs")


@pytest.mark.parametrize("name", ["c0", "c1"])
def test_config_files_match_the_benchmark_workloads(name, monkeypatch):
    from parrelray_amd import configs as CF
    monkeypatch.chdir(CONFIGS)
    P = api.Params(sim=f"{name}.params")
    f = P.frame(0)
    if name == "c0":
        ref_f, ref_t, metric = CF.frame_c0(), CF.tracer_c0(), A.PRR_SCHW_SPH
    else:
        ref_f, ref_t, metric = CF.frame_c1(), CF.tracer_kerr(precision=A.PRR_FP32_BULK), A.PRR_KERR_BL
    assert P.spacetime()["metric"] == metric
    t = P.tracer()
    for k, _ in A.Tracer._fields_:
        if ref_t.solver == A.PRR_EULER and k in ("hmin", "hmax", "atol", "rtol", "max_attempts"):
            continue  # unused by Euler
        assert getattr(t, k) == pytest.approx(getattr(ref_t, k), rel=1e-15), k
    for k in ("width", "height", "ssaa", "samples", "projection", "fxaa", "n_spheres"):
        assert getattr(f, k) == getattr(ref_f, k), k
    for k in ("alpha", "beta", "vfov", "mass", "spin"):
        assert getattr(f, k) == pytest.approx(getattr(ref_f, k), rel=1e-14, abs=1e-15), k
    assert list(f.pos) == pytest.approx(list(ref_f.pos), rel=1e-15)
    for s, rs in zip(f.spheres[:f.n_spheres], ref_f.spheres[:ref_f.n_spheres]):
        assert list(s.c) == pytest.approx(list(rs.c), rel=1e-15) and s.radius == rs.radius


@pytest.mark.parametrize("w,h,ssaa,s", [(37, 23, A.PRR_SSAA_NONE, 1), (64, 48, A.PRR_SSAA_REGULAR, 4),
                                        (50, 7, A.PRR_SSAA_REGULAR, 16), (9, 5, A.PRR_SSAA_RANDOM, 2),
                                        (13, 11, A.PRR_SSAA_REGULAR, 9)])
def test_queue_order_is_a_permutation_of_tiles(w, h, ssaa, s):
    """The integrator's queue visits every ray of a frame exactly once; each
    wave's 64 positions cover a compact tile of whole pixels where 64 / s
    pixels fit (s = 9 keeps row order)."""
    f = A.make_frame([30, 0, 0], math.pi, 0, 1.0, w, h, ssaa=ssaa, samples=s)
    n = w * h * s
    ids = api.queue_rays(f, np.arange(n))
    assert np.array_equal(np.sort(ids), np.arange(n))
    if 64 % s:
        assert np.array_equal(ids, np.arange(n))
        return
    pix = ids // s
    px, py = pix % w, pix // w
    for b in range(n // 64):
        sl = slice(64 * b, 64 * b + 64)
        # whole pixels: every sample of each pixel in the same wave
        assert np.all(np.bincount(pix[sl] - pix[sl].min()) % s == 0)
    # the first wave is a tile: spread in x and y both below 64 / s pixels (a row run's x spread)
    first = slice(0, 64)
    if n >= 64 and w >= 8 and h >= 8:
        assert np.ptp(px[first]) < 64 // s or 64 // s == 1
        assert np.ptp(py[first]) >= 1 or 64 // s == 1
