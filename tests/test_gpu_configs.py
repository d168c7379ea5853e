"""The workload configurations at their full sizes on the GPU, against the
FP64 oracle at tol 1e-12 on the same seeded inputs:

* C2: the 64-frame trajectory (389,520,384 rays) in one persistent launch,
  FP32-bulk; a 20 x 20-pixel block of every frame (all its SSAA samples);
* C3: the 512 x 512 forward render that the adjoint differentiates, whole;
* C4: the 2048 x 2048 frame with 4 x 4 SSAA and FXAA (67,108,864 rays); a
  28 x 28-pixel crop across the shadow's edge (16 samples per pixel).

Per ray: the termination class agrees except for rays within 1e-6 rad of a
class boundary (the oracle's own +-1e-6 rad perturbations change their
class), and the endpoint direction of agreeing escape / sphere rays is
within the tolerance each test states.  Per pixel: RGB8 within 1 LSB; for
FXAA frames the oracle runs SSAA and FXAA on the block, so the 2-pixel
border of the block (whose FXAA taps reach outside it) is not compared.
"""
import numpy as np
import pytest

import oracle as O
from parrelray_amd import _abi as A
from parrelray_amd import api
from parrelray_amd import configs as CF

from parity import angle, boundary_excused, gpu_context, oracle_scene

pytestmark = pytest.mark.gpu
HALO = 2


def device_rays(torch, total):
    return dict(cls=torch.empty(total, dtype=torch.int8, device="cuda"),
                endpoint=torch.empty((total, 3), dtype=torch.float64, device="cuda"),
                color=torch.empty((total, 3), dtype=torch.float32, device="cuda"))


def block_ids(f, x0, y0, n):
    """Ray ids (within the frame) of the n x n pixel block at (x0, y0), ordered (row, column, sample)."""
    s = f.samples
    jj, ii, qq = np.meshgrid(np.arange(y0, y0 + n), np.arange(x0, x0 + n), np.arange(s), indexing="ij")
    return ((jj * f.width + ii) * s + qq).ravel()


def gather(torch, rays, base, ids):
    idx = torch.from_numpy(base + ids).cuda()
    return {k: v.index_select(0, idx).cpu().numpy() for k, v in rays.items()}


def check_block(S, f, gpu, ids, gpu_rgb8, x0, y0, n):
    """Classes, endpoint angles and RGB8 of one block against the oracle; returns its statistics."""
    orc = S.render_rays(f, ids)
    bad = np.nonzero(gpu["cls"] != orc["cls"])[0]
    exc = boundary_excused(S, f, ids[bad], orc["cls"][bad])
    same = gpu["cls"] == orc["cls"]
    m = same & ((orc["cls"] == A.PRR_ESCAPE) | (orc["cls"] == A.PRR_SPHERE))
    ang = angle(gpu["endpoint"][m], orc["endpoint"][m])
    img = O.ssaa(orc["color"], n, n, f.samples)
    if f.fxaa:
        img = O.fxaa(img)
    ref8 = O.rgb8(img)
    h = HALO if f.fxaa else 0
    g8 = gpu_rgb8[y0 + h:y0 + n - h, x0 + h:x0 + n - h].astype(np.int16)
    d = np.abs(g8 - ref8[h:n - h, h:n - h].astype(np.int16)).max(axis=-1)
    return dict(rays=ids.size, mismatched=int(bad.size), unexcused=int((~exc).sum()), angles=ang,
                pixels=int(d.size), rgb8_off=int((d > 1).sum()), classes=np.bincount(orc["cls"], minlength=5))


def summarize(name, stats):
    ang = np.concatenate([s["angles"] for s in stats])
    out = dict(rays=sum(s["rays"] for s in stats), mismatched=sum(s["mismatched"] for s in stats),
               unexcused=sum(s["unexcused"] for s in stats), pixels=sum(s["pixels"] for s in stats),
               rgb8_off=sum(s["rgb8_off"] for s in stats), angle_median=float(np.median(ang)),
               angle_p99=float(np.quantile(ang, 0.99)), angle_max=float(ang.max()), compared=int(ang.size),
               classes=np.sum([s["classes"] for s in stats], axis=0).tolist())
    print(name, out, flush=True)
    return out


def test_c2_full_batch_matches_oracle():
    """All 64 C2 frames in one persistent FP32-bulk launch (tol 1e-8); in each
    frame, of six seeded candidate 20 x 20 blocks the one whose GPU classes
    are most mixed (escape against the rest) is checked against the oracle:
    classes, endpoint angle median < 1e-6 and 99th percentile < 1e-5 rad,
    RGB8 within 1 LSB on >= 99.9 % of the compared pixels."""
    torch = pytest.importorskip("torch")
    frames = CF.frames_c2()
    total = sum(f.nrays for f in frames)
    assert total == 389520384
    ctx = gpu_context(A.PRR_KERR_BL, CF.tracer_kerr(precision=A.PRR_FP32_BULK))
    rays = device_rays(torch, total)
    rgb8 = [torch.empty((f.height, f.width, 3), dtype=torch.uint8, device="cuda") for f in frames]
    print("C2: one launch of", total, "rays", flush=True)
    ctx.render_device(frames, rgb8=rgb8, rays=rays)
    assert ctx.stats()["rays"] == total
    S = oracle_scene(A.PRR_KERR_BL, CF.tracer_kerr(tol=1e-12))
    rng = np.random.default_rng(0xC2)
    n, base, stats = 20, 0, []
    for k, f in enumerate(frames):
        cands = [(int(rng.integers(0, f.width - n)), int(rng.integers(0, f.height - n))) for _ in range(6)]
        best, score, data = None, -1, None
        for x0, y0 in cands:
            ids = block_ids(f, x0, y0, n)
            g = gather(torch, rays, base, ids)
            esc = int((g["cls"] == A.PRR_ESCAPE).sum())
            if min(esc, ids.size - esc) > score:
                best, score, data = (x0, y0, ids), min(esc, ids.size - esc), g
        x0, y0, ids = best
        st = check_block(S, f, data, ids, rgb8[k].cpu().numpy(), x0, y0, n)
        stats.append(st)
        base += f.nrays
    out = summarize("C2", stats)
    assert out["unexcused"] <= 3, out
    assert out["angle_median"] < 1e-6 and out["angle_p99"] < 1e-5, out
    assert out["rgb8_off"] <= 0.001 * out["pixels"], out
    # the blocks hold both sky and shadow (the most mixed candidates)
    assert min(out["classes"][:2]) > 1000, out


@pytest.mark.parametrize("precision", [A.PRR_FP64, A.PRR_FP32_BULK])
def test_c1_full_frame_matches_oracle(precision):
    """The whole C1 frame (1920 x 1080, 2 x 2 SSAA + FXAA, Kerr a = 0.95, tol
    1e-8) in one launch, in both precision forms; of 24 seeded candidate
    24 x 24-pixel blocks, the four whose GPU classes are most mixed and the
    two with the most sphere rays are checked against the oracle (tol 1e-12):
    classes, endpoint angle median < 1e-6 and 99th percentile < 1e-5 rad,
    RGB8 within 1 LSB on >= 99.9 % of the compared pixels (the oracle runs
    SSAA and FXAA on each block; its 2-pixel border is not compared)."""
    torch = pytest.importorskip("torch")
    f = CF.frame_c1()
    assert f.nrays == 8294400 and f.fxaa == 1
    ctx = gpu_context(A.PRR_KERR_BL, CF.tracer_kerr(precision=precision))
    rays = device_rays(torch, f.nrays)
    rgb8 = torch.empty((f.height, f.width, 3), dtype=torch.uint8, device="cuda")
    ctx.render_device([f], rgb8=[rgb8], rays=rays)
    assert ctx.stats()["rays"] == f.nrays
    rng = np.random.default_rng(0xC1)
    n = 24
    cands = []
    for _ in range(24):
        x0, y0 = int(rng.integers(0, f.width - n)), int(rng.integers(0, f.height - n))
        ids = block_ids(f, x0, y0, n)
        g = gather(torch, rays, 0, ids)
        counts = np.bincount(g["cls"], minlength=5)
        esc = int(counts[A.PRR_ESCAPE])
        cands.append((min(esc, ids.size - esc), int(counts[A.PRR_SPHERE]), x0, y0, ids, g))
    mixed = sorted(cands, key=lambda c: -c[0])[:4]
    rest = [c for c in cands if all(c is not m for m in mixed)]
    sphere = sorted(rest, key=lambda c: -c[1])[:2]
    S = oracle_scene(A.PRR_KERR_BL, CF.tracer_kerr(tol=1e-12))
    img = rgb8.cpu().numpy()
    stats = [check_block(S, f, g, ids, img, x0, y0, n) for _, _, x0, y0, ids, g in mixed + sphere]
    out = summarize("C1 %s" % ("FP64" if precision == A.PRR_FP64 else "FP32-bulk"), stats)
    assert out["unexcused"] <= 2, out
    assert out["angle_median"] < 1e-6 and out["angle_p99"] < 1e-5, out
    assert out["rgb8_off"] <= 0.001 * out["pixels"], out
    assert out["classes"][A.PRR_SPHERE] > 100 and min(out["classes"][:2]) > 1000, out


def test_c3_forward_render_matches_oracle():
    """The C3 render that the adjoint differentiates (512 x 512, Kerr a = 0.7,
    camera r = 15M, i = 75 deg, sphere at phi = 0.3; FP32-bulk tol 1e-8,
    bilinear textures), every ray and pixel against the oracle: classes,
    endpoint angle median < 1e-6 and 99th percentile < 1e-5 rad, RGB8 within
    1 LSB on >= 99.9 % of pixels."""
    torch = pytest.importorskip("torch")
    f = api.frame_from_orbit(CF.frame_c3(), CF.c3_orbit())
    ctx = gpu_context(A.PRR_KERR_BL, CF.tracer_kerr(precision=A.PRR_FP32_BULK), sphere_filter=A.PRR_FILTER_BILINEAR,
                      sky_filter=A.PRR_FILTER_BILINEAR)
    rays = device_rays(torch, f.nrays)
    rgb8 = torch.empty((f.height, f.width, 3), dtype=torch.uint8, device="cuda")
    rgbf = torch.empty((f.height, f.width, 3), dtype=torch.float32, device="cuda")
    ctx.render_device([f], rgb8=[rgb8], rgbf=[rgbf], rays=rays)
    S = oracle_scene(A.PRR_KERR_BL, CF.tracer_kerr(tol=1e-12), sphere_filter=A.PRR_FILTER_BILINEAR,
                     sky_filter=A.PRR_FILTER_BILINEAR)
    ids = block_ids(f, 0, 0, f.width)
    assert np.array_equal(ids, np.arange(f.nrays))
    g = {k: v.cpu().numpy() for k, v in rays.items()}
    print("C3: oracle over", f.nrays, "rays", flush=True)
    st = check_block(S, f, g, ids, rgb8.cpu().numpy(), 0, 0, f.width)
    out = summarize("C3", [st])
    assert out["unexcused"] <= 2, out
    assert out["angle_median"] < 1e-6 and out["angle_p99"] < 1e-5, out
    assert out["rgb8_off"] <= 0.001 * out["pixels"], out
    assert min(out["classes"][:3]) > 1000, out  # sky, shadow and sphere
    # the linear image the adjoint differentiates is the SSAA of the ray colours (no FXAA in C3)
    assert np.array_equal(rgbf.cpu().numpy(), O.ssaa(g["color"], f.width, f.height, f.samples))


C4_CROP = (1970, 1010, 28)


@pytest.fixture(scope="module")
def c4_oracle():
    """The oracle's rays of the C4 crop (FP64 RKF45 tol 1e-12), shared by both precision forms."""
    f = CF.frame_c4()
    x0, y0, n = C4_CROP
    ids = block_ids(f, x0, y0, n)
    S = oracle_scene(A.PRR_KERR_BL, CF.tracer_kerr(tol=1e-12, max_steps=400000))
    return S, f, ids


@pytest.mark.parametrize("precision", [A.PRR_FP64, A.PRR_FP32_BULK])
def test_c4_crop_with_4x4_ssaa_and_fxaa_matches_oracle(c4_oracle, precision):
    """The full C4 frame (a = 0.998, camera r = 6M at 89.9 deg, 2048 x 2048,
    4 x 4 SSAA + FXAA, tol 1e-10) in one launch, in both precision forms; the
    28 x 28-pixel crop at (1970, 1010), which straddles the edge of the shadow
    on the prograde side, against the oracle: its 12,544 rays and the RGB8 of
    the inner 24 x 24 pixels after the oracle's SSAA and FXAA of the crop.
    These are the frame's most sensitive rays (they wind near the photon
    orbit, which amplifies the 1e-10 local error): classes agree, endpoint
    angle median < 1e-6 and 99th percentile < 5e-5 rad (measured 1.9e-7 and
    1.3e-5 in the FP32-bulk form)."""
    torch = pytest.importorskip("torch")
    S, f, ids = c4_oracle
    assert f.samples == 16 and f.fxaa == 1 and f.nrays == 67108864
    ctx = gpu_context(A.PRR_KERR_BL, CF.tracer_c4(precision=precision))
    rays = device_rays(torch, f.nrays)
    rgb8 = torch.empty((f.height, f.width, 3), dtype=torch.uint8, device="cuda")
    ctx.render_device([f], rgb8=[rgb8], rays=rays)
    assert ctx.stats()["rays"] == f.nrays
    x0, y0, n = C4_CROP
    g = gather(torch, rays, 0, ids)
    del rays
    st = check_block(S, f, g, ids, rgb8.cpu().numpy(), x0, y0, n)
    out = summarize("C4 crop %s" % ("FP64" if precision == A.PRR_FP64 else "FP32-bulk"), [st])
    assert out["classes"][A.PRR_ESCAPE] > 2000 and out["classes"][A.PRR_HORIZON] > 2000, out
    assert out["unexcused"] <= 2, out
    assert out["angle_median"] < 1e-6 and out["angle_p99"] < 5e-5, out
    assert out["rgb8_off"] <= 0.001 * out["pixels"], out
