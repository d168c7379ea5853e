"""Pins of the oracle's image-side functions (CPU, no GPU): the SSAA combine,
FXAA, RGB8 quantisation, the celestial-sphere lookup and the bilinear texture
filter, each against hand-computed values, exact identities or symmetries
that a wrong index, sign or transposition breaks.

References: combineSamples kernels/ssaa.cuh:109-131, luma / fxaa
kernels/ssaa.cuh:141-182, RGB8 cudamaths/image.cu:31, getColorDefault2
main/base.h:119-162, sqavg cudamaths/f3.cuh:101-103, tex2d bilinear
cudamaths/tex2d.cuh:112-119.
"""
import math

import numpy as np
import pytest

import oracle as O
from parrelray_amd import _abi as A

RNG = np.random.default_rng(11)


# ------------------------------------------------------------------ SSAA
def test_ssaa_single_sample_is_a_clamp():
    c = np.array([[-0.5, 0.25, 1.75], [0.0, 1.0, 0.5]], np.float32)
    img = O.ssaa(c, 2, 1, 1)
    assert np.array_equal(img.reshape(-1, 3), np.clip(c, 0, 1))


def test_ssaa_is_the_rms_of_clamped_squares_in_sample_order():
    # 2x1 image, 4 samples per pixel; pixel p owns rays p*4 .. p*4+3
    rays = np.zeros((8, 3), np.float32)
    rays[0:4, 0] = [0.0, 1.0, 0.0, 1.0]   # pixel 0 red: sqrt(2/4)
    rays[0:4, 1] = [2.0, 0.0, 0.0, 0.0]   # pixel 0 green: the square clamps to 1 -> sqrt(1/4)
    rays[4:8, 2] = [0.6, 0.6, 0.6, 0.6]   # pixel 1 blue: constant
    img = O.ssaa(rays, 2, 1, 4)
    assert img[0, 0, 0] == pytest.approx(math.sqrt(0.5), abs=1e-7)
    assert img[0, 0, 1] == pytest.approx(0.5, abs=1e-7)
    assert img[0, 0, 2] == 0.0
    assert img[0, 1, 2] == pytest.approx(0.6, abs=1e-7)
    assert img[0, 1, 0] == 0.0 and img[0, 1, 1] == 0.0


def test_ssaa_ignores_the_order_of_samples_within_a_pixel():
    W, H, s = 5, 3, 9
    rays = RNG.uniform(-0.2, 1.2, size=(W * H * s, 3)).astype(np.float32)
    perm = rays.reshape(W * H, s, 3)[:, ::-1, :].reshape(-1, 3)
    a, b = O.ssaa(rays, W, H, s), O.ssaa(perm, W, H, s)
    assert np.allclose(a, b, atol=1e-7)
    # and pixel (i, j) is built from rays ((j W + i) s .. + s)
    j, i = 2, 3
    blk = rays[(j * W + i) * s:(j * W + i + 1) * s]
    assert np.allclose(a[j, i], np.sqrt(np.clip(blk.astype(np.float64) ** 2, 0, 1).mean(axis=0)), atol=1e-6)


# ------------------------------------------------------------------ FXAA
def fxaa_image():
    """A 24x20 image with smooth shading, a diagonal edge and noise."""
    H, W = 20, 24
    y, x = np.mgrid[0:H, 0:W]
    img = np.zeros((H, W, 3), np.float32)
    img[..., 0] = 0.2 + 0.6 * (x + 0.7 * y > 14)
    img[..., 1] = 0.5 + 0.3 * np.sin(0.4 * x) * np.cos(0.3 * y)
    img[..., 2] = RNG.uniform(0, 1, size=(H, W))
    return img


def test_fxaa_leaves_constant_images_unchanged():
    img = np.full((9, 7, 3), 0.37, np.float32)
    assert np.array_equal(O.fxaa(img), img)


@pytest.mark.parametrize("axis", [0, 1])
def test_fxaa_leaves_straight_axis_aligned_edges_unchanged(axis):
    # the search direction of an axis-aligned edge runs along it, so every
    # tap stays on the pixel's own side of the edge
    img = np.full((12, 12, 3), 0.2, np.float32)
    if axis == 0:
        img[6:, :, :] = 0.9
    else:
        img[:, 6:, :] = 0.9
    # (up to the rounding of bilinear weights that sum to 1; a tap across the
    # edge would move a pixel by ~0.1 or more)
    assert np.abs(O.fxaa(img) - img).max() <= 1e-7


@pytest.mark.parametrize("op", ["flip_x", "flip_y", "transpose"])
def test_fxaa_commutes_with_the_symmetries_of_the_pixel_grid(op):
    # the taps are symmetric in the search direction (p +- dir / 6, p +- dir / 2),
    # so FXAA is equivariant under flips and the transpose; a swapped sign or
    # neighbour in the direction formula breaks this
    f = {"flip_x": lambda a: a[:, ::-1], "flip_y": lambda a: a[::-1, :],
         "transpose": lambda a: a.transpose(1, 0, 2)}[op]
    img = fxaa_image()
    a = f(O.fxaa(img))
    b = O.fxaa(np.ascontiguousarray(f(img)))
    assert np.abs(a - b).max() <= 2e-6


def test_fxaa_smooths_a_staircase_and_stays_in_range():
    img = fxaa_image()
    out = O.fxaa(img)
    assert out.min() >= img.min() - 1e-6 and out.max() <= img.max() + 1e-6
    # the red channel's diagonal staircase gets intermediate values
    changed = np.abs(out[..., 0] - img[..., 0]) > 1e-3
    assert changed.any()
    inter = out[..., 0][changed]
    assert np.all((inter > 0.2 - 1e-6) & (inter < 0.8 + 1e-6))


# ------------------------------------------------------------------ RGB8
def test_rgb8_rounds_half_away_from_zero_after_the_fp32_product():
    k = np.arange(256, dtype=np.float64)
    c = np.concatenate([k / 255, (k + 0.5) / 255, (k + 0.49) / 255, (k + 0.51) / 255]).astype(np.float32)
    # ties of 255 * c as the FP32 product sees them
    c = np.concatenate([c, (np.arange(256, dtype=np.float32) + np.float32(0.5)) / np.float32(255)])
    c = c[c <= 1.0]
    prod = (np.float32(255) * c).astype(np.float64)  # exact FP32 product, widened
    expect = np.floor(prod + 0.5)                    # round half up (values >= 0)
    got = O.rgb8(c.reshape(-1, 1, 1).repeat(3, axis=2))[:, 0, 0]
    assert np.array_equal(got, expect.astype(np.uint8))
    assert (prod == np.floor(prod) + 0.5).any()      # the set holds exact ties


# ------------------------------------------------------------------ textures
def one_texel_texture(w, h, i, j):
    t = np.zeros((h, w, 3), np.uint8)
    t[j, i, 0] = 255
    t[j, i, 1] = 51
    return t


def test_sky_lookup_block_rms_and_orientation():
    W, H = 16, 8
    i, j = 5, 3
    S = O.Scene(A.PRR_FLAT)
    S.set_texture(0, one_texel_texture(W, H, i, j))

    def look(px, py):  # centre of texel (px, py): phi along the width, theta along the height
        return S.sky_lookup((py + 0.5) / H * math.pi, (px + 0.5) / W * 2 * math.pi)

    # the 2x2 blocks (px..px+1, py..py+1) that contain the lit texel: RMS = sqrt(1/4) = 1/2
    for px, py in ((i, j), (i - 1, j), (i, j - 1), (i - 1, j - 1)):
        got = look(px, py)
        assert got[0] == pytest.approx(0.5, abs=1e-7) and got[1] == pytest.approx(0.1, abs=1e-7)
        assert got[2] == 0.0
    for px, py in ((i + 1, j), (i, j + 1), (i - 2, j), (j, i)):
        assert np.all(look(px, py) == 0.0), (px, py)


def test_sky_lookup_last_row_and_column_take_a_single_texel():
    W, H = 16, 8
    S = O.Scene(A.PRR_FLAT)
    S.set_texture(0, one_texel_texture(W, H, W - 1, 2))
    got = S.sky_lookup((2 + 0.5) / H * math.pi, (W - 0.5) / W * 2 * math.pi)
    assert got[0] == pytest.approx(1.0) and got[1] == pytest.approx(0.2)
    S.set_texture(0, one_texel_texture(W, H, 4, H - 1))
    got = S.sky_lookup((H - 0.5) / H * math.pi, (4 + 0.5) / W * 2 * math.pi)
    assert got[0] == pytest.approx(1.0)


def test_bilinear_reproduces_a_linear_texture():
    # texel (i, j) holds (3 i + 7 j) / 255 in red and (5 i) / 255 in green;
    # texel centres at integer coordinates: X = u W, Y = v H
    W, H = 32, 16
    jj, ii = np.mgrid[0:H, 0:W]
    t = np.zeros((H, W, 3), np.uint8)
    t[..., 0] = 3 * ii + 7 * jj
    t[..., 1] = 5 * ii
    S = O.Scene(A.PRR_FLAT)
    S.set_texture(1, t, A.PRR_FILTER_BILINEAR)
    for _ in range(200):
        X, Y = RNG.uniform(0, W - 1), RNG.uniform(0, H - 1)
        got = S.bilinear(1, X / W, Y / H)
        assert got[0] == pytest.approx((3 * X + 7 * Y) / 255, abs=2e-6)
        assert got[1] == pytest.approx(5 * X / 255, abs=2e-6)
    # u wraps: between the last column and column 0
    got = S.bilinear(1, (W - 0.25) / W, 2.0 / H)
    assert got[1] == pytest.approx((0.25 * 5 * (W - 1) + 0.75 * 0) / 255, abs=2e-6)
    # v clamps to the last row
    assert S.bilinear(1, 3.0 / W, 1.7)[0] == pytest.approx((3 * 3 + 7 * (H - 1)) / 255, abs=2e-6)


def test_sphere_colour_is_the_texture_at_the_normal_times_the_shade():
    """Sphere::getColor (scene/sphere.h:59-64): the texture at
    u = (atan2(n_y, n_x) + pi) / 2 pi, v = acos(n_z) / pi of the unit normal n
    at the hit, times clamp(0.5 + 0.5 n_z, 0, 1); untextured spheres are white.
    Flat space, a camera at r = 20 looking at two spheres: the hit point (pinned
    by the ray/sphere tests) names the sphere and its normal, the texture
    filter is pinned by the bilinear test above."""
    tex = RNG.integers(0, 256, size=(32, 64, 3), dtype=np.uint8)
    tracer = A.make_tracer(solver=A.PRR_RKF45, h=0.05, hmin=1e-9, hmax=2.0, atol=1e-12, rtol=1e-12, r_escape=60.0,
                           horizon_eps_rel=1e-3, horizon_eps_abs=0.0, lambda_max=1e9, max_steps=10**6,
                           max_attempts=10**7)
    S = O.Scene(A.PRR_FLAT, 0.0, 0.0, tracer)
    S.set_texture(1, tex, A.PRR_FILTER_BILINEAR)
    spheres = [(np.array([0.0, -3.0, 0.5]), 2.5, 1), (np.array([0.0, 3.5, -0.5]), 2.0, 0)]
    f = A.make_frame([20.0, 0.0, 0.0], math.pi, 0.0, math.radians(20.0), 48, 32, mass=0.0,
                     spheres=[(list(c), R, t) for c, R, t in spheres])
    r = S.render_rays(f, want_state=True)
    hit = r["cls"] == A.PRR_SPHERE
    assert hit.sum() > 200
    counts = [0, 0]
    for P, n, col in zip(r["state8"][hit, :3], r["endpoint"][hit], r["color"][hit]):
        k = int(np.argmin([abs(np.linalg.norm(P - c) - R) for c, R, _ in spheres]))
        c, R, t = spheres[k]
        m = (P - c) / R
        assert np.abs(n - m).max() < 1e-9
        shade = np.float32(min(max(np.float32(0.5) + np.float32(0.5) * np.float32(m[2]), 0.0), 1.0))
        if t:
            u = (math.atan2(m[1], m[0]) + math.pi) / (2 * math.pi)
            v = math.acos(max(-1.0, min(1.0, m[2]))) / math.pi
            expect = S.bilinear(t, u, v) * shade
        else:
            expect = np.full(3, shade, np.float32)
        assert np.abs(col - expect).max() < 1e-6, (k, m, col, expect)
        counts[k] += 1
    assert min(counts) > 100
