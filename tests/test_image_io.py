"""Texture files (prr_image_read, CPU only): binary PPM, PNG and baseline
JPEG decoded to RGB8, the reference's stbi_load(path, ..., 3)
(cudamaths/image.cu:13-21).  Files come from Pillow (an independent encoder)
or are assembled here from the PNG specification (one file per filter
type, checked by Pillow's decoder first); JPEG output is compared with
Pillow's own decode of the same file and with the original image.
"""
import struct
import zlib

import numpy as np
import pytest
from PIL import Image

from parrelray_amd import api
from parrelray_amd._lib import PrrError

RNG = np.random.default_rng(5)


def smooth_image(h=61, w=83):
    """Gradients, a sharp disc and mild noise: every PNG filter has work to do."""
    y, x = np.mgrid[0:h, 0:w]
    img = np.stack([255 * x / (w - 1), 255 * y / (h - 1), 128 + 100 * np.sin(0.2 * x + 0.1 * y)], axis=-1)
    img[(x - 40) ** 2 + (y - 30) ** 2 < 150] = [250, 20, 60]
    img += RNG.normal(0, 3, size=img.shape)
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def test_ppm(tmp_path):
    img = smooth_image()
    Image.fromarray(img).save(tmp_path / "a.ppm")
    assert np.array_equal(api.read_image(tmp_path / "a.ppm"), img)


@pytest.mark.parametrize("mode", ["RGB", "RGBA", "L", "LA", "P", "I;16"])
def test_png_colour_types_from_pillow(tmp_path, mode):
    img = smooth_image()
    if mode == "RGB":
        im, expect = Image.fromarray(img), img
    elif mode == "RGBA":
        a = RNG.integers(0, 256, size=img.shape[:2], dtype=np.uint8)
        im, expect = Image.fromarray(np.dstack([img, a]), "RGBA"), img
    elif mode == "L":
        g = img[..., 0]
        im, expect = Image.fromarray(g, "L"), np.repeat(g[..., None], 3, axis=2)
    elif mode == "LA":
        g = img[..., 1]
        im, expect = Image.fromarray(np.dstack([g, g // 2]), "LA"), np.repeat(g[..., None], 3, axis=2)
    elif mode == "P":
        im = Image.fromarray(img).quantize(colors=64)
        expect = np.asarray(im.convert("RGB"))
    else:  # 16-bit grey: the high byte survives
        g16 = (img[..., 2].astype(np.uint16) << 8) | img[..., 0]
        im = Image.fromarray(g16.astype(np.uint16))
        expect = np.repeat(img[..., 2:3], 3, axis=2)
    im.save(tmp_path / "a.png")
    got = api.read_image(tmp_path / "a.png")
    assert got.shape == expect.shape and np.array_equal(got, expect)


def png_bytes(rows_filtered, w, h, ctype=2, depth=8):
    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    ihdr = struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 0)
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr) + chunk(b"IDAT", zlib.compress(rows_filtered, 9))
            + chunk(b"IEND", b""))


def png_filter(img, ft):
    """Forward PNG filter ft (ISO/IEC 15948 section 9) of an RGB8 image, every row."""
    h, w, _ = img.shape
    bpp = 3
    out = bytearray()
    prev = np.zeros(w * bpp, np.int32)
    for y in range(h):
        cur = img[y].reshape(-1).astype(np.int32)
        a = np.concatenate([np.zeros(bpp, np.int32), cur[:-bpp]])
        b = prev
        c = np.concatenate([np.zeros(bpp, np.int32), prev[:-bpp]])
        if ft == 0:
            pred = np.zeros_like(cur)
        elif ft == 1:
            pred = a
        elif ft == 2:
            pred = b
        elif ft == 3:
            pred = (a + b) // 2
        else:
            p = a + b - c
            pa, pb, pc = np.abs(p - a), np.abs(p - b), np.abs(p - c)
            pred = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c))
        out.append(ft)
        out += ((cur - pred) % 256).astype(np.uint8).tobytes()
        prev = cur
    return bytes(out)


@pytest.mark.parametrize("ft", [0, 1, 2, 3, 4])
def test_png_each_filter_type(tmp_path, ft):
    img = smooth_image(37, 29)
    data = png_bytes(png_filter(img, ft), img.shape[1], img.shape[0])
    (tmp_path / "f.png").write_bytes(data)
    with Image.open(tmp_path / "f.png") as im:  # the file itself is right
        assert np.array_equal(np.asarray(im.convert("RGB")), img)
    assert np.array_equal(api.read_image(tmp_path / "f.png"), img)


@pytest.mark.parametrize("subsampling,quality,restart,maxdiff", [
    (0, 95, 0, 3), (2, 85, 0, 5), (1, 90, 0, 5), (2, 90, 3, 5), (0, 75, 1, 3)])
def test_jpeg_matches_an_independent_decoder(tmp_path, subsampling, quality, restart, maxdiff):
    # 4:4:4 (0), 4:2:2 (1), 4:2:0 (2); odd sizes exercise partial MCUs
    img = smooth_image(67, 101)
    kw = dict(quality=quality, subsampling=subsampling)
    if restart:
        kw["restart_marker_blocks"] = restart
    Image.fromarray(img).save(tmp_path / "a.jpg", **kw)
    with Image.open(tmp_path / "a.jpg") as im:
        pil = np.asarray(im.convert("RGB")).astype(int)
    got = api.read_image(tmp_path / "a.jpg").astype(int)
    assert got.shape == pil.shape
    d = np.abs(got - pil)
    # IDCT and upsampling rounding differ between decoders by a few levels
    assert d.max() <= maxdiff and d.mean() < 0.3, (int(d.max()), float(d.mean()))
    # and both are close to the original
    mse = ((got - img.astype(int)) ** 2).mean()
    assert 10 * np.log10(255 ** 2 / mse) > 28


def test_jpeg_greyscale(tmp_path):
    g = smooth_image(40, 50)[..., 1]
    Image.fromarray(g, "L").save(tmp_path / "g.jpg", quality=90)
    with Image.open(tmp_path / "g.jpg") as im:
        pil = np.asarray(im).astype(int)
    got = api.read_image(tmp_path / "g.jpg").astype(int)
    assert np.array_equal(got[..., 0], got[..., 1]) and np.array_equal(got[..., 0], got[..., 2])
    assert np.abs(got[..., 0] - pil).max() <= 2


def test_progressive_jpeg_and_missing_files_are_refused(tmp_path):
    Image.fromarray(smooth_image()).save(tmp_path / "p.jpg", quality=90, progressive=True)
    with pytest.raises(PrrError, match="progressive"):
        api.read_image(tmp_path / "p.jpg")
    with pytest.raises(PrrError, match="cannot read"):
        api.read_image(tmp_path / "nothing.png")
    (tmp_path / "x.txt").write_text("hello")
    with pytest.raises(PrrError, match="not a PPM"):
        api.read_image(tmp_path / "x.txt")


@pytest.mark.parametrize("fmt", ["jpg", "png"])
def test_truncated_and_corrupted_files_fail_cleanly(tmp_path, fmt):
    """Damaged files give an error status (or, for a cut inside padding, an
    image): the decoder never reads past the file."""
    Image.fromarray(smooth_image(40, 50)).save(tmp_path / f"a.{fmt}", **({"quality": 90} if fmt == "jpg" else {}))
    data = (tmp_path / f"a.{fmt}").read_bytes()
    rng = np.random.default_rng(1)
    cases = [data[:n] for n in (3, 10, 40, 100, 200, len(data) // 2, len(data) - 3)]
    for _ in range(20):  # random byte flips after the signature
        b = bytearray(data)
        for i in rng.integers(8, len(b), size=4):
            b[i] = int(rng.integers(0, 256))
        cases.append(bytes(b))
    for k, blob in enumerate(cases):
        p = tmp_path / f"bad{k}.{fmt}"
        p.write_bytes(blob)
        try:
            img = api.read_image(p)
            assert img.ndim == 3 and img.shape[2] == 3
        except PrrError:
            pass
