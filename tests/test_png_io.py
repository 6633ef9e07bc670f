"""PNG input and result images (ske_png_*): decoded against arrays written
by an independent codec (Pillow), for every colour type / bit depth; the
grey level is R+G+B of the 8-bit RGBA decode (src/helper.h:167-186).  Host
code only: runs without a GPU."""
import ctypes
import os
import tempfile

import numpy as np
import pytest

from skewengine_amd import _lib

Image = pytest.importorskip("PIL.Image")


def _read(path):
    L = _lib.lib()
    w, h = ctypes.c_int(0), ctypes.c_int(0)
    _lib.check(L.ske_png_read(path.encode(), ctypes.byref(w), ctypes.byref(h), None))
    out = np.zeros((h.value, w.value), np.float32)
    _lib.check(L.ske_png_read(path.encode(), ctypes.byref(w), ctypes.byref(h), out.ctypes.data))
    return out


@pytest.fixture
def tmp():
    with tempfile.TemporaryDirectory() as d:
        yield d


def test_rgb_rgba_grey_and_alpha(tmp):
    rng = np.random.default_rng(0)
    rgb = rng.integers(0, 256, (37, 53, 3), dtype=np.uint8)
    a = rng.integers(0, 256, (37, 53), dtype=np.uint8)
    want = rgb.astype(np.int32).sum(axis=2)
    cases = {
        "rgb.png": Image.fromarray(rgb, "RGB"),
        "rgba.png": Image.fromarray(np.dstack([rgb, a]), "RGBA"),
    }
    for name, im in cases.items():
        p = os.path.join(tmp, name)
        im.save(p)
        assert np.array_equal(_read(p), want), name
    g = rgb[..., 0]
    for mode, arr in (("L", g), ("LA", np.dstack([g, a]))):
        p = os.path.join(tmp, mode + ".png")
        Image.fromarray(arr, mode).save(p)
        assert np.array_equal(_read(p), 3 * g.astype(np.int32)), mode


def test_sixteen_bit_keeps_high_byte(tmp):
    rng = np.random.default_rng(1)
    g16 = rng.integers(0, 65536, (20, 31), dtype=np.uint16)
    p = os.path.join(tmp, "g16.png")
    im = Image.fromarray(g16)
    assert im.mode.startswith("I;16")
    im.save(p)
    assert np.array_equal(_read(p), 3 * (g16 >> 8).astype(np.int32))


def test_one_bit_and_palette(tmp):
    rng = np.random.default_rng(2)
    bits = rng.integers(0, 2, (19, 45), dtype=np.uint8) * 255
    p = os.path.join(tmp, "b1.png")
    Image.fromarray(bits, "L").convert("1").save(p)
    assert np.array_equal(_read(p), 3 * bits.astype(np.int32))
    pal = rng.integers(0, 256, (16, 3), dtype=np.uint8)
    idx = rng.integers(0, 16, (23, 29), dtype=np.uint8)
    im = Image.fromarray(idx, "P")
    im.putpalette(pal.reshape(-1).tolist())
    for bits_ in (4, 8):
        p = os.path.join(tmp, "p%d.png" % bits_)
        im.save(p, bits=bits_)
        assert np.array_equal(_read(p), pal[idx].astype(np.int32).sum(axis=2)), bits_


def test_writer_round_trips_through_pillow(tmp):
    L = _lib.lib()
    rng = np.random.default_rng(3)
    g = rng.integers(0, 256, (13, 70), dtype=np.uint8)
    rgb = rng.integers(0, 256, (13, 70, 3), dtype=np.uint8)
    for arr, ch in ((g, 1), (rgb, 3)):
        p = os.path.join(tmp, "w%d.png" % ch)
        a = np.ascontiguousarray(arr)
        _lib.check(L.ske_png_write(p.encode(), 70, 13, ch, a.ctypes.data))
        back = np.asarray(Image.open(p))
        assert np.array_equal(back, arr)


def test_result_images_follow_show_results(tmp):
    L = _lib.lib()
    rng = np.random.default_rng(4)
    pixels = rng.integers(0, 766, (24, 24)).astype(np.float32)
    out = rng.standard_normal((24, 24)).astype(np.float32) * 1000
    _lib.check(L.ske_result_pngs(tmp.encode(), _lib.KERNEL_RADON, 24, 24, pixels.ctypes.data, out.ctypes.data))
    s1 = np.asarray(Image.open(os.path.join(tmp, "salida1.png")))
    assert np.array_equal(s1, (pixels.astype(np.int32) // 3).astype(np.uint8))
    s2 = np.asarray(Image.open(os.path.join(tmp, "salida2.png")))
    mn, mx = np.float32(out.min()), np.float32(out.max())
    esc = np.float32(255) / (mx - mn)
    v = (out - mn).astype(np.int32).astype(np.float32) * esc
    want = v.astype(np.int32).astype(np.uint8)
    assert s2.shape == (24, 24, 3) and np.array_equal(s2[..., 0], want) and np.array_equal(s2[..., 2], want)
    with pytest.raises(_lib.SkeError):
        _lib.check(L.ske_result_pngs(tmp.encode(), _lib.KERNEL_SDEM, 24, 24, pixels.ctypes.data, out.ctypes.data))


def test_rejects_corrupt_and_missing_files(tmp):
    p = os.path.join(tmp, "bad.png")
    with open(p, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n" + b"\x00" * 30)
    with pytest.raises(_lib.SkeError):
        _read(p)
    with pytest.raises(_lib.SkeError):
        _read(os.path.join(tmp, "none.png"))
    good = os.path.join(tmp, "g.png")
    Image.fromarray(np.zeros((8, 8), np.uint8), "L").save(good)
    data = bytearray(open(good, "rb").read())
    data[-20] ^= 0xFF  # flip a byte inside the image data: CRC mismatch
    with open(p, "wb") as f:
        f.write(bytes(data))
    with pytest.raises(_lib.SkeError):
        _read(p)
