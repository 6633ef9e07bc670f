"""CPU check of the invariant the standard-geometry projector relies on
(kernels_fbp.hip: k_fbp_pad, ray_sample): for every ray sample that
ray_range admits, the four bilinear taps lie within 3 pixels of the image, so
a zero border of kProjPad = 4 pixels replaces all bounds tests.

The sample coordinates are recomputed here in float32 with the kernel's
operation order (two explicit fma per sample, otherwise no contraction, IEEE
division), vectorised over bins and samples, for small and odd image sides, detector counts from 2 to 2n+3 and
angles on the axes, the diagonals and at random."""
import math

import numpy as np
import pytest

PAD = 4  # kProjPad in skewengine_amd/csrc/fbp.h

f32 = np.float32


def _footprint(n, ndet, c, sn):
    """Lowest and highest tap row/column index over all samples of one angle."""
    delta = f32(2.0 / n)
    nt = ndet
    det_c = f32(0.5) * f32(ndet - 1)
    tmax = f32(0.5) * f32(ndet - 1) * delta
    c, sn = f32(c), f32(sn)
    k = np.arange(ndet, dtype=np.float32)
    s = (k - det_c) * delta
    B = f32(1.0) + delta
    tlo = np.full(ndet, -tmax, np.float32)
    thi = np.full(ndet, tmax, np.float32)
    empty = np.zeros(ndet, bool)
    with np.errstate(divide="ignore", invalid="ignore"):
        if abs(sn) > f32(1e-7):
            t1 = (s * c - B) / sn
            t2 = (s * c + B) / sn
            tlo = np.maximum(tlo, np.minimum(t1, t2))
            thi = np.minimum(thi, np.maximum(t1, t2))
        else:
            empty |= np.abs(s * c) > B
        if abs(c) > f32(1e-7):
            t1 = (-B - s * sn) / c
            t2 = (B - s * sn) / c
            tlo = np.maximum(tlo, np.minimum(t1, t2))
            thi = np.minimum(thi, np.maximum(t1, t2))
        else:
            empty |= np.abs(s * sn) > B
    half = f32(0.5) * f32(nt - 1)
    ok = (thi >= tlo) & ~empty
    i0 = np.maximum(np.ceil(tlo / delta + half).astype(np.int64) - 1, 0)
    i1 = np.minimum(np.floor(thi / delta + half).astype(np.int64) + 1, nt - 1)
    lo, hi = 0, n - 1
    inv = f32(1.0) / delta
    for b in np.nonzero(ok & (i1 >= i0))[0]:
        i = np.arange(i0[b], i1[b] + 1).astype(np.float32)
        # ray_coord / ray_sample: u = fma(-i, sn, U0), v = fma(-i, c, V0); the
        # product of two floats is exact in double, then one rounding to float
        u0 = ((s[b] * c + f32(1.0)) * inv - f32(0.5)) + half * sn
        v0 = ((f32(1.0) - s[b] * sn) * inv - f32(0.5)) + half * c
        u = (-i.astype(np.float64) * np.float64(sn) + np.float64(u0)).astype(np.float32)
        v = (-i.astype(np.float64) * np.float64(c) + np.float64(v0)).astype(np.float32)
        ju, iv = np.floor(u).astype(np.int64), np.floor(v).astype(np.int64)
        lo = min(lo, int(ju.min()), int(iv.min()))
        hi = max(hi, int(ju.max()) + 1, int(iv.max()) + 1)
    return lo, hi


@pytest.mark.parametrize("n", [2, 3, 8, 31, 64, 100])
def test_projector_taps_inside_zero_border(n):
    rng = np.random.default_rng(n)
    angles = [a * math.pi / 4 for a in range(8)] + [math.pi * a / 60 for a in range(60)]
    angles += list(rng.uniform(-10, 10, 16))
    for ndet in sorted({2, n, int(1.4143 * n) + 1, 2 * n + 3}):
        if ndet < 2:
            continue
        for th in angles:
            lo, hi = _footprint(n, ndet, math.cos(th), math.sin(th))
            assert lo >= -PAD + 1 and hi <= n - 1 + PAD - 1, (n, ndet, th, lo, hi)
