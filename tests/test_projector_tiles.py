"""CPU check of the bounds the tiled standard-geometry projector relies on
(kernels_fbp.hip: tile_krange, sample_range, k_fbp_project_tiles): for every
ray sample whose bilinear base pixel lies in a tile, the tile's bin range
holds its bin, the bin's sample range for that tile holds its index, and the
bin range fits the kw partial sums a (slice, angle, tile) owns; the tiles
cover the base pixels [-1, n)^2 exactly once.  A sample outside [-1, n)^2
reads only zero taps, so these bounds make the tiled sum equal the direct one
term for term.

Sample coordinates are recomputed in float32 with the kernel's operation
order (ray_coord without contraction, then one fma per coordinate: the
product of two floats is exact in double, then one rounding to float), for
small and odd image sides, detector counts from 2 to 2n+3, and angles on the
axes, the diagonals and at random."""
import math

import numpy as np
import pytest

f32 = np.float32
KPT_MAX = 129  # kPtMax in skewengine_amd/csrc/fbp.h


def proj_tiling(n):
    """proj_tiling() of fbp.h."""
    ntx = (n + 1 + KPT_MAX - 1) // KPT_MAX
    ts = (n + 1 + ntx - 1) // ntx
    kw = (ts * 1415 + 999) // 1000 + 8
    return ntx, ts, kw


def _tile_krange(n, ndet, det_c, c, sn, x0, x1, y0, y1):
    m = f32(0.5) * f32(n) - f32(0.5)
    ua, ub, va, vb = f32(x0) - m, f32(x1) - m, f32(y0) - m, f32(y1) - m
    p = [ua * c - va * sn, ub * c - va * sn, ua * c - vb * sn, ub * c - vb * sn]
    klo = max(int(math.floor(f32(min(p)) + det_c)) - 2, 0)
    khi = min(int(math.ceil(f32(max(p)) + det_c)) + 2, ndet - 1)
    return klo, khi


def _sample_range(n, nt, c, sn, P, x0, x1, y0, y1):
    """Vectorised over the bin offsets P (float32 array)."""
    m = f32(0.5) * f32(n) - f32(0.5)
    X0, X1 = f32(x0) - m - f32(1), f32(x1) - m + f32(1)
    Y0, Y1 = f32(y0) - m - f32(1), f32(y1) - m + f32(1)
    qlo = np.full(P.shape, f32(-1e30), np.float32)
    qhi = np.full(P.shape, f32(1e30), np.float32)
    pc, ps = P * c, P * sn
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        rc, rs = f32(1) / c, f32(1) / sn  # reciprocals rounded to float
        if abs(sn) > f32(1e-6):
            q1, q2 = (X0 - pc) * rs, (X1 - pc) * rs
            qlo, qhi = np.maximum(qlo, np.minimum(q1, q2)), np.minimum(qhi, np.maximum(q1, q2))
        else:
            bad = (pc < X0) | (pc > X1)
            qhi = np.where(bad, qlo - f32(1), qhi)
        if abs(c) > f32(1e-6):
            q1, q2 = (Y0 + ps) * rc, (Y1 + ps) * rc
            qlo, qhi = np.maximum(qlo, np.minimum(q1, q2)), np.minimum(qhi, np.maximum(q1, q2))
        else:
            bad = (-ps < Y0) | (-ps > Y1)
            qhi = np.where(bad, qlo - f32(1), qhi)
    half = f32(0.5) * f32(nt - 1)
    ok = qhi >= qlo
    qhi, qlo = np.where(ok, qhi, f32(0)), np.where(ok, qlo, f32(0))
    ilo = np.maximum(np.floor(half - qhi).astype(np.int64) - 2, 0)
    ihi = np.minimum(np.ceil(half - qlo).astype(np.int64) + 2, nt - 1)
    return np.where(ok, ilo, 0), np.where(ok, ihi, -1)


def _check_angle(n, ndet, th):
    delta = f32(2.0 / n)
    nt = ndet
    det_c = f32(0.5) * f32(ndet - 1)
    c, sn = f32(math.cos(th)), f32(math.sin(th))
    half = f32(0.5) * f32(nt - 1)
    inv = f32(1.0) / delta
    k = np.arange(ndet)
    P = k.astype(np.float32) - det_c
    s = P * delta
    u0 = ((s * c + f32(1.0)) * inv - f32(0.5)) + half * sn
    v0 = ((f32(1.0) - s * sn) * inv - f32(0.5)) + half * c
    i = np.arange(nt, dtype=np.float32)
    u = (-i[None, :].astype(np.float64) * np.float64(sn) + u0[:, None].astype(np.float64)).astype(np.float32)
    v = (-i[None, :].astype(np.float64) * np.float64(c) + v0[:, None].astype(np.float64)).astype(np.float32)
    bu, bv = np.floor(u).astype(np.int64), np.floor(v).astype(np.int64)
    ntx, ts, kw = proj_tiling(n)
    covered = np.zeros((ndet, nt), np.int64)
    for ty in range(ntx):
        for tx in range(ntx):
            x0, y0 = -1 + tx * ts, -1 + ty * ts
            x1, y1 = min(x0 + ts, n), min(y0 + ts, n)
            inside = (bu >= x0) & (bu < x1) & (bv >= y0) & (bv < y1)
            covered += inside
            klo, khi = _tile_krange(n, ndet, det_c, c, sn, x0, x1, y0, y1)
            assert khi - klo + 1 <= kw, (n, ndet, th, tx, ty, klo, khi, kw)
            kk, ii = np.nonzero(inside)
            if kk.size == 0:
                continue
            assert kk.min() >= klo and kk.max() <= khi, (n, ndet, th, tx, ty)
            ilo, ihi = _sample_range(n, nt, c, sn, P, x0, x1, y0, y1)
            assert np.all(ii >= ilo[kk]) and np.all(ii <= ihi[kk]), (n, ndet, th, tx, ty)
    dom = (bu >= -1) & (bu < n) & (bv >= -1) & (bv < n)
    assert np.array_equal(covered, dom.astype(np.int64))


@pytest.mark.parametrize("n", [2, 3, 8, 31, 64, 100, 129, 130, 300])
def test_projector_tile_bounds(n):
    rng = np.random.default_rng(n)
    angles = [a * math.pi / 4 for a in range(8)] + [math.pi * a / 24 for a in range(24)]
    angles += list(rng.uniform(-10, 10, 8))
    for ndet in sorted({2, n, int(1.4143 * n) + 1, 2 * n + 3}):
        if ndet < 2:
            continue
        for th in angles:
            _check_angle(n, ndet, th)


def test_projector_tiling_covers_the_base_pixels():
    for n in list(range(1, 600)) + [1023, 1024, 2048]:
        ntx, ts, kw = proj_tiling(n)
        assert 1 <= ts <= KPT_MAX and ntx * ts >= n + 1 and (ntx - 1) * ts < n + 1

