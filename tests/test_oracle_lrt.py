"""Pins of the oracle's local Radon transform (pointRadon,
src/kernelRADN.h:114-182): a numpy restatement built on numpy's FFT (the
library routine the reference's radix-2 transform must agree with), the
exact scaling invariant LRT(2 img) = 4 LRT(img), and the zero image."""
import numpy as np
import pytest

import oracle
from skewengine_amd import synth


def _numpy_point_radon(sec, i, j):
    p = sec["p"]
    H, w = p.skewHeight, p.skewWidth
    S = sec["S"][:H * w].reshape(H, w).astype(np.float64)
    first, last = sec["first"], sec["last"]
    sigma = (64 / 5.0) / p.scale
    x = np.zeros(64)
    for kk in range(64):
        k = i - 32 + kk
        acc = 0.0
        if 0 <= k < H:
            lo, hi = max(j - 32, int(first[k])), min(j + 32, int(last[k]) - 1)
            if hi >= lo:
                acc = S[k, lo:hi + 1].sum()
        x[kk] = acc * np.exp(-((kk - 32) ** 2) / (2 * sigma * sigma))
    X = np.fft.fft(x) / 8.0
    f = np.arange(64)
    y = np.where(f > 25.6, 0.0, np.sin(f / 24.6 * np.pi / 2) * X.real)
    Y = np.fft.fft(y) / 8.0
    spec = np.abs(Y) ** 2
    for m in range(5, 32):
        if spec[m] > spec[m - 1]:
            return spec[m - 1], spec
    return 0.0, spec


@pytest.mark.parametrize("angle", [0, 17, 45, 46, 90, 133, 170])
def test_point_radon_matches_numpy_fft(angle):
    img = synth.ellipse_phantom(48, seed=angle + 1)
    sec = oracle.sector(oracle.RADON_LOCAL, img, angle)
    p = sec["p"]
    H, w = p.skewHeight, p.skewWidth
    SO = sec["SO"][:H * w].reshape(H, w)
    rng = np.random.default_rng(angle)
    checked = 0
    for i in rng.choice(H, size=min(H, 24), replace=False):
        f, l = int(sec["first"][i]), int(sec["last"][i])
        for j in range(f, l, 5):
            v, spec = oracle.point_radon(sec, int(i), j)
            nv, nspec = _numpy_point_radon(sec, int(i), j)
            scale = max(nspec.max(), 1e-300)
            assert np.allclose(spec, nspec, rtol=0, atol=1e-12 * scale)
            # the peak rule picks the same bin unless two bins tie to rounding
            if abs(v - nv) > 1e-12 * scale:
                d = np.abs(np.diff(nspec[4:32]))
                assert d.min() <= 1e-10 * scale, (i, j, v, nv)
            assert SO[i, j] == np.float32(v)
            checked += 1
    assert checked > 50


def test_scaling_is_exact_and_zero_maps_to_zero():
    img = synth.ellipse_phantom(40, seed=9)
    a, _ = oracle.execute(oracle.RADON_LOCAL, img, nthreads=1, s0=0, s1=180)
    b, _ = oracle.execute(oracle.RADON_LOCAL, 2 * img, nthreads=1, s0=0, s1=180)
    assert np.array_equal(b, 4 * a)
    z, _ = oracle.execute(oracle.RADON_LOCAL, np.zeros((40, 40), np.float32), nthreads=1, s0=0, s1=180)
    assert not np.any(z)
    assert np.any(a)


@pytest.mark.parametrize("angle", [3, 45, 100, 160])
def test_sample_vectors_match_numpy(angle):
    """The four vectors recorded for a sample point (kernelRADN.h:164-170),
    restated with numpy's FFT; the axis vector uses the extents of row k of
    the window for column j + k - i of row i, as the reference reads them."""
    n = 48
    img = synth.ellipse_phantom(n, seed=angle)
    sec = oracle.sector(oracle.RADON_LOCAL, img, angle)
    p = sec["p"]
    H, w = p.skewHeight, p.skewWidth
    S = sec["S"][:H * w].reshape(H, w).astype(np.float64)
    first, last = sec["first"], sec["last"]
    sigma = (64 / 5.0) / p.scale
    for (x, y) in [(20, 24), (5, 40), (33, 7)]:
        i, j = oracle.skewed_cell(p, sec["target"], x, y, n, n)
        # the cell's source is (x, y): its skewed value carries that pixel
        t = int(sec["target"][j])
        assert 0 <= i - t < p.dim_o
        got = oracle.point_radon_samples(sec, i, j)
        g = np.exp(-((np.arange(64) - 32) ** 2) / (2 * sigma * sigma))
        win = np.zeros(64)
        axis = np.zeros(64)
        for kk in range(64):
            k = i - 32 + kk
            if 0 <= k < H:
                lo, hi = max(j - 32, int(first[k])), min(j + 32, int(last[k]) - 1)
                win[kk] = S[k, lo:hi + 1].sum() if hi >= lo else 0.0
                c = j + kk - 32
                if first[k] <= c < last[k]:
                    axis[kk] = g[kk] * S[i, c]
        x1 = win * g
        X = np.fft.fft(x1) / 8.0
        f = np.arange(64)
        y2 = np.where(f > 25.6, 0.0, np.sin(f / 24.6 * np.pi / 2) * X.real)
        Y = np.fft.fft(y2) / 8.0
        ref = np.stack([x1, axis, y2, Y.real])
        scale = np.maximum(np.abs(ref).max(axis=1, keepdims=True), 1e-30)
        assert np.all(np.abs(got - ref) <= 1e-6 * scale), (x, y)


def test_skewed_cell_inverts_the_reference_coordinate_map():
    # getXY (src/kernelRADN.h:28-41) applied to skewed_cell(x, y) gives (x, y)
    n = 37
    for angle in (0, 20, 45, 60, 90, 91, 135, 179):
        p, target, weight, first, last = oracle.compute_params(n, n, angle)
        tr = p.sectorType in (1, 2)
        rx = p.sectorType in (1, 3)
        ry = p.sectorType in (1, 2)
        for (x, y) in [(0, 0), (36, 0), (0, 36), (10, 29), (36, 36)]:
            i, j = oracle.skewed_cell(p, target, x, y, n, n)
            yy = j if tr else i - target[j]
            xx = i - target[j] if tr else j
            xx = n - 1 - xx if rx else xx
            yy = n - 1 - yy if ry else yy
            assert (xx, yy) == (x, y), angle
