"""Pins of the FP64 total-viewshed oracle (the reference's kernelV3,
src/kernelSDEM.h:168-222, restated in double), the yardstick the GPU's
viewshed totals are reported against: an independent numpy formulation on
small DEMs (maps from the pinned planner oracle, per-target visibility by
brute force, explicit deskew and reduce), flat terrain, and closeness to
the FP32 farm."""
import numpy as np
import pytest

import oracle
from skewengine_amd import synth


def observer_f64(row, first, last, j, povh):
    """Visible area of one observer in double from explicit per-target
    visibility (a target is visible when its elevation angle exceeds all
    nearer ones and -2000; each visible run [s, e] closed before the row end
    adds e'^2 - s^2, e' the first hidden distance, the observer's own run
    starting at 0)."""
    h = row[j] + povh
    cv = 0.0
    for direction in (+1, -1):
        ks = np.arange(j + direction, last if direction > 0 else first - 1, direction)
        if len(ks) == 0:
            continue
        d = np.abs(ks - j).astype(np.float64)
        ang = (row[ks] - h) / d
        prev = np.maximum.accumulate(np.concatenate([[-2000.0], ang[:-1]]))
        vis = ang > prev
        start = 0.0
        visible = True
        for k in range(len(ks)):
            if vis[k] and not visible:
                start = d[k]
            if not vis[k] and visible:
                cv += d[k] ** 2 - start ** 2
            visible = bool(vis[k])
    return cv


def farm_f64_numpy(img, s0, s1, povh):
    n = img.shape[0]
    src4 = [a.reshape(n, n).astype(np.float64) for a in oracle.prepare(img)]
    accs = [np.zeros((n, n)) for _ in range(4)]
    for a in range(s0, s1):
        p, target, weight, first, last = oracle.compute_params(n, n, a)
        di, H = p.dim_i, p.skewHeight
        t = target[:di]
        w = p.skewness * np.arange(di) - t  # double weights
        S = np.zeros((H + 1, di))
        src = src4[p.sectorType]
        for i in range(p.dim_o):
            for j in range(di):
                S[i + t[j], j] += (1 - w[j]) * src[i, j]
                S[i + t[j] + 1, j] += w[j] * src[i, j]
        SO = np.zeros((H + 1, di))
        for i in range(H):
            for j in range(first[i], last[i]):
                SO[i, j] = observer_f64(S[i], int(first[i]), int(last[i]), j, povh) * (p.scale * p.scale)
        acc = accs[p.sectorType]
        for i in range(p.dim_o):
            for j in range(di):
                acc[i, j] += (1 - w[j]) * SO[i + t[j], j] + w[j] * SO[i + t[j] + 1, j]
    out = np.zeros((n, n))
    for i in range(n):
        ci = n - 1 - i
        for j in range(n):
            cj = n - 1 - j
            out[i, j] = accs[0][i, j] + accs[1].flat[cj * n + ci] + accs[2].flat[j * n + ci] + accs[3].flat[i * n + cj]
    return out


@pytest.mark.parametrize("seed,s0,s1", [(1, 0, 180), (2, 40, 50), (3, 85, 140)])
def test_fp64_farm_matches_numpy_formulation(seed, s0, s1):
    img = synth.dem(12, 0.8, seed, 1500.0)
    got = oracle.execute_sdem_f64(img, nthreads=3, s0=s0, s1=s1, povh=0.15)
    ref = farm_f64_numpy(img, s0, s1, 0.15)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-9)


def test_fp64_sea_level_terrain_is_zero():
    # zero heights stay zero through the skew; from 0.15 above the ground every
    # angle -0.15/d rises with d, so every target is visible and nothing closes
    img = np.zeros((20, 20), np.float32)
    assert np.all(oracle.execute_sdem_f64(img, nthreads=2, povh=0.15) == 0.0)


def test_fp64_close_to_fp32_farm():
    img = synth.dem(64, 0.8, 5, 1500.0)
    f64 = oracle.execute_sdem_f64(img, nthreads=4, povh=0.15)
    f32, _ = oracle.execute(oracle.SDEM, img, nthreads=1, povh=0.15)
    # float rounding of the angles flips a few near-tie visibility decisions
    assert np.max(np.abs(f32 - f64)) <= 1e-3 * np.max(np.abs(f64))
    assert np.mean(np.abs(f32 - f64)) <= 1e-6 * np.max(np.abs(f64))
