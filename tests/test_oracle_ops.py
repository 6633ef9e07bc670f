"""Pins of the oracle's operators and transforms against what the mathematics
fixes: numpy's FFT for helper::fft, a brute-force line-of-sight for
kernelV3, the skew/deskew closed form for the identity operator, the
definition of the cepstrum chain, hand-built peak-rule vectors, and
driver-level invariants (thread count, identity ~ 180 x input)."""
import numpy as np
import pytest

import oracle
from skewengine_amd import synth


# ---------------------------------------------------------------- helper::fft
@pytest.mark.parametrize("n", [2, 8, 32, 64, 512, 1024])
def test_fft_matches_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(x)
    got = oracle.fft(x)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref)) * np.log2(n)
    got_n = oracle.fft(x, normalize=True)
    assert np.allclose(got_n, ref / np.sqrt(n), rtol=0, atol=1e-12 * np.max(np.abs(ref)))
    got_i = oracle.ifft(x)
    assert np.allclose(got_i, np.fft.ifft(x), rtol=0, atol=1e-12 * np.max(np.abs(x)))


def test_fft_sign_convention():
    # forward transform uses e^{-2 pi i k n / N}: a unit impulse at n=1 gives
    # X_k = e^{-2 pi i k / N}
    x = np.zeros(16, complex)
    x[1] = 1
    X = oracle.fft(x)
    assert np.allclose(X, np.exp(-2j * np.pi * np.arange(16) / 16), atol=1e-14)


# ------------------------------------------------------- getBlurParameter2
def test_blur_parameter_monotone_is_zero():
    v = np.linspace(10, 0, 16)
    assert oracle.blur_parameter2(v) == (0.0, 0)


def test_blur_parameter_skips_first_ascending_sample():
    # first ascent at i=7; v[7] is the global max but is skipped: the value is
    # max(v[6], v[8:]) (kernelBLUR.h:100 sets imax=i++).
    v = np.array([9, 8, 7, 6, 5, 4, 3, 100, 2, 5, 1, 0, 0, 0, 0, 0], float)
    val, imax = oracle.blur_parameter2(v)
    assert val == 5.0 and imax == 9
    v2 = v.copy()
    v2[9] = 1
    val, imax = oracle.blur_parameter2(v2)
    assert val == 3.0 and imax == 7  # v[6]; imax stays at the ascent index


def test_blur_parameter_ignores_ascents_before_5():
    v = np.array([0, 5, 9, 1, 8, 7, 6, 5, 4, 3, 2, 1, 0.5, 0.4, 0.3, 0.2])
    assert oracle.blur_parameter2(v)[0] == 0.0


# ---------------------------------------------------------- lineCepstrum
def cepstrum_by_definition(j, row, first, last):
    x = np.array([row[i] if first <= i < last else 0.0 for i in range(j - 32, j + 32)], float)
    X = np.fft.fft(x) / 8.0
    Lg = np.log(np.abs(X[:32]) ** 2)
    C = np.fft.fft(Lg) / np.sqrt(32)
    v = -np.minimum(0.0, C.real[:16])
    i = 5
    while i < 16 and not v[i] > v[i - 1]:
        i += 1
    if i == 16:
        V = 0.0
    else:
        V = max([v[i - 1]] + list(v[i + 1:]))
    return np.float32(V * 10000)


def test_line_cepstrum_matches_definition():
    rng = np.random.default_rng(5)
    img, length, angle = synth.blurred_texture(n=128, seed=3)
    row = img[40].astype(np.float32)
    n_ok = 0
    for j in list(range(0, 128, 7)) + [3, 120]:
        for first, last in [(0, 128), (10, 100)]:
            if not first <= j < last:
                continue
            got = oracle.line_cepstrum(j, row, first, last)
            ref = cepstrum_by_definition(j, row, first, last)
            assert got == pytest.approx(ref, rel=1e-9, abs=1e-6), (j, first, last)
            n_ok += 1
    assert n_ok > 10
    del rng


# ------------------------------------------------------------- kernelV3
def observer_brute_force(row, first, last, j, povh):
    """Visible area of one observer from explicit per-target visibility:
    target k is visible when its (float32) elevation angle exceeds every
    angle strictly between the observer and k (and -2000); each maximal
    visible run [s, e] adds (e+1)^2 - s^2 unless it reaches the row end,
    with the run touching the observer starting at distance 0."""
    row = np.asarray(row, np.float32)
    h = np.float32(row[j] + np.float32(povh))
    cv = np.float32(0)
    for direction in (+1, -1):
        ks = np.arange(j + direction, last if direction > 0 else first - 1, direction)
        if len(ks) == 0:
            continue
        d = np.abs(ks - j).astype(np.float32)
        ang = (row[ks] - h) / d
        prev = np.maximum.accumulate(np.concatenate([[np.float32(-2000)], ang[:-1]]))
        vis = ang > prev
        L = len(ks)
        dist = np.arange(1, L + 1)
        k = 0
        # the initial run (observer counts as visible, distance 0)
        while k < L and vis[k]:
            k += 1
        if k < L:
            cv += np.float32(dist[k] ** 2)
        while k < L:
            while k < L and not vis[k]:
                k += 1
            if k >= L:
                break
            s = dist[k]
            while k < L and vis[k]:
                k += 1
            if k < L:
                cv += np.float32(dist[k] ** 2 - s ** 2)
    return cv


def test_viewshed_flat_terrain_is_zero():
    row = np.full(50, 3.0, np.float32)
    for j in (0, 10, 49):
        cv, vr, vl = oracle.sdem_observer(row, 0, 50, j, 0.15)
        assert cv == 0.0 and np.all(vr == 1) and np.all(vl == 1)


def test_viewshed_hand_profile():
    row = np.array([0, 0, 0, 10, 0, 0], np.float32)
    cv, vr, vl = oracle.sdem_observer(row, 0, 6, 0, 0.0)
    # d=1 visible, d=2 level -> closes (4), d=3 opens, d=4 closes (16-9)
    assert list(vr) == [1, 0, 1, 0, 0]
    assert cv == 4 + 7


@pytest.mark.parametrize("seed", range(6))
def test_viewshed_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    L = 40
    row = (rng.random(L) * 5).astype(np.float32)
    if seed % 2:
        row = np.cumsum(rng.standard_normal(L)).astype(np.float32)
    first, last = 3, L - 2
    for j in range(first, last):
        cv, vr, vl = oracle.sdem_observer(row, first, last, j, 0.15)
        assert cv == observer_brute_force(row, first, last, j, 0.15), j


def test_viewshed_mirror_symmetry():
    rng = np.random.default_rng(9)
    row = (rng.random(60) * 20).astype(np.float32)
    for j in range(60):
        cv, vr, vl = oracle.sdem_observer(row, 0, 60, j, 0.15)
        cv2, vr2, vl2 = oracle.sdem_observer(row[::-1].copy(), 0, 60, 59 - j, 0.15)
        assert np.array_equal(vr, vl2) and np.array_equal(vl, vr2)


# --------------------------------------------------- skew / deskew / identity
def test_skew_deskew_closed_form():
    # identity operator: deskew(skew(x)) = ((1-w)^2+w^2) x_i + w(1-w)(x_{i-1}+x_{i+1})
    rng = np.random.default_rng(3)
    img = rng.random((60, 60)).astype(np.float32)
    for a in (7, 30, 45, 60, 100, 150):
        s = oracle.sector(oracle.IDENTITYIMG, img, a)
        p = s["p"]
        src = [*oracle.prepare(img)][p.sectorType].reshape(p.dim_o, p.dim_i).astype(float)
        O = np.zeros(p.dim_o * p.dim_i, np.float32)
        oracle.deskew(p, s["target"], s["weight"], s["SO"], O, 0)
        O = O.reshape(p.dim_o, p.dim_i)
        t = s["target"][:p.dim_i]
        w = s["weight"][:p.dim_i].astype(float)
        f, l = s["first"], s["last"]
        checked = 0
        for i in range(1, p.dim_o - 1):
            for j in range(p.dim_i):
                r = i + t[j]
                if not (f[r] <= j < l[r] and f[r + 1] <= j < l[r + 1]):
                    continue
                ref = ((1 - w[j]) ** 2 + w[j] ** 2) * src[i, j] + w[j] * (1 - w[j]) * (src[i - 1, j] + src[i + 1, j])
                assert abs(O[i, j] - ref) <= 1e-6 * (1 + abs(ref))
                checked += 1
        assert checked > 0.9 * (p.dim_o - 2) * p.dim_i


def test_identity_constant_image_gives_180():
    # main.cu:50-54: the identity run reproduces the input x 180 (interior).
    img = np.full((40, 40), 2.0, np.float32)
    out, _ = oracle.execute(oracle.IDENTITYIMG, img, nthreads=1)
    assert np.allclose(out[5:-5, 5:-5], 360.0, rtol=1e-5)


def test_execute_thread_count_invariance():
    img = synth.dem(64, 0.8, 1, 1500.0)
    o1, _ = oracle.execute(oracle.SDEM, img, nthreads=1)
    o8, _ = oracle.execute(oracle.SDEM, img, nthreads=6)
    assert np.max(np.abs(o1 - o8)) <= 1e-5 * np.max(np.abs(o1))


# ----------------------------------------------------------------- radon
def test_radon_operator_matches_numpy_filter():
    img = synth.ellipse_phantom(64, seed=2)
    for a in (0, 20, 45, 77, 120):
        s = oracle.sector(oracle.RADON, img, a)
        p = s["p"]
        H = p.skewHeight
        S = s["S"][:H * p.dim_i].reshape(H, p.dim_i).astype(float)
        f, l = s["first"][:H].astype(int), s["last"][:H].astype(int)
        P = np.array([S[r, f[r]:l[r]].sum() for r in range(H)])
        H2 = 1 << int(np.ceil(np.log2(H)))
        X = np.fft.fft(np.concatenate([P, np.zeros(H2 - H)]))
        k = np.arange(H2)
        X = np.where(k > H // 2, 0, k / H * X)
        q = np.fft.ifft(X).real[:H]
        SO = s["SO"][:H * p.dim_i].reshape(H, p.dim_i)
        for r in range(H):
            if f[r] < l[r]:
                assert np.allclose(SO[r, f[r]:l[r]], np.float32(q[r]), rtol=1e-6, atol=1e-6 * np.abs(q).max())
