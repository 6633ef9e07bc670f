"""CPU check of the bound the back-projection adjoint relies on
(kernels_learned.hip: k_bp_adjoint_gather): for every pixel p and angle, the
two bins its back-projection weight reaches, k = floor(kf_p) and
floor(kf_p) + 1, each test p on their walk -- p's index along its row (or
column, when |cos| < |sin|) lies in the run of candidates the bin's thread
computes for that line.  kf and the run bounds are recomputed in float32 in
the kernel's operation order (no contraction)."""
import math

import numpy as np
import pytest

f32 = np.float32


def _check(n, ndet, th):
    delta = f32(2.0 / n)
    inv = f32(1) / delta
    det_c = f32(0.5) * f32(ndet - 1)
    c, sn = f32(math.cos(th)), f32(math.sin(th))
    rows = abs(c) >= abs(sn)
    slope = c if rows else -sn
    rsl = f32(1) / slope
    cnt = int(f32(2) * abs(rsl)) + 3
    idx = np.arange(n)
    x = (idx.astype(f32) + f32(0.5)) * delta - f32(1)   # x of column j
    y = f32(1) - (idx.astype(f32) + f32(0.5)) * delta   # y of row i
    kf = (x[None, :] * c + y[:, None] * sn) * inv + det_c  # [i, j], as the back-projection computes it
    k0 = np.floor(kf).astype(np.int64)
    # per line u: kf at index 0, approximated as the kernel does
    fu = (idx.astype(f32) + f32(0.5)) * delta
    if rows:
        kf0 = ((f32(0.5) * delta - f32(1)) * c + (f32(1) - fu) * sn) * inv + det_c
    else:
        kf0 = ((fu - f32(1)) * c + (f32(1) - f32(0.5) * delta) * sn) * inv + det_c
    for dk in (0, 1):
        k = k0 + dk  # the bin, and the pixel's position on its line
        ok = (k >= 0) & (k < ndet)
        u = np.broadcast_to(idx[:, None] if rows else idx[None, :], k.shape)  # line index
        t = np.broadcast_to(idx[None, :] if rows else idx[:, None], k.shape)  # position along the line
        kk = k.astype(f32)
        ta = (kk - f32(1) - kf0[u]) * rsl
        tb = (kk + f32(1) - kf0[u]) * rsl
        t0 = np.floor(np.minimum(ta, tb)).astype(np.int64) - 1
        inside = (t >= t0) & (t < t0 + cnt)
        assert np.all(inside[ok]), (n, ndet, th, dk, int(np.sum(~inside & ok)))


@pytest.mark.parametrize("n", [8, 31, 64, 96, 128, 160, 512])
def test_adjoint_gather_tests_every_contributing_pixel(n):
    rng = np.random.default_rng(n)
    angles = [math.pi * a / 48 for a in range(96)] + [math.pi / 4, 3 * math.pi / 4] + list(rng.uniform(-10, 10, 8))
    for ndet in sorted({n, int(1.4143 * n) + 1, 2 * n + 3}):
        for th in angles:
            _check(n, ndet, th)
