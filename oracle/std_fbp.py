"""FP64 CPU oracle of the standard parallel-beam projector / FBP
(test infrastructure only).

There is no reference code for this geometry (the reference's Radon works in
skew space, src/kernelRADN.h:212-259); this restatement of the GPU
definitions in float64 numpy is pinned by the closed-form projections of
ellipses and by the Ram-Lak kernel's vanishing sum (tests/test_oracle_fbp.py).
"""
import numpy as np


def angle_table(n_angles, angles=None):
    """cos/sin as the device table holds them (float32-rounded), as float64."""
    th = np.pi * np.arange(n_angles) / n_angles if angles is None else np.asarray(angles, np.float64)
    return np.cos(th).astype(np.float32).astype(np.float64), np.sin(th).astype(np.float32).astype(np.float64)


def _bilinear(img, u, v):
    n = img.shape[0]
    j0 = np.floor(u).astype(np.int64)
    i0 = np.floor(v).astype(np.int64)
    wu = u - j0
    wv = v - i0
    out = np.zeros(u.shape)

    def tap(i, j):
        ok = (i >= 0) & (i < n) & (j >= 0) & (j < n)
        val = np.zeros(u.shape)
        val[ok] = img[i[ok], j[ok]]
        return val

    a, b, c, d = tap(i0, j0), tap(i0, j0 + 1), tap(i0 + 1, j0), tap(i0 + 1, j0 + 1)
    top = a + wu * (b - a)
    bot = c + wu * (d - c)
    out = top + wv * (bot - top)
    return out


def project(img, n_angles, n_det, angles=None, angle_idx=None):
    """p(a,k) = delta * sum_t f(s_k cos - t sin, s_k sin + t cos), t on the
    bin pitch, bilinear f."""
    img = np.asarray(img, np.float64)
    n = img.shape[0]
    delta = 2.0 / n
    det_c = 0.5 * (n_det - 1)
    s = (np.arange(n_det) - det_c) * delta
    t = (np.arange(n_det) - det_c) * delta
    cs, sn = angle_table(n_angles, angles)
    idx = range(n_angles) if angle_idx is None else angle_idx
    out = np.zeros((len(list(idx)), n_det))
    for r, a in enumerate(idx):
        x = s[:, None] * cs[a] - t[None, :] * sn[a]
        y = s[:, None] * sn[a] + t[None, :] * cs[a]
        u = (x + 1.0) / delta - 0.5
        v = (1.0 - y) / delta - 0.5
        out[r] = delta * _bilinear(img, u, v).sum(axis=1)
    return out


def ramlak_taps(n_det, n):
    delta = 2.0 / n
    m = np.arange(-(n_det - 1), n_det)
    h = np.zeros(m.shape)
    h[m == 0] = 1.0 / (4 * delta * delta)
    odd = (m % 2) != 0
    h[odd] = -1.0 / (m[odd].astype(np.float64) ** 2 * np.pi ** 2 * delta * delta)
    return delta * h


def filter_rows(sino, taps):
    """q(a,k) = sum_k' p(a,k') h(k-k')."""
    sino = np.asarray(sino, np.float64)
    n_det = sino.shape[-1]
    full = np.array([np.convolve(row, taps) for row in sino.reshape(-1, n_det)])
    return full[:, n_det - 1:2 * n_det - 1].reshape(sino.shape)


def backproject(q, n, angles=None):
    """f(x,y) = (pi/nang) sum_a q(a, s), s = x cos + y sin, linear in s."""
    q = np.asarray(q, np.float64)
    n_angles, n_det = q.shape
    delta = 2.0 / n
    det_c = 0.5 * (n_det - 1)
    cs, sn = angle_table(n_angles, angles)
    x = (np.arange(n) + 0.5) * delta - 1.0
    y = 1.0 - (np.arange(n) + 0.5) * delta
    X, Y = x[None, :], y[:, None]
    out = np.zeros((n, n))
    for a in range(n_angles):
        kf = (X * cs[a] + Y * sn[a]) / delta + det_c
        k0 = np.floor(kf).astype(np.int64)
        w = kf - k0
        row = q[a]
        v0 = np.where((k0 >= 0) & (k0 < n_det), row[np.clip(k0, 0, n_det - 1)], 0.0)
        v1 = np.where((k0 + 1 >= 0) & (k0 + 1 < n_det), row[np.clip(k0 + 1, 0, n_det - 1)], 0.0)
        out += v0 + w * (v1 - v0)
    return out * (np.pi / n_angles)


def fbp(sino, n, taps=None, angles=None):
    n_det = sino.shape[-1]
    taps = ramlak_taps(n_det, n) if taps is None else taps
    return backproject(filter_rows(sino, taps), n, angles)


def ellipse_sinogram(params, n_angles, n_det, n, angles=None):
    """Closed-form line integrals of additive ellipses (x0, y0, a, b, phi, rho):
    p(s,theta) = 2 rho a b sqrt(m^2 - s'^2) / m^2 for |s'| < m,
    m^2 = a^2 cos^2(theta-phi) + b^2 sin^2(theta-phi), s' = s - (x0 cos + y0 sin)."""
    delta = 2.0 / n
    s = (np.arange(n_det) - 0.5 * (n_det - 1)) * delta
    th = np.pi * np.arange(n_angles) / n_angles if angles is None else np.asarray(angles, np.float64)
    out = np.zeros((n_angles, n_det))
    for (x0, y0, a, b, phi, rho) in params:
        m2 = (a * np.cos(th - phi)) ** 2 + (b * np.sin(th - phi)) ** 2
        sp = s[None, :] - (x0 * np.cos(th) + y0 * np.sin(th))[:, None]
        inside = sp ** 2 < m2[:, None]
        val = np.zeros_like(sp)
        val[inside] = (2 * rho * a * b * np.sqrt((m2[:, None] - sp ** 2)[inside]) / np.broadcast_to(m2[:, None], sp.shape)[inside])
        out += val
    return out
