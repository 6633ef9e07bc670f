"""Seeded synthetic inputs for the benchmark configurations.

All generators compute in float64 and return float32 grids, so the same
bytes feed the GPU path and the CPU oracle.
"""
import numpy as np

try:  # multi-threaded FFTs when scipy is present
    import scipy.fft as _fft

    def _irfft2(a, s):
        return _fft.irfft2(a, s=s, workers=-1)

    def _fft2(a):
        return _fft.fft2(a, workers=-1)

    def _ifft2(a):
        return _fft.ifft2(a, workers=-1)
except ImportError:  # pragma: no cover
    def _irfft2(a, s):
        return np.fft.irfft2(a, s=s)

    def _fft2(a):
        return np.fft.fft2(a)

    def _ifft2(a):
        return np.fft.ifft2(a)


def fbm(n, hurst, seed, lo=0.0, hi=1.0):
    """Fractional Brownian surface by spectral synthesis: amplitude
    |k|^-(H+1), DC removed, uniform random phase, real inverse FFT, min-max
    rescaled to [lo, hi]."""
    rng = np.random.default_rng(seed)
    ky = np.fft.fftfreq(n)[:, None]
    kx = np.fft.rfftfreq(n)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = 1.0
    amp = k ** (-(hurst + 1.0))
    amp[0, 0] = 0.0
    phase = rng.uniform(0.0, 2.0 * np.pi, size=amp.shape)
    z = _irfft2(amp * np.exp(1j * phase), s=(n, n))
    z -= z.min()
    z *= (hi - lo) / z.max()
    z += lo
    return z


def dem(n, hurst, seed, hmax_m, step_m=10.0):
    """fBm DEM with heights in [0, hmax_m] metres, stored in cell units
    (metres / step) as the reference stores DEM heights (src/kernelSDEM.h:47)."""
    return (fbm(n, hurst, seed, 0.0, hmax_m) / step_m).astype(np.float32)


def ellipse_phantom(n=512, n_ellipses=10, seed=0, return_params=False):
    """Additive random ellipses on [-1,1]^2 sampled at pixel centres: centres
    uniform in the unit disk, semi-axes U[0.05,0.4], rotation U[0,pi),
    intensity U[-0.5,1.0].  Row i is y (top = +1), column j is x."""
    rng = np.random.default_rng(seed)
    c = (np.arange(n) + 0.5) / n * 2.0 - 1.0
    X = c[None, :]
    Y = -c[:, None]
    img = np.zeros((n, n))
    params = []
    for _ in range(n_ellipses):
        r = np.sqrt(rng.uniform())
        t = rng.uniform(0, 2 * np.pi)
        x0, y0 = r * np.cos(t), r * np.sin(t)
        a, b = rng.uniform(0.05, 0.4, size=2)
        phi = rng.uniform(0, np.pi)
        rho = rng.uniform(-0.5, 1.0)
        cp, sp = np.cos(phi), np.sin(phi)
        u = (X - x0) * cp + (Y - y0) * sp
        v = -(X - x0) * sp + (Y - y0) * cp
        img += rho * ((u / a) ** 2 + (v / b) ** 2 <= 1.0)
        params.append((x0, y0, a, b, phi, rho))
    img = img.astype(np.float32)
    return (img, params) if return_params else img


def motion_psf(n, length, angle_deg, supersample=16):
    """Linear motion-blur PSF of `length` pixels at `angle_deg`, rasterised
    with sub-pixel line sampling, unit sum, centred at (0,0) with wrap-around
    (for circular convolution of an n x n image)."""
    m = max(int(np.ceil(length * supersample)), 2)
    t = (np.arange(m) + 0.5) / m - 0.5
    th = np.deg2rad(angle_deg)
    xs = t * length * np.cos(th)
    ys = -t * length * np.sin(th)
    psf = np.zeros((n, n))
    x0 = np.floor(xs).astype(int)
    y0 = np.floor(ys).astype(int)
    fx = xs - x0
    fy = ys - y0
    for dx, dy, w in ((0, 0, (1 - fx) * (1 - fy)), (1, 0, fx * (1 - fy)), (0, 1, (1 - fx) * fy), (1, 1, fx * fy)):
        np.add.at(psf, ((y0 + dy) % n, (x0 + dx) % n), w)
    return psf / psf.sum()


def blurred_texture(n=1024, hurst=0.6, seed=0, noise=0.01):
    """fBm texture in [0,1] convolved (circularly) with a motion PSF of length
    U[5,40] px and angle U[0,180) deg, plus N(0, noise^2).  Returns
    (image float32, length, angle)."""
    rng = np.random.default_rng(seed + 7919)
    length = rng.uniform(5.0, 40.0)
    angle = rng.uniform(0.0, 180.0)
    tex = fbm(n, hurst, seed, 0.0, 1.0)
    psf = motion_psf(n, length, angle)
    img = np.real(_ifft2(_fft2(tex) * _fft2(psf)))
    img += rng.normal(0.0, noise, size=img.shape)
    return img.astype(np.float32), length, angle


def viewshed_batch_sizes(count=32, seed=4):
    """Sizes and Hurst exponents of the batched-viewshed set: sides uniform
    from {1000, 2000, 3000, 4000}, H ~ U[0.7, 0.9]."""
    rng = np.random.default_rng(seed)
    sides = rng.choice([1000, 2000, 3000, 4000], size=count)
    hursts = rng.uniform(0.7, 0.9, size=count)
    return [int(s) for s in sides], [float(h) for h in hursts]
