"""ctypes binding of the CPU oracle (test infrastructure only).

Each wrapper names the reference function it restates; the C source
(oracle/ske_oracle.c) carries the file:line citations.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libske_oracle.so")

SDEM, IDENTITYDEM, RADON, IDENTITYIMG, CEPSTRUM = 0, 1, 2, 3, 4
RADON_LOCAL = 5  # the Radon operator with the reference's LRT flag set


class Params(ctypes.Structure):
    """Per-sector fields set by computeParams (src/skewEngine.cuh:503-539)."""
    _fields_ = [
        ("a", ctypes.c_int), ("newAngle", ctypes.c_int), ("sectorType", ctypes.c_int),
        ("isT", ctypes.c_int), ("isM", ctypes.c_int),
        ("dim_i", ctypes.c_int), ("dim_o", ctypes.c_int), ("dim_l", ctypes.c_int),
        ("offset", ctypes.c_int), ("skewHeight", ctypes.c_int), ("skewWidth", ctypes.c_int),
        ("scale", ctypes.c_double), ("skewness", ctypes.c_double), ("fAngle", ctypes.c_float),
    ]


def build():
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        f32p, f64p, i32p, u16p, u8p = (P(ctypes.c_float), P(ctypes.c_double), P(ctypes.c_int),
                                       P(ctypes.c_uint16), P(ctypes.c_uint8))
        L.orc_fangle.restype = ctypes.c_float
        L.orc_fangle.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_compute_params.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P(Params), i32p, f32p]
        L.orc_skew_shape.argtypes = [P(Params), i32p, u16p, u16p]
        L.orc_prepare.argtypes = [f32p, ctypes.c_int, ctypes.c_int, f32p, f32p, f32p]
        L.orc_skew.argtypes = [P(Params), i32p, f32p, f32p, f32p]
        L.orc_deskew.argtypes = [P(Params), i32p, f32p, f32p, f32p, ctypes.c_int]
        L.orc_reduce.argtypes = [ctypes.c_int, ctypes.c_int, f32p, f32p, f32p, f32p, f32p, ctypes.c_int]
        L.orc_fft.argtypes = [f64p, f64p, ctypes.c_uint, ctypes.c_int]
        L.orc_ifft.argtypes = [f64p, f64p, ctypes.c_uint]
        L.orc_blur_parameter2.restype = ctypes.c_double
        L.orc_blur_parameter2.argtypes = [f64p, ctypes.c_int, i32p]
        L.orc_line_cepstrum.restype = ctypes.c_float
        L.orc_line_cepstrum.argtypes = [ctypes.c_int, f32p, ctypes.c_int, ctypes.c_int]
        L.orc_line_cepstrum_ex.restype = ctypes.c_float
        L.orc_line_cepstrum_ex.argtypes = [ctypes.c_int, f32p, ctypes.c_int, ctypes.c_int, f64p, f64p]
        L.orc_point_radon.restype = ctypes.c_double
        L.orc_point_radon.argtypes = [P(Params), u16p, u16p, f32p, ctypes.c_int, ctypes.c_int, f64p]
        L.orc_point_radon_samples.restype = ctypes.c_double
        L.orc_point_radon_samples.argtypes = [P(Params), u16p, u16p, f32p, ctypes.c_int, ctypes.c_int, f32p]
        L.orc_sdem_observer.restype = ctypes.c_float
        L.orc_sdem_observer.argtypes = [f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, u8p, u8p]
        L.orc_default_threads.restype = ctypes.c_int
        L.orc_execute.restype = ctypes.c_int
        L.orc_execute.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, f32p, f32p, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_float, f64p]
        L.orc_execute_sdem_f64.restype = ctypes.c_int
        L.orc_execute_sdem_f64.argtypes = [ctypes.c_int, ctypes.c_int, f32p, f64p, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_double]
        L.orc_sector.restype = ctypes.c_int
        L.orc_sector.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, f32p, ctypes.c_int, ctypes.c_float,
                                 P(Params), i32p, f32p, u16p, u16p, f32p, f32p]
        L.orc_work.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, f64p, f64p]
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def fangle(dimx, dimy):
    return lib().orc_fangle(dimx, dimy)


def compute_params(dimx, dimy, angle):
    """computeParams + skewShape for one sector -> (params, target, weight, first, last)."""
    dim_l = max(dimx, dimy)
    p = Params()
    target = np.zeros(dim_l, np.int32)
    weight = np.zeros(dim_l, np.float32)
    first = np.zeros(2 * dim_l + 1, np.uint16)
    last = np.zeros(2 * dim_l + 1, np.uint16)
    lib().orc_compute_params(dimx, dimy, angle, ctypes.byref(p), _p(target, ctypes.c_int), _p(weight, ctypes.c_float))
    lib().orc_skew_shape(ctypes.byref(p), _p(target, ctypes.c_int), _p(first, ctypes.c_uint16),
                         _p(last, ctypes.c_uint16))
    return p, target, weight, first, last


def prepare(img):
    """prepare(): the four orientations (src/skewEngine.cuh:962-987)."""
    img = np.ascontiguousarray(img, np.float32)
    dimy, dimx = img.shape
    in1 = np.empty(dimx * dimy, np.float32)
    in2 = np.empty(dimx * dimy, np.float32)
    in3 = np.empty(dimx * dimy, np.float32)
    lib().orc_prepare(_p(img, ctypes.c_float), dimx, dimy, _p(in1, ctypes.c_float),
                      _p(in2, ctypes.c_float), _p(in3, ctypes.c_float))
    return img.reshape(-1), in1, in2, in3


def skew(p, target, weight, src):
    S = np.zeros(p.skewHeight * p.dim_i + p.dim_i, np.float32)
    src = np.ascontiguousarray(src, np.float32)
    lib().orc_skew(ctypes.byref(p), _p(target, ctypes.c_int), _p(weight, ctypes.c_float),
                   _p(src, ctypes.c_float), _p(S, ctypes.c_float))
    return S


def deskew(p, target, weight, SO, O, is_max=0):
    SO = np.ascontiguousarray(SO, np.float32)
    assert O.dtype == np.float32 and O.flags.c_contiguous
    lib().orc_deskew(ctypes.byref(p), _p(target, ctypes.c_int), _p(weight, ctypes.c_float),
                     _p(SO, ctypes.c_float), _p(O, ctypes.c_float), is_max)
    return O


def reduce(dimx, dimy, out, Os, is_max=0):
    ptrs = [None if o is None else _p(o, ctypes.c_float) for o in Os]
    lib().orc_reduce(dimx, dimy, _p(out, ctypes.c_float), *ptrs, is_max)
    return out


def fft(x, normalize=False):
    """helper::fft (src/helper.h:59-106) on a complex vector."""
    x = np.asarray(x, np.complex128)
    re = np.ascontiguousarray(x.real.copy())
    im = np.ascontiguousarray(x.imag.copy())
    lib().orc_fft(_p(re, ctypes.c_double), _p(im, ctypes.c_double), len(x), 1 if normalize else 0)
    return re + 1j * im


def ifft(x):
    x = np.asarray(x, np.complex128)
    re = np.ascontiguousarray(x.real.copy())
    im = np.ascontiguousarray(x.imag.copy())
    lib().orc_ifft(_p(re, ctypes.c_double), _p(im, ctypes.c_double), len(x))
    return re + 1j * im


def blur_parameter2(v):
    v = np.ascontiguousarray(v, np.float64)
    imax = ctypes.c_int(0)
    val = lib().orc_blur_parameter2(_p(v, ctypes.c_double), len(v), ctypes.byref(imax))
    return val, imax.value


def line_cepstrum(j, row, first, last):
    row = np.ascontiguousarray(row, np.float32)
    return lib().orc_line_cepstrum(j, _p(row, ctypes.c_float), first, last)


def line_cepstrum_ex(j, row, first, last):
    """lineCepstrum (src/kernelBLUR.h:153-228) with its intermediates:
    (value, the 32 log powers ln|X_k|^2, the 16 cepstrum values Re C_k whose
    -min(0, .) are the peak rule's samples)."""
    row = np.ascontiguousarray(row, np.float32)
    logps = np.zeros(32, np.float64)
    ceps = np.zeros(16, np.float64)
    v = lib().orc_line_cepstrum_ex(j, _p(row, ctypes.c_float), first, last, _p(logps, ctypes.c_double),
                                   _p(ceps, ctypes.c_double))
    return v, logps, ceps


def sdem_observer(row, first, last, j, povh):
    """One observer of kernelV3 (src/kernelSDEM.h:180-218): (cv, vis_right, vis_left)."""
    row = np.ascontiguousarray(row, np.float32)
    vr = np.zeros(max(last - j - 1, 0) + 1, np.uint8)
    vl = np.zeros(max(j - first, 0) + 1, np.uint8)
    cv = lib().orc_sdem_observer(_p(row, ctypes.c_float), first, last, j, np.float32(povh),
                                 _p(vr, ctypes.c_uint8), _p(vl, ctypes.c_uint8))
    return cv, vr[:max(last - j - 1, 0)], vl[:max(j - first, 0)]


def point_radon(sec, i, j):
    """pointRadon (src/kernelRADN.h:114-182) of skewed cell (i, j) of a
    sector() result: (value, the 64 squared magnitudes the peak rule scans)."""
    spec = np.zeros(64, np.float64)
    v = lib().orc_point_radon(ctypes.byref(sec["p"]), _p(sec["first"], ctypes.c_uint16),
                              _p(sec["last"], ctypes.c_uint16), _p(sec["S"], ctypes.c_float), i, j,
                              _p(spec, ctypes.c_double))
    return v, spec


def point_radon_samples(sec, i, j):
    """The four vectors pointRadon records for a sample point
    (src/kernelRADN.h:164-170) at skewed cell (i, j) of a sector() result:
    float32 array (4, 64) = windowed sums, axis, filtered spectrum, real
    part of its transform."""
    out = np.zeros(4 * 64, np.float32)
    lib().orc_point_radon_samples(ctypes.byref(sec["p"]), _p(sec["first"], ctypes.c_uint16),
                                  _p(sec["last"], ctypes.c_uint16), _p(sec["S"], ctypes.c_float), i, j,
                                  _p(out, ctypes.c_float))
    return out.reshape(4, 64)


def skewed_cell(p, target, x, y, dimx, dimy):
    """The skewed cell (i, j) of sector p whose source is image cell (x, y):
    getXY (src/kernelRADN.h:28-41) inverted."""
    tr = p.sectorType in (1, 2)
    rx = p.sectorType in (1, 3)
    ry = p.sectorType in (1, 2)
    xx = dimx - 1 - x if rx else x
    yy = dimy - 1 - y if ry else y
    j = yy if tr else xx
    i = (xx if tr else yy) + int(target[j])
    return i, j


def default_threads():
    return lib().orc_default_threads()


def execute(alg, img, nthreads=0, s0=0, s1=180, povh=0.15):
    """The OpenMP sector farm (src/main.cu:71-172): returns (out, seconds)."""
    img = np.ascontiguousarray(img, np.float32)
    dimy, dimx = img.shape
    out = np.zeros(dimx * dimy, np.float32)
    secs = ctypes.c_double(0)
    rc = lib().orc_execute(alg, dimx, dimy, _p(img, ctypes.c_float), _p(out, ctypes.c_float), nthreads,
                           s0, s1, np.float32(povh), ctypes.byref(secs))
    if rc != 0:
        raise RuntimeError("oracle execute failed")
    return out.reshape(dimy, dimx), secs.value


def execute_sdem_f64(img, nthreads=0, s0=0, s1=180, povh=0.15):
    """FP64 restatement of the total-viewshed farm (kernelV3,
    src/kernelSDEM.h:168-222): the same maps and scans in double; povh is
    taken as the double nearest the given value."""
    img = np.ascontiguousarray(img, np.float32)
    dimy, dimx = img.shape
    out = np.zeros(dimx * dimy, np.float64)
    rc = lib().orc_execute_sdem_f64(dimx, dimy, _p(img, ctypes.c_float), _p(out, ctypes.c_double), nthreads,
                                    s0, s1, float(povh))
    if rc != 0:
        raise RuntimeError("oracle FP64 execute failed")
    return out.reshape(dimy, dimx)


def sector(alg, img, angle, povh=0.15):
    """One sector's skew + operator on a fresh engine."""
    img = np.ascontiguousarray(img, np.float32)
    dimy, dimx = img.shape
    dim_l = max(dimx, dimy)
    N = dimx * dimy
    p = Params()
    target = np.zeros(dim_l, np.int32)
    weight = np.zeros(dim_l, np.float32)
    first = np.zeros(2 * dim_l + 1, np.uint16)
    last = np.zeros(2 * dim_l + 1, np.uint16)
    S = np.zeros(2 * N + dim_l, np.float32)
    SO = np.zeros(2 * N + dim_l, np.float32)
    rc = lib().orc_sector(alg, dimx, dimy, _p(img, ctypes.c_float), angle, np.float32(povh), ctypes.byref(p),
                          _p(target, ctypes.c_int), _p(weight, ctypes.c_float), _p(first, ctypes.c_uint16),
                          _p(last, ctypes.c_uint16), _p(S, ctypes.c_float), _p(SO, ctypes.c_float))
    if rc != 0:
        raise RuntimeError("oracle sector failed")
    return dict(p=p, target=target, weight=weight, first=first, last=last, S=S, SO=SO)


def work(dimx, dimy, s0=0, s1=180):
    """(cells, steps) from the exact maps: sum L_r and sum L_r(L_r-1)."""
    c = ctypes.c_double(0)
    s = ctypes.c_double(0)
    lib().orc_work(dimx, dimy, s0, s1, ctypes.byref(c), ctypes.byref(s))
    return c.value, s.value
