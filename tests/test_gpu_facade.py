"""The C++ facade include/skewEngine.hpp (the reference's skewEngine<T>
class and inputData<T>, src/skewEngine.cuh:133-139,253-372), exercised by
tests/facade_check.hip: the reference's sector farm with a built-in
operator, a host operator (kernelcpu form) and a device operator
(kernelcuda form), against the one-thread oracle."""
import ctypes
import os

import numpy as np
import pytest

import oracle
from skewengine_amd import _lib, synth

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "libfacade_check.so")


@pytest.fixture(scope="module")
def fc():
    _lib.lib()  # libskewengine.so first (the facade library links it)
    if not os.path.exists(LIB):
        raise ImportError("tests/_build/libfacade_check.so is not built (python -m skewengine_amd.build)")
    L = ctypes.CDLL(LIB)
    vp = ctypes.c_void_p
    L.facade_farm.argtypes = [ctypes.c_int] * 3 + [vp] + [ctypes.c_int] * 4 + [vp]
    L.facade_plan.argtypes = [ctypes.c_int, vp, ctypes.c_int, vp, vp, vp, vp, vp, vp]
    L.facade_prepare.argtypes = [ctypes.c_int, vp, vp]
    L.facade_fields.argtypes = [ctypes.c_int, vp, ctypes.c_int, vp, vp, vp, vp, vp, vp]
    return L


def _farm(fc, mode, op, grid, s0=0, s1=180, is_max=0, flags=0):
    g = np.ascontiguousarray(grid, np.float32)
    out = np.zeros_like(g)
    assert fc.facade_farm(mode, op, g.shape[0], g.ctypes.data, s0, s1, is_max, flags, out.ctypes.data) == 0
    return out


def test_farm_with_builtin_host_and_device_operators(fc):
    grid = synth.dem(48, 0.8, 5, 1500.0)
    ref, _ = oracle.execute(oracle.IDENTITYDEM, grid, nthreads=1)
    builtin = _farm(fc, 0, _lib.KERNEL_IDENTITYDEM, grid)
    host = _farm(fc, 1, -1, grid)
    device = _farm(fc, 2, -1, grid)
    assert np.array_equal(builtin, ref)
    assert np.array_equal(host, ref)
    assert np.array_equal(device, ref)
    sdem, _ = oracle.execute(oracle.SDEM, grid, nthreads=1, povh=0.15)
    assert np.array_equal(_farm(fc, 0, _lib.KERNEL_SDEM, grid), sdem)
    part, _ = oracle.execute(oracle.IDENTITYDEM, grid, nthreads=1, s0=40, s1=95)
    assert np.array_equal(_farm(fc, 1, -1, grid, 40, 95), part)


def test_prepare_and_public_fields(fc):
    n = 37
    grid = np.random.default_rng(0).standard_normal((n, n)).astype(np.float32)
    out = np.zeros((3, n, n), np.float32)
    assert fc.facade_prepare(n, grid.ctypes.data, out.ctypes.data) == 0
    r = oracle.prepare(grid)
    for k in range(3):
        assert np.array_equal(out[k].ravel(), r[k + 1])
    for angle, fn in [(a, f) for a in (0, 30, 45, 46, 100, 179) for f in (fc.facade_fields, fc.facade_plan)]:
        ints = np.zeros(8, np.int32)
        first = np.zeros(2 * n + 1, np.uint16)
        last = np.zeros(2 * n + 1, np.uint16)
        target = np.zeros(n, np.int32)
        weight = np.zeros(n, np.float32)
        sc = np.zeros(2, np.float64)
        assert fn(n, grid.ctypes.data, angle, ints.ctypes.data, first.ctypes.data, last.ctypes.data,
                  target.ctypes.data, weight.ctypes.data, sc.ctypes.data) == 0
        ref = oracle.sector(oracle.IDENTITYDEM, grid, angle)
        p = ref["p"]
        assert list(ints) == [p.a, p.newAngle, p.sectorType, p.dim_i, p.dim_o, p.offset, p.skewHeight, p.skewWidth]
        H = p.skewHeight
        assert np.array_equal(first[:H], ref["first"][:H]) and np.array_equal(last[:H], ref["last"][:H])
        assert np.array_equal(target, ref["target"][:n]) and np.array_equal(weight, ref["weight"][:n])
        assert sc[0] == p.scale and sc[1] == p.skewness


def test_lazy_mirrors_row_pointer_reduce_and_reset(fc):
    """kernelcpu without a prior skewAlloc() still sees the first sector's
    tables; reduce(T**) equals the flat reduce; resetSkew() after a sector's
    kernel() leaves that sector out (src/skewEngine.cuh:541-570, 843-854)."""
    grid = synth.dem(40, 0.8, 9, 1500.0)
    ref, _ = oracle.execute(oracle.IDENTITYDEM, grid, nthreads=1)
    assert np.array_equal(_farm(fc, 1, -1, grid, flags=1), ref)
    assert np.array_equal(_farm(fc, 1, -1, grid, 7, 8, flags=1),
                          oracle.execute(oracle.IDENTITYDEM, grid, nthreads=1, s0=7, s1=8)[0])
    assert np.array_equal(_farm(fc, 0, _lib.KERNEL_IDENTITYDEM, grid, flags=2), ref)
    sdem, _ = oracle.execute(oracle.SDEM, grid, nthreads=1, povh=0.15)
    assert np.array_equal(_farm(fc, 0, _lib.KERNEL_SDEM, grid, flags=2), sdem)
    part, _ = oracle.execute(oracle.IDENTITYDEM, grid, nthreads=1, s0=20, s1=60)
    assert np.array_equal(_farm(fc, 2, -1, grid, 20, 61, flags=4), part)
