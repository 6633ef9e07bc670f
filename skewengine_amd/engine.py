"""Python mirror of the reference's skewEngine<float> API over the C ABI.

Method names and argument meaning follow src/skewEngine.cuh:281-296 and the
driver of src/main.cu:71-172; the work itself runs in libskewengine.so on the
GPU.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib, SectorParams

KERNEL_SDEM = _lib.KERNEL_SDEM
KERNEL_IDENTITYDEM = _lib.KERNEL_IDENTITYDEM
KERNEL_RADON = _lib.KERNEL_RADON
KERNEL_IDENTITYIMG = _lib.KERNEL_IDENTITYIMG
KERNEL_CEPSTRUM = _lib.KERNEL_CEPSTRUM

#: POVh for a 1.5 m observer on a 10 m grid: (float)(obsheight/step), src/kernelSDEM.h:159.
POVH_DEFAULT = float(np.float32(1.5 / 10.0))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class SkewEngine:
    """skewEngine<float>(dimx, dimy, inData, isGPU=true, deviceId).

    `grid` is a (dimy, dimx) float32 array; the engine keeps the four
    oriented copies (prepare) in device memory.
    """

    def __init__(self, grid, device_id=0):
        grid = _f32(grid)
        if grid.ndim != 2:
            raise ValueError("grid must be 2-D (dimy, dimx)")
        self.dimy, self.dimx = grid.shape
        self._npts = 0
        self._h = ctypes.c_void_p()
        check(lib().ske_create(self.dimx, self.dimy, grid.ctypes.data, device_id, ctypes.byref(self._h)))

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if self._h:
            check(lib().ske_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- the reference's per-sector contract ---------------------------------
    def set_kernel(self, kernel_id, povh=POVH_DEFAULT):
        """The kernelcpu/kernelcuda assignment of src/main.cu:107-136."""
        check(lib().ske_set_kernel(self._h, kernel_id, float(povh)))

    def skew(self, angle):
        check(lib().ske_skew(self._h, int(angle)))

    def kernel(self):
        check(lib().ske_kernel(self._h))

    def deskew(self, is_max=0):
        check(lib().ske_deskew(self._h, int(is_max)))

    def reduce(self, output, is_max=0):
        """reduce(T* output, int isMax): accumulate into `output` in place."""
        if output.dtype != np.float32 or not output.flags.c_contiguous or output.size != self.dimx * self.dimy:
            raise ValueError("output must be a contiguous float32 array of dimx*dimy")
        check(lib().ske_reduce(self._h, output.ctypes.data, int(is_max)))
        return output

    # -- public fields of the current sector ---------------------------------
    def params(self):
        p = SectorParams()
        check(lib().ske_get_params(self._h, ctypes.byref(p)))
        return p

    def tables(self):
        p = self.params()
        target = np.zeros(p.dim_i, np.int32)
        weight = np.zeros(p.dim_i, np.float32)
        first = np.zeros(p.skewHeight, np.uint16)
        last = np.zeros(p.skewHeight, np.uint16)
        check(lib().ske_get_tables(self._h, target.ctypes.data, weight.ctypes.data, first.ctypes.data,
                                   last.ctypes.data))
        return target, weight, first, last

    def skew_buffers(self):
        """(skewInput, skewOutput) as (skewHeight, dim_i) arrays."""
        p = self.params()
        S = np.zeros((p.skewHeight, p.dim_i), np.float32)
        SO = np.zeros((p.skewHeight, p.dim_i), np.float32)
        check(lib().ske_get_skew(self._h, S.ctypes.data, SO.ctypes.data))
        return S, SO

    # -- all sectors in one call ------------------------------------------------
    def run(self, kernel_id, sector_begin=0, sector_end=180, povh=POVH_DEFAULT, copy_out=True):
        """All sectors through skew -> operator -> deskew -> reduce on the GPU."""
        out = np.empty((self.dimy, self.dimx), np.float32) if copy_out else None
        check(lib().ske_run(self._h, int(kernel_id), int(sector_begin), int(sector_end), float(povh),
                            out.ctypes.data if copy_out else None))
        return out

    def synchronize(self):
        check(lib().ske_synchronize(self._h))

    def set_option(self, option, value):
        check(lib().ske_set_option(self._h, int(option), int(value)))

    def get_option(self, option):
        v = ctypes.c_longlong(0)
        check(lib().ske_get_option(self._h, int(option), ctypes.byref(v)))
        return v.value

    def profile(self, name):
        ms = ctypes.c_double(0)
        cnt = ctypes.c_int(0)
        check(lib().ske_profile_get(self._h, name.encode(), ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value

    def profile_reset(self):
        check(lib().ske_profile_reset(self._h))

    def set_sample_points(self, points):
        """Record the local Radon transform's vectors at image cells
        [(x, y), ...] (saveSampleData, src/global.cpp:47-56)."""
        xy = np.ascontiguousarray(np.asarray(points, np.int32).reshape(-1))
        check(lib().ske_set_sample_points(self._h, len(xy) // 2, xy.ctypes.data if len(xy) else None))
        self._npts = len(xy) // 2

    def sample_data(self):
        """The record: float32 (180, points, 4, 64)."""
        out = np.zeros((180, self._npts, 4, 64), np.float32)
        check(lib().ske_get_sample_data(self._h, out.ctypes.data))
        return out

    def counter(self, name):
        """Work counter accumulated since the last profile reset (e.g.
        "op_sdem_lane_steps" with OPT_COUNT_STEPS on)."""
        v = ctypes.c_double(0)
        check(lib().ske_counter_get(self._h, name.encode(), ctypes.byref(v)))
        return v.value


class SkewBatch:
    """Many square grids of uneven sizes through one all-sector run.

    No reference counterpart: each grid's result equals SkewEngine(grid).run(...).
    """

    def __init__(self, grids, device_id=0):
        self._grids = [_f32(g) for g in grids]
        for g in self._grids:
            if g.ndim != 2 or g.shape[0] != g.shape[1]:
                raise ValueError("batch grids must be square 2-D arrays")
        self.sides = [g.shape[0] for g in self._grids]
        n = len(self._grids)
        sides = (ctypes.c_int * n)(*self.sides)
        ptrs = (ctypes.c_void_p * n)(*[g.ctypes.data for g in self._grids])
        self._h = ctypes.c_void_p()
        check(lib().ske_batch_create(n, sides, ptrs, device_id, ctypes.byref(self._h)))
        self._grids = None  # copied to the device

    def close(self):
        if self._h:
            check(lib().ske_batch_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.sides)

    def run(self, kernel_id, sector_begin=0, sector_end=180, povh=POVH_DEFAULT):
        """Enqueue the run (asynchronous)."""
        check(lib().ske_batch_run(self._h, int(kernel_id), int(sector_begin), int(sector_end), float(povh)))

    def output(self, index):
        n = self.sides[index]
        out = np.empty((n, n), np.float32)
        check(lib().ske_batch_output(self._h, int(index), out.ctypes.data))
        return out

    def outputs(self):
        return [self.output(i) for i in range(len(self))]

    def synchronize(self):
        check(lib().ske_batch_synchronize(self._h))

    def set_option(self, option, value):
        check(lib().ske_batch_set_option(self._h, int(option), int(value)))

    def get_option(self, option):
        v = ctypes.c_longlong(0)
        check(lib().ske_batch_get_option(self._h, int(option), ctypes.byref(v)))
        return v.value

    def profile(self, name):
        ms = ctypes.c_double(0)
        cnt = ctypes.c_int(0)
        check(lib().ske_batch_profile_get(self._h, name.encode(), ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value

    def profile_reset(self):
        check(lib().ske_batch_profile_reset(self._h))

    def counter(self, name):
        """Work counter accumulated since the last profile reset (e.g.
        "op_sdem_lane_steps" with OPT_COUNT_STEPS on)."""
        v = ctypes.c_double(0)
        check(lib().ske_batch_counter_get(self._h, name.encode(), ctypes.byref(v)))
        return v.value

    def autotune(self, kernel_id=KERNEL_SDEM, povh=POVH_DEFAULT):
        """Pick the viewshed launch shape: returns (waves, persistent_wgs_per_cu, occupancy, best_ms)."""
        w = ctypes.c_int(0)
        p = ctypes.c_int(0)
        o = ctypes.c_int(0)
        ms = ctypes.c_double(0)
        check(lib().ske_batch_autotune(self._h, int(kernel_id), float(povh), ctypes.byref(w), ctypes.byref(p),
                                       ctypes.byref(o), ctypes.byref(ms)))
        return w.value, p.value, o.value, ms.value


def execute(kernel_id, grid, povh=POVH_DEFAULT, sector_begin=0, sector_end=180):
    """The whole reference driver (allocate, prepare, sector farm, reduce) on one GPU."""
    grid = _f32(grid)
    dimy, dimx = grid.shape
    out = np.empty_like(grid)
    check(lib().ske_execute(int(kernel_id), dimx, dimy, grid.ctypes.data, out.ctypes.data, float(povh),
                            int(sector_begin), int(sector_end)))
    return out


def work(dimx, dimy, s0=0, s1=180):
    """(cells, steps) of the exact maps for sectors [s0, s1)."""
    c = ctypes.c_double(0)
    s = ctypes.c_double(0)
    check(lib().ske_work(dimx, dimy, s0, s1, ctypes.byref(c), ctypes.byref(s)))
    return c.value, s.value


def plan_sector(dimx, dimy, angle):
    """computeParams(angle) + skewShape() (src/skewEngine.cuh:503-539,
    572-596) on the host: (params, target, weight, first, last) with
    target/weight of dim_i entries and first/last of skewHeight entries."""
    p = SectorParams()
    check(lib().ske_plan_sector(dimx, dimy, angle, ctypes.byref(p), None, None, None, None))
    target = np.zeros(p.dim_i, np.int32)
    weight = np.zeros(p.dim_i, np.float32)
    first = np.zeros(p.skewHeight, np.uint16)
    last = np.zeros(p.skewHeight, np.uint16)
    check(lib().ske_plan_sector(dimx, dimy, angle, ctypes.byref(p), target.ctypes.data, weight.ctypes.data,
                                first.ctypes.data, last.ctypes.data))
    return p, target, weight, first, last


def write_sample_files(directory, data):
    """punto<p>{o,i,l,c}.txt of showResultsRADN (src/kernelRADN.h:263-382)
    from a record of shape (180, points, 4, 64)."""
    data = np.ascontiguousarray(data, np.float32)
    check(lib().ske_write_sample_files(str(directory).encode(), data.shape[1], data.ctypes.data))


def device_count():
    n = ctypes.c_int(0)
    check(lib().ske_device_count(ctypes.byref(n)))
    return n.value


def run_cli(args):
    """The DLL entry `run(argc, argv)` (src/main.cu:199-202)."""
    argv = [b"skewengine"] + [a.encode() for a in args]
    arr = (ctypes.c_char_p * len(argv))(*argv)
    return lib().run(len(argv), arr)
