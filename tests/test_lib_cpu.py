"""The C-ABI library loads on a machine without a GPU, exports every symbol
include/skewengine.h declares, and its host-only entry points (planning,
argument validation) work without touching a device."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
import skewengine_amd as ske
from skewengine_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "skewengine.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"typedef[^;]*;", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(\w+)\s*\([^;]*\)\s*;", text, flags=re.M)
    return sorted(set(n for n in names if not n.startswith("(")))


def test_library_loads_and_exports_header_symbols():
    L = _lib.lib()
    names = header_functions()
    assert "run" in names and "ske_run" in names and len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert sorted(names) == sorted(_lib.EXPORTS)


def test_version():
    assert _lib.lib().ske_version() >= 10000


@pytest.mark.parametrize("n", [7, 64, 500, 512])
def test_host_planner_work_matches_oracle(n):
    assert ske.work(n, n) == oracle.work(n, n)
    assert ske.work(n, n, 10, 100) == oracle.work(n, n, 10, 100)


def test_argument_errors_without_device():
    L = _lib.lib()
    img = np.zeros((8, 8), np.float32)
    out = np.zeros_like(img)
    h = ctypes.c_void_p()
    assert L.ske_create(8, 8, None, 0, ctypes.byref(h)) == _lib.SKE_OK - 1  # SKE_EINVAL
    assert L.ske_execute(9, 8, 8, img.ctypes.data, out.ctypes.data, 0.15, 0, 180) == -1
    assert L.ske_execute(0, 8, 8, img.ctypes.data, out.ctypes.data, 0.15, 0, 181) == -1
    assert L.ske_execute(0, 8, 6, img.ctypes.data, out.ctypes.data, 0.15, 0, 180) == -5  # non-square
    assert L.ske_execute(0, 1, 1, img.ctypes.data, out.ctypes.data, 0.15, 0, 180) == -1
    assert L.ske_execute(0, 8, 6, img.ctypes.data, out.ctypes.data, 0.15, 0, 180) == -5
    assert b"square" in L.ske_last_error()
    assert L.ske_work(20000, 20000, 0, 180, None, None) == -1
    with pytest.raises(ske.SkeError):
        ske.work(8, 8, 5, 200)


def test_run_rejects_dims_beyond_the_input(tmp_path):
    """--dims may only crop a square corner of the file (readBilData,
    src/kernelSDEM.h:43-50): larger, non-square or negative sizes are refused
    before any device work."""
    bil = tmp_path / "dem.bil"
    np.zeros((16, 16), np.int16).tofile(bil)
    flt = tmp_path / "dem.flt"
    np.zeros((16, 16), np.float32).tofile(flt)
    for f in (bil, flt):
        for dims in ("17x17", "32x32", "8x12", "-4x-4", "1x1", "16001x16001"):
            assert ske.run_cli(["--input", str(f), "--dims", dims]) == _lib.SKE_OK - 1, (f, dims)
    # a BIL file too small to hold a 2x2 grid
    tiny = tmp_path / "tiny.bil"
    np.zeros(3, np.int16).tofile(tiny)
    assert ske.run_cli(["--input", str(tiny)]) == _lib.SKE_OK - 1


@pytest.mark.parametrize("n", [5, 64, 500])
def test_plan_sector_matches_oracle_maps(n):
    """ske_plan_sector (computeParams + skewShape on the host, through the C
    ABI) equals the oracle's maps bit for bit on every sector."""
    for a in range(180):
        p, t, w, f, l = ske.plan_sector(n, n, a)
        rp, rt, rw, rf, rl = oracle.compute_params(n, n, a)
        assert (p.a, p.newAngle, p.sectorType, p.dim_i, p.dim_o, p.offset, p.skewHeight, p.skewWidth) == (
            rp.a, rp.newAngle, rp.sectorType, rp.dim_i, rp.dim_o, rp.offset, rp.skewHeight, rp.skewWidth)
        assert p.scale == rp.scale and p.skewness == rp.skewness
        H = p.skewHeight
        assert np.array_equal(t, rt[:n]) and np.array_equal(w, rw[:n])
        assert np.array_equal(f, rf[:H]) and np.array_equal(l, rl[:H]), a
    with pytest.raises(ske.SkeError):
        ske.plan_sector(n, n, 180)
    with pytest.raises(ske.SkeError):
        ske.plan_sector(n, n + 1, 0)


def test_sample_file_layout(tmp_path):
    """showResultsRADN's text layout (src/kernelRADN.h:280-360): for each
    point and vector two blocks of 180 lines, values [32 + j] then [32 - j]
    for j = 0..31; the spectrum file ('l') repeats the [32 - j] block."""
    npts = 2
    data = np.arange(180 * npts * 4 * 64, dtype=np.float32).reshape(180, npts, 4, 64) * 0.5
    ske.write_sample_files(tmp_path, data)
    names = sorted(os.listdir(tmp_path))
    assert names == sorted("punto%d%s.txt" % (p, k) for p in (1, 2) for k in "oilc")
    for p in range(npts):
        for v, kind in enumerate("oilc"):
            text = open(tmp_path / ("punto%d%s.txt" % (p + 1, kind))).read()
            lines = text.split("\n")
            assert len(lines) == 361 and lines[-1] == "" and lines[0].endswith(" ")
            got = np.array([[float(t) for t in ln.split()] for ln in lines[:-1]])
            j = np.arange(32)
            up, down = data[:, p, v][:, 32 + j], data[:, p, v][:, 32 - j]
            exp = np.concatenate([down if kind == "l" else up, down])
            assert np.array_equal(got, exp.astype(np.float64)), (p, kind)
    # values print like std::ostream's default float format (%g, 6 digits)
    one = np.zeros((180, 1, 4, 64), np.float32)
    one[0, 0, 0, 32] = 1234567.0
    one[0, 0, 0, 33] = 1e-5
    one[0, 0, 0, 34] = -0.0
    ske.write_sample_files(tmp_path, one)
    first = open(tmp_path / "punto1o.txt").readline()
    assert first.startswith("1.23457e+06 1e-05 -0 0 ")
    with pytest.raises(ske.SkeError):
        ske.write_sample_files(tmp_path / "missing_dir", one)
