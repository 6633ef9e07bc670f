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
