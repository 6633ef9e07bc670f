"""Build libskewengine.so in-tree with hipcc for gfx950.

    python -m skewengine_amd.build

All sources of csrc/ (HIP kernels, host planner, engine, C ABI, driver) go
into one shared library next to this file.  Device code is compiled without
FMA contraction (the reference CPU path rounds every product and sum) and
without packed-FP32 instruction selection.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libskewengine.so")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SKE_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["kernels_core.hip", "kernels_sdem.hip", "kernels_radn.hip", "kernels_blur.hip", "kernels_fbp.hip",
               "kernels_learned.hip"]
CPP_SOURCES = ["planner.cpp", "engine.cpp", "noalloc.cpp", "fbp.cpp", "learned.cpp", "png_io.cpp", "driver.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
          "-I" + INCLUDE, "-I" + CSRC]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# Per-source extras: the learned CNN keeps its matrix-core accumulators in
# ordinary VGPRs (no accumulation-register copies after every MFMA).
EXTRA = {"kernels_learned.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r.stdout


def build(verbose=False):
    """Compile every source (always, even when the library exists)."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        _run([HIPCC, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + COMMON + DEVICE + EXTRA.get(src, []))
        objs.append(obj)
    for src in CPP_SOURCES:
        obj = os.path.join(objdir, src.replace(".cpp", ".o"))
        _run([HIPCC, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + COMMON + DEVICE)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-o", tmp] + objs + ["--offload-arch=" + ARCH, "-lamdhip64", "-lz"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_facade_check(verbose=False):
    """Compile tests/facade_check.hip, a program written against the C++
    facade include/skewEngine.hpp, into tests/_build/libfacade_check.so
    (loaded by the GPU tests; needs libskewengine.so first)."""
    tests = os.path.join(ROOT, "tests")
    out = os.path.join(tests, "_build", "libfacade_check.so")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    _run([HIPCC, "-x", "hip", os.path.join(tests, "facade_check.hip"), "-o", out + ".tmp", "-shared", "-fPIC", "-O2",
          "-std=c++17", "-I" + INCLUDE] + DEVICE + ["-L" + HERE, "-lskewengine", "-Wl,-rpath,$ORIGIN/../../skewengine_amd"])
    os.replace(out + ".tmp", out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
    build_facade_check(verbose=True)
