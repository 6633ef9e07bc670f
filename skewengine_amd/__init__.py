"""MI355X-native skewEngine: skew -> operator -> deskew -> reduce on gfx950.

The compute lives in libskewengine.so (HIP kernels + C++ host engine behind
the C ABI of include/skewengine.h); this package is its Python binding and
the synthetic-input generators.
"""
from .engine import (  # noqa: F401
    KERNEL_CEPSTRUM, KERNEL_IDENTITYDEM, KERNEL_IDENTITYIMG, KERNEL_RADON, KERNEL_SDEM, POVH_DEFAULT,
    SkewBatch, SkewEngine, device_count, execute, plan_sector, run_cli, work, write_sample_files,
)
from ._lib import SkeError, lib  # noqa: F401
