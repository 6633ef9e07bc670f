"""CPU oracle for the skewEngine path (test infrastructure only).

A ctypes binding of ``oracle/_build/libske_oracle.so`` -- a C/OpenMP
restatement of the reference CPU path (src/skewEngine.cuh,
src/kernel{SDEM,RADN,BLUR,UNIT}.h, src/helper.h, src/main.cu).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py may import this package.  The product package (skewengine_amd) never
imports, links or executes it.
"""
from .oracle import *  # noqa: F401,F403
