"""ORACLE -- test infrastructure only.

ctypes loader of the oracle (liblesoracle.so, a CPU FP64 restatement of the
reference timestep built from orc_*.c by oracle/Makefile).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module, and only as the checker / CPU baseline.

Parity status of each oracle routine (what pins it) is listed in
tests/test_oracle_*.py and in the repository's design notes.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

from overlappingbls_amd.abi import LesContext, LesParams, bind

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liblesoracle.so")
LIB_FMA = os.path.join(HERE, "liblesoracle_fma.so")
LIB_R2 = os.path.join(HERE, "liblesoracle_r2.so")
_PATHS = {"": LIB, "fma": LIB_FMA, "r2": LIB_R2}
_libs = {}
_fns = {}


def build(force: bool = False) -> str:
    """Compile both oracle builds with gcc via oracle/Makefile."""
    if force:
        for p in _PATHS.values():
            if os.path.exists(p):
                os.remove(p)
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def _load(variant: str):
    if variant not in _fns:
        path = _PATHS[variant]
        if not os.path.exists(path):
            build()
        _libs[variant] = C.CDLL(path)
        _fns[variant] = bind(_libs[variant], "orc_")
    return _fns[variant]


def create(prm: LesParams, fma: bool = False, variant: str = "") -> LesContext:
    """Oracle context with the same method surface as the HIP library's.
    Variants measure the algorithm's own rounding sensitivity: "fma" (fused
    multiply-adds, fma=True) and "r2" (radix-2 transforms: another FFT
    summation order)."""
    v = "fma" if fma else variant
    return LesContext(_load(v), prm, "oracle_" + v if v else "oracle")


def threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
