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
_lib = None
_fns = None


def build(force: bool = False) -> str:
    """Compile the oracle with gcc via oracle/Makefile."""
    if force and os.path.exists(LIB):
        os.remove(LIB)
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def _load():
    global _lib, _fns
    if _fns is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        _fns = bind(_lib, "orc_")
    return _fns


def create(prm: LesParams) -> LesContext:
    """Oracle context with the same method surface as the HIP library's."""
    return LesContext(_load(), prm, "oracle")


def threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
