"""Loader of the in-tree HIP library (overlappingbls_amd/_lib/liblesgpu.so).

There is no CPU fallback: if the library is missing or no GPU is visible,
creating a context raises.
"""
from __future__ import annotations

import ctypes as C
import os

from .abi import LesContext, LesParams, bind

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_lib", "liblesgpu.so")
_lib = None
_fns = None


def load():
    """Load liblesgpu.so (raises FileNotFoundError if it was not built)."""
    global _lib, _fns
    if _fns is None:
        if not os.path.exists(LIB):
            raise FileNotFoundError(f"{LIB} is missing: run overlappingbls_amd.build.build() first")
        _lib = C.CDLL(LIB)
        _fns = bind(_lib, "les_")
    return _fns


def library():
    load()
    return _lib


def create(prm: LesParams) -> LesContext:
    """Device context for one domain (les_create)."""
    return LesContext(load(), prm, "gpu")
