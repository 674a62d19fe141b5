"""MI355X-native boundary-layer LES timestep (HIP kernels behind a C-ABI).

The C-ABI is include/les_gpu.h; `lib.create` opens a device context and
`config` builds run parameters and seeded initial fields.
"""
from . import abi, config  # noqa: F401

__all__ = ["abi", "config"]
