"""MI355X Milne–Eddington inversion engine: Python binding of libmeinv.so.

The product boundary is the C ABI declared in include/meinv.h; this package
binds it (``abi``), mirrors the reference's inversion entry point
(``engine.Inversion``) and its control-file keywords (``control``), and holds
the benchmark workloads (``workloads``).
"""
from .abi import (ME_FP32, ME_FP64, ME_INIT_HEURISTIC, ME_INIT_LEARNED, MeConfig, MeStats,
                  load_library, reference_defaults)
from .engine import Inversion, MeError, Stats

__all__ = ["ME_FP32", "ME_FP64", "ME_INIT_HEURISTIC", "ME_INIT_LEARNED", "MeConfig", "MeStats",
           "load_library", "reference_defaults", "Inversion", "MeError", "Stats"]
