"""MI355X-native matched-filter engine with the API of
beridel/fast_matched_filter (``matched_filter``, ``CPU_LOADED``,
``GPU_LOADED``).  All compute runs in libfmf_hip.so (HIP, gfx950)."""
from . import _lib
from .fast_matched_filter import (matched_filter, matched_filter_batched, prepare_inputs,
                                  test_matched_filter, ARCHS)

HIP_LOADED = _lib.load() is not None
# Names of the reference module (fast_matched_filter/__init__.py:11-12): both
# report whether the one native library of this package is loadable.
GPU_LOADED = HIP_LOADED
CPU_LOADED = HIP_LOADED

__all__ = ["matched_filter", "matched_filter_batched", "prepare_inputs", "test_matched_filter",
           "ARCHS", "HIP_LOADED",
           "GPU_LOADED", "CPU_LOADED"]

__version__ = "1.0.0"
