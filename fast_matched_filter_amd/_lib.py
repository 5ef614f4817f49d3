"""ctypes binding of libfmf_hip.so (the HIP library built for gfx950).

Mirrors the reference binding (fast_matched_filter/fast_matched_filter.py:
15-63): the same 13 common argument types plus the per-entry extras.  Unlike
the reference, a missing or unloadable library is an error at call time
(``require()`` raises) instead of a printed message: this package has no other
compute path to fall back on.
"""
import ctypes as ct
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libfmf_hip.so")

KERNEL_FP32 = 0
KERNEL_BF16X3 = 1
KERNEL_BF16 = 2      # screening pass: one bf16 product, no threshold
FLAG_EXACT_ORDER = 1
FLAG_THRESHOLD_AFTER_SQRT = 2
FLAG_STRICT_KERNEL = 4   # bf16x3: refuse instead of running the FP32 kernel
KERNEL_NAMES = {0: "fp32", 1: "bf16x3", 2: "bf16"}
INJECT_ALLOC = 1    # fmf_debug_inject points
INJECT_STREAM = 2

STATUS = {0: "ok", 1: "invalid argument", 2: "out of memory",
          3: "device error", 4: "unsupported"}

_F = ct.POINTER(ct.c_float)
_I = ct.POINTER(ct.c_int)
_D = ct.POINTER(ct.c_double)
_SZ = ct.c_size_t

# fast_matched_filter.py:19-33
ARGTYPES = [_F, _F, _I, _F, _F, _SZ, _SZ, _SZ, _SZ, _SZ, _SZ, _SZ, _F]

_lib = None
_load_error = None


def _bind(lib):
    lib.matched_filter.argtypes = ARGTYPES + [ct.c_int, ct.c_int]
    lib.matched_filter_precise.argtypes = ARGTYPES + [ct.c_int]
    lib.matched_filter_no_sum.argtypes = ARGTYPES
    lib.matched_filter_precise_no_sum.argtypes = ARGTYPES + [ct.c_int]
    lib.cumsum_square_data.argtypes = [_F, _SZ, _F, _SZ, _SZ, _D]
    lib.fmf_matched_filter_host.argtypes = ARGTYPES + [ct.c_int] * 4
    lib.fmf_workspace_bytes.argtypes = [_SZ] * 5 + [ct.c_int] * 3
    lib.fmf_workspace_bytes.restype = _SZ
    lib.fmf_matched_filter_device.argtypes = (
        [ct.c_void_p] * 5 + [_SZ] * 7 + [ct.c_void_p] + [ct.c_int] * 4
        + [ct.c_void_p, _SZ, ct.c_void_p])
    lib.fmf_noalloc_scratch_bytes.argtypes = []
    lib.fmf_noalloc_scratch_bytes.restype = _SZ
    lib.fmf_matched_filter_device_noalloc.argtypes = (
        [ct.c_void_p] * 5 + [_SZ] * 7 + [ct.c_void_p] + [ct.c_int] * 4 + [ct.c_void_p])
    lib.fmf_matched_filter_device_noalloc.restype = ct.c_int
    lib.fmf_batched_workspace_bytes.argtypes = [_I, ct.c_int] + [_SZ] * 5
    lib.fmf_batched_workspace_bytes.restype = _SZ
    lib.fmf_matched_filter_batched_device.argtypes = (
        [ct.c_void_p] * 6 + [_I, ct.c_int] + [_SZ] * 7 + [ct.c_void_p, ct.c_int, ct.c_int,
                                                         ct.c_longlong, ct.c_void_p, _SZ, ct.c_void_p])
    lib.fmf_matched_filter_batched_device.restype = ct.c_int
    lib.fmf_set_mfma_tiles.argtypes = [ct.c_int, ct.c_int]
    lib.fmf_set_mfma_tiles.restype = ct.c_int
    lib.fmf_get_mfma_tiles.argtypes = [ct.c_int]
    lib.fmf_get_mfma_tiles.restype = ct.c_int
    lib.fmf_kernel_timer_enable.argtypes = [ct.c_int]
    lib.fmf_kernel_timer_enable.restype = ct.c_int
    lib.fmf_kernel_timer_read.argtypes = [ct.c_int, _D, _I]
    lib.fmf_kernel_timer_read.restype = ct.c_int
    lib.fmf_last_error.restype = ct.c_char_p
    lib.fmf_last_error.argtypes = []
    lib.fmf_last_kernel.restype = ct.c_int
    lib.fmf_last_kernel.argtypes = []
    lib.fmf_debug_inject.restype = ct.c_int
    lib.fmf_debug_inject.argtypes = [ct.c_int, ct.c_int]
    lib.fmf_copy_to_host.restype = ct.c_int
    lib.fmf_copy_to_host.argtypes = [ct.c_void_p, ct.c_void_p, _SZ, ct.c_void_p]
    for name in ("matched_filter", "matched_filter_precise",
                 "matched_filter_no_sum", "matched_filter_precise_no_sum",
                 "cumsum_square_data", "fmf_matched_filter_host",
                 "fmf_matched_filter_device"):
        getattr(lib, name).restype = ct.c_int
    lib.fmf_detect_workspace_bytes.argtypes = [ct.c_longlong, ct.c_longlong]
    lib.fmf_detect_workspace_bytes.restype = _SZ
    lib.fmf_lag_ranges_device.argtypes = ([ct.c_void_p, ct.c_longlong, ct.c_longlong, ct.c_int]
                                          + [ct.c_longlong] * 3 + [ct.c_void_p, ct.c_void_p])
    lib.fmf_lag_ranges_device.restype = ct.c_int
    lib.fmf_median_mad_device.argtypes = ([ct.c_void_p, ct.c_longlong, ct.c_longlong]
                                          + [ct.c_void_p] * 4 + [_SZ, ct.c_void_p])
    lib.fmf_median_mad_device.restype = ct.c_int
    lib.fmf_detect_device.argtypes = ([ct.c_void_p, ct.c_longlong, ct.c_longlong]
                                      + [ct.c_void_p] * 3 + [ct.c_float, ct.c_longlong]
                                      + [ct.c_void_p] * 3 + [ct.c_longlong]
                                      + [ct.c_void_p] * 2 + [_SZ, ct.c_void_p])
    lib.fmf_detect_device.restype = ct.c_int
    if hasattr(lib, "matched_filter_variable_precise"):
        lib.matched_filter_variable_precise.argtypes = (
            ARGTYPES[0:6] + [_I] + ARGTYPES[7:] + [ct.c_int])
        lib.matched_filter_variable_precise.restype = ct.c_int
    return lib


def load():
    """Load the library once; returns it or None (and remembers why).

    torch (used by the device-resident entry points for memory and streams)
    ships its own copy of the HIP runtime with the same soname.  Importing it
    first lets this library bind to that copy, so one process never holds two
    HIP runtimes over one HSA runtime (the second would see no device)."""
    global _lib, _load_error
    if _lib is None and _load_error is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        try:
            _lib = _bind(ct.CDLL(LIB_PATH))
        except OSError as exc:  # not built, or built for another host
            _load_error = exc
    return _lib


def require():
    lib = load()
    if lib is None:
        raise RuntimeError(
            "libfmf_hip.so is not loadable ({}): build it with "
            "`make -C fast_matched_filter_amd/csrc` or "
            "`python -c 'import __graft_entry__ as g; g.build()'`".format(_load_error))
    return lib


def last_kernel():
    """Name of the correlation kernel the calling thread's last call ran
    ('fp32', 'bf16x3', 'bf16'), or None if it failed before running one."""
    return KERNEL_NAMES.get(int(require().fmf_last_kernel()))


def check(rc, what):
    """Raise on a non-zero status from the library."""
    if rc != 0:
        msg = require().fmf_last_error().decode(errors="replace")
        kind = MemoryError if rc == 2 else (ValueError if rc == 1 else RuntimeError)
        raise kind("{} failed ({}): {}".format(what, STATUS.get(rc, rc), msg))
