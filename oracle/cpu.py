"""ORACLE — test infrastructure only.

ctypes bindings to the C restatement (``libfmf_oracle.so``) and to the
compiled reference (``_ref/matched_filter_CPU*.so``), plus a restatement of
the reference wrapper's input normalisation
(fast_matched_filter/fast_matched_filter.py:180-274) so that both libraries
can be driven from the same NumPy inputs as the product wrapper.

Pins: the C restatement is bit-identical to the compiled reference on every
golden case (tests/test_oracle.py); ``prepare`` is pinned by the reference's
own format equivalences run on the restatement, by the FP64 oracle fed the
raw 4D arrays, and by planted mistakes those catch
(tests/test_oracle_prepare.py).
"""
import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libfmf_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "matched_filter_CPU.so")
REF_PORTABLE_SO = os.path.join(HERE, "_ref", "matched_filter_CPU_x86-64-v3.so")

_F = ct.POINTER(ct.c_float)
_I = ct.POINTER(ct.c_int)
_ARGS = [_F, _F, _I, _F, _F] + [ct.c_size_t] * 7 + [_F]
# fast_matched_filter.py:41 — variable_precise takes int* lengths in slot 6
_ARGS_VAR = _ARGS[0:6] + [_I] + _ARGS[7:] + [ct.c_int]

_libs = {}


def _load(path, prefix):
    key = (path, prefix)
    if key in _libs:
        return _libs[key]
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    lib = ct.CDLL(path)
    for name, args in (
        ("matched_filter", _ARGS),
        ("matched_filter_precise", _ARGS + [ct.c_int]),
        ("matched_filter_no_sum", _ARGS),
        ("matched_filter_precise_no_sum", _ARGS + [ct.c_int]),
        ("matched_filter_variable_precise", _ARGS_VAR),
    ):
        fn = getattr(lib, prefix + name)
        fn.argtypes = args
        fn.restype = None
    _libs[key] = lib
    return lib


def oracle_lib():
    """The C restatement (always built by ``make -C oracle``)."""
    lib = _load(ORACLE_SO, "orc_")
    lib.orc_cumsum_square_data.argtypes = [_F, ct.c_size_t, ct.c_size_t,
                                           ct.POINTER(ct.c_double)]
    lib.orc_cumsum_square_data.restype = None
    return lib


def cumsum_square_data(data):
    """Per-channel running sum of squares (S*C, L+1) from the restatement of
    matched_filter.c:660-705."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    L = data.shape[-1]
    n_ch = data.size // L
    out = np.zeros(n_ch * (L + 1), dtype=np.float64)
    oracle_lib().orc_cumsum_square_data(data.ctypes.data_as(_F), L, n_ch,
                                        out.ctypes.data_as(ct.POINTER(ct.c_double)))
    return out.reshape(n_ch, L + 1)


def reference_lib(portable=False):
    """The compiled reference, if ``make -C oracle ref`` has built it."""
    return _load(REF_PORTABLE_SO if portable else REF_SO, "")


def reference_available(portable=False):
    return os.path.exists(REF_PORTABLE_SO if portable else REF_SO)


def prepare(templates, moveouts, weights, data, step, n_samples_template=None):
    """Restates fast_matched_filter.py:180-274: infer (T, S, C), reshape,
    broadcast moveouts/weights, n_corr, sum of squares in the input dtype cast
    to float32, C-contiguous float32/int32 flat arrays."""
    if templates.ndim > data.ndim:
        n_templates = int(templates.shape[0])
        n_stations = int(templates.shape[1])
        n_components = int(templates.shape[2]) if templates.ndim == 4 else 1
    else:
        n_templates = 1
        n_stations = int(templates.shape[0])
        n_components = int(templates.shape[1]) if templates.ndim == 3 else 1
    if n_samples_template is None:
        n_tp = templates.shape[-1]
        n_max = n_tp
    else:
        n_max = int(np.max(n_samples_template))
        n_tp = np.ascontiguousarray(np.asarray(n_samples_template).flatten(), dtype=np.int32)
    templates = templates.reshape(n_templates, n_stations, n_components, n_max)
    n_data = data.shape[-1]
    data = data.reshape(n_stations, n_components, n_data)
    full = n_templates * n_stations * n_components
    moveouts = np.asarray(moveouts)
    weights = np.asarray(weights)
    if moveouts.size * n_components == full and moveouts.size != full:
        moveouts = np.repeat(moveouts, n_components)
    if weights.size * n_components == full and weights.size != full:
        weights = np.repeat(weights, n_components)
    n_corr = int((n_data - n_max) / step + 1)
    sst = np.sum(templates ** 2, axis=-1)
    return dict(
        templates=np.ascontiguousarray(templates.flatten(), dtype=np.float32),
        sst=np.ascontiguousarray(sst.flatten(), dtype=np.float32),
        moveouts=np.ascontiguousarray(moveouts.flatten(), dtype=np.int32),
        weights=np.ascontiguousarray(weights.flatten(), dtype=np.float32),
        data=np.ascontiguousarray(data.flatten(), dtype=np.float32),
        step=int(step), n_tp=n_tp, n_max=n_max, n_data=n_data,
        T=n_templates, S=n_stations, C=n_components, n_corr=n_corr,
    )


def run(templates, moveouts, weights, data, step, *, kind="cpu",
        normalize="short", network_sum=True, n_samples_template=None,
        lib="oracle"):
    """Run one reference algorithm on the CPU.

    kind: 'cpu' | 'precise' | 'variable_precise' (the reference ``arch``).
    lib:  'oracle' (C restatement), 'ref' (compiled reference, native ISA) or
          'ref_portable' (compiled reference, x86-64-v3).
    Returns the output reshaped as the reference wrapper does
    (fast_matched_filter.py:318-321).
    """
    p = prepare(templates, moveouts, weights, data, step, n_samples_template)
    if lib == "oracle":
        L, pre = oracle_lib(), "orc_"
    else:
        L, pre = reference_lib(portable=(lib == "ref_portable")), ""
    T, S, C, n_corr = p["T"], p["S"], p["C"], p["n_corr"]
    norm = 1 if normalize == "full" else 0
    out = np.zeros(T * n_corr * (1 if network_sum else S * C), dtype=np.float32)
    args = [p["templates"].ctypes.data_as(_F), p["sst"].ctypes.data_as(_F),
            p["moveouts"].ctypes.data_as(_I), p["data"].ctypes.data_as(_F),
            p["weights"].ctypes.data_as(_F), p["step"],
            p["n_tp"] if kind != "variable_precise" else p["n_tp"].ctypes.data_as(_I),
            p["n_data"], T, S, C, n_corr, out.ctypes.data_as(_F)]
    if kind == "cpu":
        assert norm == 0
        getattr(L, pre + ("matched_filter" if network_sum else "matched_filter_no_sum"))(*args)
    elif kind == "precise":
        getattr(L, pre + ("matched_filter_precise" if network_sum
                          else "matched_filter_precise_no_sum"))(*args, norm)
    elif kind == "variable_precise":
        assert network_sum
        getattr(L, pre + "matched_filter_variable_precise")(*args, norm)
    else:
        raise ValueError(kind)
    if network_sum:
        return out.reshape(T, n_corr)
    return out.reshape(T, n_corr, S, C)
