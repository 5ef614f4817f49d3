"""Host-side checks that need no GPU: the C-ABI library loads and exports
every symbol include/fmf_hip.h declares, argument validation fails with a
status instead of touching the device, and the Python wrapper reproduces the
reference wrapper's input handling (fast_matched_filter.py:152-361)."""
import contextlib
import ctypes as ct
import io
import os
import re

import numpy as np
import pytest

import fast_matched_filter_amd as fmf
from fast_matched_filter_amd import _lib
from fast_matched_filter_amd.fast_matched_filter import _check_zeros, prepare_inputs
from oracle import cpu as ocpu
from oracle import fp64

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "fmf_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|size_t|const char \*)\s*\**\s*(\w+)\s*\(", text, re.M)))


def test_library_loads_and_exports_header_symbols():
    lib = _lib.require()
    names = _declared_symbols()
    assert "matched_filter" in names and "fmf_matched_filter_device" in names
    for name in names:
        assert hasattr(lib, name), name
    assert lib.fmf_version() >= 10000


def test_invalid_arguments_return_status_without_device():
    lib = _lib.require()
    f = np.zeros(8, dtype=np.float32)
    i = np.zeros(8, dtype=np.int32)
    F = ct.POINTER(ct.c_float)
    I = ct.POINTER(ct.c_int)
    args = [f.ctypes.data_as(F), f.ctypes.data_as(F), i.ctypes.data_as(I),
            f.ctypes.data_as(F), f.ctypes.data_as(F)]
    # step = 0
    rc = lib.fmf_matched_filter_host(*args, 0, 4, 8, 1, 1, 1, 5, f.ctypes.data_as(F), 0, 1, 0, 0)
    assert rc == 1 and b"step" in lib.fmf_last_error()
    # unknown kernel
    rc = lib.fmf_matched_filter_host(*args, 1, 4, 8, 1, 1, 1, 5, f.ctypes.data_as(F), 0, 1, 7, 0)
    assert rc == 1 and b"kernel" in lib.fmf_last_error()
    # exact order is an FP32 / short-normalization mode
    rc = lib.fmf_matched_filter_host(*args, 1, 4, 8, 1, 1, 1, 5, f.ctypes.data_as(F), 1, 1, 0, 1)
    assert rc == 4
    with pytest.raises(ValueError):
        _lib.check(1, "probe")
    assert lib.fmf_workspace_bytes(800, 10000, 4, 2, 3, 0, 0, 0) > 0


def test_mixed_items_validation_without_device():
    """fmf_matched_filter_items_mixed_device refuses null pointers, negative
    item counts, tile shapes other than 212 and 412 and a short workspace
    before it touches the device."""
    from fast_matched_filter_amd import screener
    lib = screener.lib()
    f = np.zeros(64, dtype=np.float32)
    it = np.zeros(4, dtype=np.int32)
    P = lambda a: ct.c_void_p(a.ctypes.data)
    ins = [P(f)] * 5
    sizes = (1, 8, 64, 1, 1, 1, 57)  # step, N, L, T, S, C, n_corr

    def call(ptrs=ins, n_items=1, n_screen=1, shape=212, ws_bytes=1 << 30):
        return lib.fmf_matched_filter_items_mixed_device(
            *ptrs, *sizes, P(f), P(it), n_items, P(it), n_screen, shape, P(f), ws_bytes, None)

    assert call(ptrs=[None] + ins[1:]) == 1 and b"null" in lib.fmf_last_error()
    assert call(n_items=-1) == 1 and b"counts" in lib.fmf_last_error()
    assert call(n_screen=-2) == 1
    for shape in (0, 422, 43, 213):
        assert call(shape=shape) == 1 and b"212 or 412" in lib.fmf_last_error()
    assert call(ws_bytes=16) == 1 and b"workspace" in lib.fmf_last_error()


@pytest.mark.parametrize("form", ["4d", "traces", "single", "single_traces"])
def test_prepare_inputs_matches_reference_normalisation(form):
    rng = np.random.Generator(np.random.PCG64(3))
    T, S, C, N, L = 3, 4, 2, 16, 300
    tp = rng.standard_normal((T, S, C, N)).astype(np.float32)
    d = rng.standard_normal((S, C, L)).astype(np.float32)
    mo = rng.integers(0, 9, size=(T, S)).astype(np.float32)  # float pads, as in the fixtures
    w = np.full((T, S), 0.125)
    if form == "traces":
        tp, d = tp.reshape(T, S * C, N), d.reshape(S * C, L)
        mo, w = np.repeat(mo, C, axis=1), np.repeat(w, C, axis=1)
    elif form == "single":
        tp, mo, w = tp[0], mo[0], w[0]
    elif form == "single_traces":
        tp, d = tp[0].reshape(S * C, N), d.reshape(S * C, L)
        mo, w = np.repeat(mo[0], C), np.repeat(w[0], C)
    a = prepare_inputs(tp, mo, w, d, 2)
    b = ocpu.prepare(tp, mo, w, d, 2)
    for k in ("templates", "sst", "moveouts", "weights", "data"):
        np.testing.assert_array_equal(a[k], b[k])
    assert (a["T"], a["S"], a["C"], a["n_corr"]) == (b["T"], b["S"], b["C"], b["n_corr"])


def test_incompatible_dimensions_print_and_return_none():
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        r = prepare_inputs(np.zeros((2, 3, 4, 5, 6)), np.zeros((2, 3)), np.zeros((2, 3)),
                           np.zeros((3, 10)), 1)
    assert r is None and "not compatible" in buf.getvalue()


def test_wrapper_argument_errors():
    z = np.zeros((1, 2, 1, 4), dtype=np.float32)
    d = np.zeros((2, 1, 50), dtype=np.float32)
    m = np.zeros((1, 2), dtype=np.int32)
    with pytest.raises(AssertionError):
        fmf.matched_filter(z, m, m, d, 1, arch="cpu", normalize="full")
    with pytest.raises(AssertionError):
        fmf.matched_filter(z, m, m, d, 1, normalize="zero-mean")
    with pytest.raises(ValueError):
        fmf.matched_filter(z, m, m, d, 1, arch="tpu")


def test_check_zeros_messages():
    """fast_matched_filter.py:322-361 on a crafted result."""
    cc = np.ones((2, 100), dtype=np.float32)
    cc[0, :30] = 0.0
    cc[1, :5] = 0.0
    mo = np.zeros(2 * 3, dtype=np.int32)
    w = np.ones(2 * 3, dtype=np.float32)
    for mode, expect in (("first", ["30 correlation"]), ("all", ["30 correlation"]),
                         (False, []), ("bogus", ["check_zeros should be", "30 correlation"])):
        buf = io.StringIO()
        with contextlib.redirect_stdout(buf):
            _check_zeros(cc, mode, True, 100, mo, w, 1, 2)
        out = buf.getvalue()
        for e in expect:
            assert e in out
        assert "5 correlation" not in out
    # per-channel layout: the reference indexes the flattened weights
    # (fast_matched_filter.py:348-351), so the threshold is 10 when the first
    # weight is positive, whatever the number of active channels
    cc4 = np.ones((1, 100, 2, 3), dtype=np.float32)
    cc4[0, :, 0, 0] = 0.0  # 100 zeros over the sliced range
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        _check_zeros(cc4, "first", False, 100, mo[:6], w[:6], 1, 1)
    assert "correlation computations were skipped" in buf.getvalue()
    # 50 zeros over 6 active channels: above 10, below 10 * 6
    cc4 = np.ones((1, 100, 2, 3), dtype=np.float32)
    cc4[0, :50, 0, 0] = 0.0
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        _check_zeros(cc4, "first", False, 100, mo[:6], w[:6], 1, 1)
    assert "50 correlation computations were skipped" in buf.getvalue()
    # first weight 0: threshold 0, a single zero prints
    w0 = w[:6].copy()
    w0[0] = 0.0
    cc4 = np.ones((1, 100, 2, 3), dtype=np.float32)
    cc4[0, 7, 1, 2] = 0.0
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        _check_zeros(cc4, "first", False, 100, mo[:6], w0, 1, 1)
    assert "1 correlation computations were skipped" in buf.getvalue()


def test_bench_optional_legs_report_errors():
    """An optional leg of the bench line returns its result, or its error as
    {"error": ...} without ending the run."""
    import bench
    assert bench.optional_leg("ok", lambda: {"value": 1.0}) == {"value": 1.0}
    assert bench.optional_leg("tuple", lambda: (1, 2)) == (1, 2)

    def boom():
        raise ValueError("no weights")
    r = bench.optional_leg("boom", boom)
    assert r == {"error": "ValueError: no weights"}


def test_bench_lag_counts_follow_reference_rule():
    import bench
    rng = np.random.Generator(np.random.PCG64(8))
    mo = rng.integers(-40, 300, size=(20, 4, 3))
    for step in (1, 2, 3):
        L, N = 5000, 101
        n_corr = (L - N) // step + 1
        counts, olo, ohi = bench.lag_counts(mo, step, N, L, n_corr)
        for t in range(20):
            start, stop = fp64.lag_range(mo[t], step, N, L)
            assert counts[t] == len(range(start, stop, step))
            assert olo[t] == start // step


def test_harness_has_the_reference_signature():
    """test_matched_filter takes the reference's keyword arguments with the
    reference's defaults (fast_matched_filter.py:365-377), plus a seed."""
    import inspect
    from fast_matched_filter_amd import test_matched_filter as harness
    p = inspect.signature(harness).parameters
    ref = dict(n_templates=1, n_stations=1, n_components=1, template_duration=10,
               data_duration=86400, sampling_rate=100, step=1, arch="cpu",
               check_zeros="first", normalize="short", network_sum=True)
    assert list(p)[:len(ref)] == list(ref)
    for k, v in ref.items():
        assert p[k].default == v
    assert p["seed"].default is None
