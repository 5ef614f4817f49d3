"""libfmf_hip.so driven exactly as the reference's own ctypes binding drives
its native libraries (fast_matched_filter/fast_matched_filter.py):

* the 13 argument types of :19-33, set as the GPU symbol's argtypes (:54),
  and the call ``_libGPU.matched_filter(*args, normalize, int(network_sum))``
  of :314-316, whose two extra ints ctypes passes with its default
  conversion — short and full normalization, network sum and per channel;
* ``matched_filter_precise`` / ``_no_sum`` / ``_precise_no_sum`` with the
  argtypes of :38-40 and the calls of :299-306;
* the arguments built as :253-290 builds them (the oracle's restatement of
  that preparation, oracle.cpu.prepare).

Expected values: the golden vectors the compiled reference wrote
(tests/golden/reference_golden.npz) for the CPU symbols, within 1e-4; for the
GPU symbol with 'full', which the reference computes on the GPU only (mean
removed in the kernel, threshold before the square root,
matched_filter.cu:116-140), the FP64 oracle within 1e-4, and the reference's
'precise' full output within 1e-4 where the two threshold rules agree.

``cumsum_square_data`` (matched_filter.c:660-677): bit for bit against the
oracle's restatement (pinned bit-identical to the reference's running sum on
the CPU) and exactly equal to integer prefix sums on integer-valued data.
"""
import ctypes as ct
import os

import numpy as np
import pytest

from tests.golden import cases

pytestmark = pytest.mark.gpu

TOL = 1e-4

# fast_matched_filter.py:19-33
ARGTYPES = [
    ct.POINTER(ct.c_float),  # templates
    ct.POINTER(ct.c_float),  # sum of squares of templates
    ct.POINTER(ct.c_int),    # moveouts
    ct.POINTER(ct.c_float),  # data
    ct.POINTER(ct.c_float),  # weights
    ct.c_size_t,             # step
    ct.c_size_t,             # n_samples_template
    ct.c_size_t,             # n_samples_data
    ct.c_size_t,             # n_templates
    ct.c_size_t,             # n_stations
    ct.c_size_t,             # n_components
    ct.c_size_t,             # n_corr
    ct.POINTER(ct.c_float),  # cc_sums or cc
]


@pytest.fixture(scope="module")
def libs():
    """Two handles on the same file, bound the way :35-55 binds the CPU and
    the GPU library."""
    import fast_matched_filter_amd._lib as L
    import torch  # noqa: F401  (the HIP runtime torch loads, as the package does)
    path = L.LIB_PATH
    assert os.path.exists(path)
    cpu_like = ct.cdll.LoadLibrary(path)
    cpu_like.matched_filter_precise.argtypes = ARGTYPES + [ct.c_int]
    cpu_like.matched_filter_no_sum.argtypes = ARGTYPES
    cpu_like.matched_filter_precise_no_sum.argtypes = ARGTYPES + [ct.c_int]
    gpu_like = ct.cdll.LoadLibrary(path)
    gpu_like.matched_filter.argtypes = ARGTYPES
    return cpu_like, gpu_like


@pytest.fixture(scope="module")
def inputs():
    return {name: build() for name, build in cases.CASES.items() if name != "varlen"}


def _args(inp, network_sum):
    """The argument tuple of fast_matched_filter.py:276-290, and the output."""
    from oracle.cpu import prepare
    p = prepare(inp["templates"], inp["moveouts"], inp["weights"], inp["data"], inp["step"])
    T, S, C, n_corr = p["T"], p["S"], p["C"], p["n_corr"]
    cc = np.zeros(T * n_corr * (1 if network_sum else S * C), dtype=np.float32)
    args = (p["templates"].ctypes.data_as(ct.POINTER(ct.c_float)),
            p["sst"].ctypes.data_as(ct.POINTER(ct.c_float)),
            p["moveouts"].ctypes.data_as(ct.POINTER(ct.c_int)),
            p["data"].ctypes.data_as(ct.POINTER(ct.c_float)),
            p["weights"].ctypes.data_as(ct.POINTER(ct.c_float)),
            p["step"], p["n_tp"], p["n_data"], T, S, C, n_corr,
            cc.ctypes.data_as(ct.POINTER(ct.c_float)))
    shape = (T, n_corr) if network_sum else (T, n_corr, S, C)
    return args, cc, shape, p


def _keep(*arrays):
    """The prepared arrays must outlive the call (ctypes holds raw pointers)."""
    return arrays


@pytest.mark.parametrize("name,network_sum", [
    ("edge", True), ("edge", False), ("selfdet", True), ("selfdet", False), ("noisy", True)])
def test_gpu_symbol_short(libs, golden, inputs, name, network_sum):
    _, gpu = libs
    key = "%s__cpu_%s__full" % (name, "sum" if network_sum else "nosum")
    args, cc, shape, p = _args(inputs[name], network_sum)
    print("case: reference GPU binding, %s short network_sum=%s" % (name, network_sum), flush=True)
    normalize = 0
    rc = gpu.matched_filter(*args, normalize, int(network_sum))  # fast_matched_filter.py:314-316
    _keep(p)
    assert rc == 0
    out = cc.reshape(shape)
    assert np.abs(out - golden[key]).max() <= TOL


@pytest.mark.parametrize("name", ["edge", "selfdet", "noisy"])
@pytest.mark.parametrize("network_sum", [True, False])
def test_gpu_symbol_full(libs, golden, inputs, name, network_sum):
    from oracle import fp64
    _, gpu = libs
    inp = inputs[name]
    args, cc, shape, p = _args(inp, network_sum)
    print("case: reference GPU binding, %s full network_sum=%s" % (name, network_sum), flush=True)
    rc = gpu.matched_filter(*args, 1, int(network_sum))
    _keep(p)
    assert rc == 0
    out = cc.reshape(shape)
    T, S, C = p["T"], p["S"], p["C"]
    exact = fp64.matched_filter(p["templates"].reshape(T, S, C, -1),
                                p["moveouts"].reshape(T, S, C), p["weights"].reshape(T, S, C),
                                p["data"].reshape(S, C, -1), p["step"], "full", network_sum)
    assert np.abs(out - exact).max() <= TOL
    key = "%s__precise_full_%s__full" % (name, "sum" if network_sum else "nosum")
    if key in golden:
        # same values wherever the threshold rules (before / after the
        # square root) agree: every window whose energy product is above 1e-6
        assert np.abs(out - golden[key]).max() <= TOL


@pytest.mark.parametrize("name", ["edge", "selfdet", "noisy"])
@pytest.mark.parametrize("normalize", [0, 1])
def test_cpu_symbol_precise(libs, golden, inputs, name, normalize):
    cpu, _ = libs
    args, cc, shape, p = _args(inputs[name], True)
    print("case: reference CPU binding, precise %s normalize=%d" % (name, normalize), flush=True)
    rc = cpu.matched_filter_precise(*(args + (normalize,)))  # :301-303
    _keep(p)
    assert rc == 0
    key = "%s__precise_%s_sum__full" % (name, "full" if normalize else "short")
    assert np.abs(cc.reshape(shape) - golden[key]).max() <= TOL


@pytest.mark.parametrize("name", ["edge", "selfdet"])
def test_cpu_symbol_no_sum(libs, golden, inputs, name):
    cpu, _ = libs
    args, cc, shape, p = _args(inputs[name], False)
    print("case: reference CPU binding, no_sum %s" % name, flush=True)
    rc = cpu.matched_filter_no_sum(*args)  # :299-300
    _keep(p)
    assert rc == 0
    assert np.abs(cc.reshape(shape) - golden[name + "__cpu_nosum__full"]).max() <= TOL


@pytest.mark.parametrize("name", ["edge", "selfdet"])
@pytest.mark.parametrize("normalize", [0, 1])
def test_cpu_symbol_precise_no_sum(libs, golden, inputs, name, normalize):
    cpu, _ = libs
    args, cc, shape, p = _args(inputs[name], False)
    print("case: reference CPU binding, precise_no_sum %s normalize=%d" % (name, normalize),
          flush=True)
    rc = cpu.matched_filter_precise_no_sum(*(args + (normalize,)))  # :304-306
    _keep(p)
    assert rc == 0
    key = "%s__precise_%s_nosum__full" % (name, "full" if normalize else "short")
    assert np.abs(cc.reshape(shape) - golden[key]).max() <= TOL


def _cumsum(lib, x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    S, C, L = x.shape
    out = np.full(S * C * (L + 1), np.nan, dtype=np.float64)
    w = np.ones(S * C, dtype=np.float32)
    lib.cumsum_square_data.argtypes = [ct.POINTER(ct.c_float), ct.c_size_t,
                                       ct.POINTER(ct.c_float), ct.c_size_t, ct.c_size_t,
                                       ct.POINTER(ct.c_double)]
    lib.cumsum_square_data.restype = ct.c_int
    rc = lib.cumsum_square_data(x.ctypes.data_as(ct.POINTER(ct.c_float)), L,
                                w.ctypes.data_as(ct.POINTER(ct.c_float)), S, C,
                                out.ctypes.data_as(ct.POINTER(ct.c_double)))
    assert rc == 0
    return out.reshape(S * C, L + 1)


def test_cumsum_square_data_bit_identical(libs, oracle_built):
    cpu, _ = libs
    rng = np.random.Generator(np.random.PCG64(71))
    x = (rng.standard_normal((3, 2, 300_001)) * 37.0).astype(np.float32)
    x[1, 0, 5000:9000] = 0.0
    x[2, 1, :] += np.float32(1e3)  # large offset: compensation matters
    print("case: cumsum_square_data (3, 2, 300001)", flush=True)
    got = _cumsum(cpu, x)
    ref = oracle_built.cumsum_square_data(x.reshape(6, -1))
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    assert np.all(got[:, 0] == 0.0)
    assert np.all(got[2, 5001:9001] == got[2, 5000])  # constant over the gap (row 2 = (1, 0))


def test_cumsum_square_data_exact_on_integers(libs):
    cpu, _ = libs
    rng = np.random.Generator(np.random.PCG64(72))
    x = rng.integers(-1000, 1000, size=(2, 3, 70_000)).astype(np.float32)
    print("case: cumsum_square_data integers (2, 3, 70000)", flush=True)
    got = _cumsum(cpu, x)
    sq = x.reshape(6, -1).astype(np.int64) ** 2
    exact = np.concatenate([np.zeros((6, 1), np.int64), np.cumsum(sq, axis=1)], axis=1)
    np.testing.assert_array_equal(got, exact.astype(np.float64))


@pytest.mark.parametrize("name", ["edge", "selfdet"])
def test_arch_gpu_full_against_fp64(inputs, name):
    """arch='gpu' with normalize='full' (the reference sends it to its GPU
    kernel, matched_filter.cu:116-140) runs the FP32 kernel with 'full' and
    the threshold before the square root: FP64 oracle within 1e-4."""
    import fast_matched_filter_amd as fmf
    from oracle import fp64
    inp = inputs[name]
    print("case: arch=gpu full %s" % name, flush=True)
    for network_sum in (True, False):
        out = fmf.matched_filter(inp["templates"], inp["moveouts"], inp["weights"], inp["data"],
                                 inp["step"], arch="gpu", normalize="full",
                                 network_sum=network_sum, check_zeros=False)
        from oracle.cpu import prepare
        p = prepare(inp["templates"], inp["moveouts"], inp["weights"], inp["data"], inp["step"])
        T, S, C = p["T"], p["S"], p["C"]
        exact = fp64.matched_filter(p["templates"].reshape(T, S, C, -1),
                                    p["moveouts"].reshape(T, S, C),
                                    p["weights"].reshape(T, S, C), p["data"].reshape(S, C, -1),
                                    p["step"], "full", network_sum)
        assert out.shape == exact.shape
        assert np.abs(out - exact).max() <= TOL
