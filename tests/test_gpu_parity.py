"""HIP path vs the reference, through libfmf_hip.so on the GPU.

Expected values come from the golden vectors the compiled reference wrote
(tests/golden/reference_golden.npz) and from the oracle on the same seeded
inputs.  Bars:
  * exact_order=True (FP32 kernel, short normalization): bit-identical to the
    reference 'cpu' build.
  * FP32 kernel, default (FMA) order: max |dcc| <= 1e-4 vs the reference, and
    the same argmax lag at every injected event.
  * 'precise' / 'full': max |dcc| <= 1e-4 vs the reference 'precise' arch.
"""
import numpy as np
import pytest

from tests.golden import cases

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _hip():
    import fast_matched_filter_amd as fmf
    assert fmf.HIP_LOADED, "libfmf_hip.so must load on an MI355X"
    return fmf


def _run(inp, **kw):
    fmf = _hip()
    print("case: T=%s step=%s kw=%s" % (inp["templates"].shape, inp["step"], kw), flush=True)
    return fmf.matched_filter(inp["templates"], inp["moveouts"], inp["weights"],
                              inp["data"], inp["step"], check_zeros=False, **kw)


@pytest.fixture(scope="module")
def inputs():
    return {name: build() for name, build in cases.CASES.items()}


@pytest.mark.parametrize("name", ["edge", "selfdet", "noisy"])
def test_exact_order_bit_identical(golden, inputs, name):
    inp = inputs[name]
    ref = golden[name + "__cpu_sum__full"]
    out = _run(inp, arch="hip", exact_order=True)
    assert out.shape == ref.shape
    mism = np.count_nonzero(out.view(np.uint32) != ref.view(np.uint32))
    assert mism == 0, "bit mismatches: %d, max |d| %g" % (mism, np.abs(out - ref).max())


def test_exact_order_small_digest(golden, inputs):
    import hashlib
    out = _run(inputs["small"], arch="hip", exact_order=True)
    assert hashlib.sha256(out.tobytes()).hexdigest() == str(golden["small__cpu_sum__sha"])


@pytest.mark.parametrize("name", ["edge", "selfdet", "noisy"])
@pytest.mark.parametrize("arch", ["hip", "gpu", "cpu"])
def test_fp32_short_sum(golden, inputs, name, arch):
    ref = golden[name + "__cpu_sum__full"]
    out = _run(inputs[name], arch=arch)
    assert np.abs(out - ref).max() <= TOL


@pytest.mark.parametrize("name", ["edge", "selfdet"])
def test_fp32_short_nosum(golden, inputs, name):
    ref = golden[name + "__cpu_nosum__full"]
    out = _run(inputs[name], arch="hip", network_sum=False)
    assert out.shape == ref.shape
    assert np.abs(out - ref).max() <= TOL


@pytest.mark.parametrize("name", ["edge", "selfdet", "noisy"])
@pytest.mark.parametrize("normalize", ["short", "full"])
def test_precise(golden, inputs, name, normalize):
    ref = golden["%s__precise_%s_sum__full" % (name, normalize)]
    out = _run(inputs[name], arch="precise", normalize=normalize)
    assert np.abs(out - ref).max() <= TOL


@pytest.mark.parametrize("name", ["edge", "selfdet"])
@pytest.mark.parametrize("normalize", ["short", "full"])
def test_precise_nosum(golden, inputs, name, normalize):
    ref = golden["%s__precise_%s_nosum__full" % (name, normalize)]
    out = _run(inputs[name], arch="precise", normalize=normalize, network_sum=False)
    assert np.abs(out - ref).max() <= TOL


def test_small_workload_events(golden, inputs):
    """Config 'small': every-97th-lag samples, per-template max/argmax and the
    windows around each injected event match the reference."""
    inp = inputs["small"]
    out = _run(inp, arch="hip")
    g = "small__cpu_sum__"
    assert np.abs(out[:, ::97] - golden[g + "every97"]).max() <= TOL
    assert np.abs(out.max(axis=1) - golden[g + "max"]).max() <= TOL
    np.testing.assert_array_equal(out.argmax(axis=1), golden[g + "argmax"])
    np.testing.assert_array_equal((out == 0).sum(axis=1), golden[g + "zeros"])
    wins = []
    for t, lag in inp["events"]:
        w = out[t, max(0, lag - 64):lag + 65]
        wins.append(np.pad(w, (0, 129 - len(w))))
        # the injected lag is the local maximum of its window
        assert np.argmax(out[t, lag - 50:lag + 51]) == 50
    assert np.abs(np.stack(wins) - golden[g + "event_windows"]).max() <= TOL


def test_full_normalization_small(golden, inputs):
    out = _run(inputs["small"], arch="hip", normalize="full")
    g = "small__precise_full_sum__"
    assert np.abs(out[:, ::97] - golden[g + "every97"]).max() <= TOL
    np.testing.assert_array_equal(out.argmax(axis=1), golden[g + "argmax"])


def test_self_detection(inputs):
    """Templates cut from the data: cc equals the sum of the channel weights
    at the cut lag (fast_matched_filter.py:495-508 harness invariant).  The
    (T, S) weights are repeated over the C components by the wrapper."""
    inp = inputs["selfdet"]
    out = _run(inp, arch="hip")
    C = inp["templates"].shape[2]
    for t, lag in inp["events"]:
        wsum = float(np.sum(inp["weights"][t])) * C
        assert abs(out[t, lag // inp["step"]] - wsum) <= 1e-5 * wsum
        assert out[t].argmax() == lag // inp["step"]


def test_formats_match(inputs):
    """tests/test_python.py:135-153: single-template and traces formats give
    the batch result."""
    inp = inputs["noisy"]
    fmf = _hip()
    w = inp["weights"]
    batch = fmf.matched_filter(inp["templates"], inp["moveouts"], w, inp["data"], 1, arch="hip")
    single = fmf.matched_filter(inp["templates"][0], inp["moveouts"][0], w[0], inp["data"], 1, arch="hip")
    T, S, C, N = inp["templates"].shape
    traces = fmf.matched_filter(inp["templates"].reshape(T, S * C, N), inp["moveouts"].reshape(T, S * C),
                                w.reshape(T, S * C), inp["data"].reshape(S * C, -1), 1, arch="hip")
    np.testing.assert_array_equal(single, batch[0] if single.ndim == 1 else batch)
    np.testing.assert_array_equal(traces, batch)
    assert not np.isnan(batch).any()
    assert batch.max() <= 1.0 and batch.min() >= -1.0


def test_nan_in_weight_zero_channel_is_ignored(inputs):
    inp = dict(inputs["edge"])
    data = inp["data"].copy()
    base = _run(inp, arch="hip")
    # channel (2,0) has weight 0 for template 2 only: poisoning it changes
    # templates 0-1 but template 2 must not see NaN
    data[2, 0, :] = np.nan
    inp["data"] = data
    out = _run(inp, arch="hip")
    assert not np.isnan(out[2]).any()
    assert np.array_equal(out[2], base[2])


# ---- bf16x3 MFMA kernel ('hip_mfma') ---------------------------------------
# Same bar as the FP32 kernel: max |dcc| <= 1e-4 against the reference build
# and the same argmax at every injected event; against the FP64 oracle the
# split products keep the error at FP32 level (<= 2e-6 on these cases).

@pytest.mark.parametrize("name", ["selfdet", "noisy", "edge"])
def test_mfma_short_sum(golden, inputs, name):
    ref = golden[name + "__cpu_sum__full"]
    out = _run(inputs[name], arch="hip_mfma")
    assert np.abs(out - ref).max() <= TOL


@pytest.mark.parametrize("name", ["selfdet", "noisy"])
def test_mfma_full_sum(golden, inputs, name):
    ref = golden[name + "__precise_full_sum__full"]
    out = _run(inputs[name], arch="hip_mfma", normalize="full")
    assert np.abs(out - ref).max() <= TOL


def test_mfma_nosum(golden, inputs):
    ref = golden["selfdet__cpu_nosum__full"]
    out = _run(inputs["selfdet"], arch="hip_mfma", network_sum=False)
    assert out.shape == ref.shape
    assert np.abs(out - ref).max() <= TOL


def test_mfma_small_workload_vs_fp64(golden, inputs):
    from oracle import fp64
    inp = inputs["small"]
    out = _run(inp, arch="hip_mfma")
    g = "small__cpu_sum__"
    assert np.abs(out[:, ::97] - golden[g + "every97"]).max() <= TOL
    np.testing.assert_array_equal(out.argmax(axis=1), golden[g + "argmax"])
    for t, lag in inp["events"]:
        assert np.argmax(out[t, lag - 50:lag + 51]) == 50
    exact = fp64.matched_filter(inp["templates"], inp["moveouts"], inp["weights"],
                                inp["data"], 1)
    err = np.abs(out - exact)
    print("mfma vs fp64: max %.3g rms %.3g" % (err.max(), np.sqrt((err ** 2).mean())))
    assert err.max() <= 2e-6


def test_mfma_self_detection(inputs):
    inp = inputs["selfdet"]
    out = _run(inp, arch="hip_mfma")
    C = inp["templates"].shape[2]
    for t, lag in inp["events"]:
        wsum = float(np.sum(inp["weights"][t])) * C
        assert abs(out[t, lag // inp["step"]] - wsum) <= 1e-5 * wsum


# ---- arch='variable_precise' -----------------------------------------------

@pytest.mark.parametrize("normalize", ["short", "full"])
def test_variable_precise(golden, inputs, normalize):
    inp = inputs["varlen"]
    ref = golden["varlen__varprec_%s__full" % normalize]
    fmf = _hip()
    out = fmf.matched_filter(inp["templates"], inp["moveouts"], inp["weights"],
                             inp["data"], inp["step"], arch="variable_precise",
                             n_samples_template=inp["lengths"], normalize=normalize,
                             check_zeros=False)
    assert out.shape == ref.shape
    scale = max(1.0, float(np.abs(ref).max()))
    assert np.abs(out - ref).max() <= 1e-4 * scale
    # the tanh tail outside the lag range is reproduced exactly
    np.testing.assert_array_equal(out[:, -3:], ref[:, -3:])
