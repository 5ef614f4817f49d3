"""Which kernel runs, and the library's promises about it (GPU).

* arch='hip_mfma' runs the bf16x3 MFMA kernel where its tiling takes the
  call and reports it (``_lib.last_kernel()``); where it does not (step 3,
  N > 1344 - 31*step on the default 32x32 tile shapes) the FP32 kernel runs
  with a RuntimeWarning and the
  result equals arch='hip' bit for bit; ``strict_kernel=True`` raises
  instead, as does FMF_FLAG_STRICT_KERNEL through the C ABI.
* Per-channel output whose template block holds more than 2^29 values
  (n_corr * S * C) stays on the MFMA kernel: every channel at sampled lags,
  the last valid lag included, within 2e-6 of FP64.
* Calls from several threads at once, while another thread changes the
  process-wide tile shape between two 32x32 shapes (which sum in the same
  order), equal the same calls made one after the other, bit for bit.
"""
import threading
import warnings

import numpy as np
import pytest

from tests.golden import cases

pytestmark = pytest.mark.gpu


def _fmf():
    import fast_matched_filter_amd as fmf
    assert fmf.HIP_LOADED
    return fmf


def _edge():
    return cases.edge()


def test_bf16x3_reports_mfma():
    from fast_matched_filter_amd import _lib
    fmf = _fmf()
    inp = cases.selfdet()
    print("case: selfdet hip_mfma step 2", flush=True)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        fmf.matched_filter(inp["templates"], inp["moveouts"], inp["weights"], inp["data"],
                           inp["step"], arch="hip_mfma", check_zeros=False)
    assert _lib.last_kernel() == "bf16x3"
    fmf.matched_filter(inp["templates"], inp["moveouts"], inp["weights"], inp["data"],
                       inp["step"], arch="hip", check_zeros=False)
    assert _lib.last_kernel() == "fp32"


def test_step3_falls_back_with_warning_and_strict_refuses():
    from fast_matched_filter_amd import _lib
    fmf = _fmf()
    inp = _edge()  # step 3
    args = (inp["templates"], inp["moveouts"], inp["weights"], inp["data"], inp["step"])
    print("case: edge hip_mfma step 3", flush=True)
    with pytest.warns(RuntimeWarning, match="FP32 kernel ran"):
        out = fmf.matched_filter(*args, arch="hip_mfma", check_zeros=False)
    assert _lib.last_kernel() == "fp32"
    ref = fmf.matched_filter(*args, arch="hip", check_zeros=False)
    assert np.array_equal(out, ref)
    with pytest.raises(RuntimeError, match="STRICT"):
        fmf.matched_filter(*args, arch="hip_mfma", check_zeros=False, strict_kernel=True)
    assert _lib.last_kernel() is None


def test_long_template_falls_back_with_warning():
    from fast_matched_filter_amd import _lib
    fmf = _fmf()
    rng = np.random.Generator(np.random.PCG64(90))
    N, L = 1400, 20_000
    tp = rng.standard_normal((1, 2, 1, N)).astype(np.float32)
    d = rng.standard_normal((2, 1, L)).astype(np.float32)
    mo = np.array([[[0], [17]]], dtype=np.int32)
    w = np.full((1, 2, 1), 0.5, dtype=np.float32)
    print("case: N=1400 hip_mfma step 1", flush=True)
    with pytest.warns(RuntimeWarning):
        out = fmf.matched_filter(tp, mo, w, d, 1, arch="hip_mfma", check_zeros=False)
    assert _lib.last_kernel() == "fp32"
    with pytest.raises(RuntimeError):
        fmf.matched_filter(tp, mo, w, d, 1, arch="hip_mfma", check_zeros=False,
                           strict_kernel=True)
    # the longest template the 32x32 tiling takes at step 1 stays on the MFMA
    # kernel, one tap more does not
    with pytest.warns(RuntimeWarning):
        fmf.matched_filter(tp[..., :1314], mo, w, d, 1, arch="hip_mfma", check_zeros=False)
    n_ok = 1344 - 31
    out = fmf.matched_filter(tp[..., :n_ok], mo, w, d, 1, arch="hip_mfma", check_zeros=False)
    assert _lib.last_kernel() == "bf16x3"
    from oracle import fp64
    lags = np.array([0, 5, 777, L - n_ok - 17])
    exact = fp64.cc_at(tp[..., :n_ok], mo, w, d, 0, lags)
    assert np.abs(out[0, lags] - exact).max() <= 2e-6


def test_per_channel_output_beyond_2_29_values():
    """64 channels x 8 400 000 lags = 5.4e8 values in one template's block."""
    from fast_matched_filter_amd import _lib
    from oracle import fp64
    fmf = _fmf()
    rng = np.random.Generator(np.random.PCG64(91))
    S, N, L = 64, 64, 8_400_000
    tp = rng.standard_normal((1, S, 1, N)).astype(np.float32)
    d = rng.standard_normal((S, 1, L)).astype(np.float32)
    mo = rng.integers(0, 50, size=(1, S, 1)).astype(np.int32)
    w = np.full((1, S, 1), 1.0 / S, dtype=np.float32)
    n_corr = L - N + 1
    assert n_corr * S >= 2 ** 29
    print("case: per-channel output, S=64, L=8.4e6, hip_mfma", flush=True)
    out = fmf.matched_filter(tp, mo, w, d, 1, arch="hip_mfma", network_sum=False,
                             check_zeros=False)
    assert _lib.last_kernel() == "bf16x3"
    assert out.shape == (1, n_corr, S, 1)
    stop = 1 + L - N - int(mo.max())
    lags = np.unique(np.concatenate([[0, 1, stop - 1], rng.integers(0, stop, 20)]))
    for ch in (0, 17, S - 1):
        one = np.zeros_like(w)
        one[0, ch, 0] = 1.0
        exact = fp64.cc_at(tp, mo, one, d, 0, lags)
        err = np.abs(out[0, lags, ch, 0] - exact).max()
        print("channel %d max |dcc| %.3g" % (ch, err), flush=True)
        assert err <= 2e-6
    assert not out[0, stop:].any()


def test_concurrent_calls_with_tile_shape_changes():
    from fast_matched_filter_amd import _lib
    fmf = _fmf()
    lib = _lib.require()
    inps = [cases.selfdet(), cases.noisy(), cases.small(), cases.selfdet()]
    inps[3] = dict(inps[3], step=1)

    def run(inp):
        return fmf.matched_filter(inp["templates"], inp["moveouts"], inp["weights"],
                                  inp["data"], inp["step"], arch="hip_mfma", check_zeros=False)
    before = {s: int(lib.fmf_get_mfma_tiles(s)) for s in (1, 2)}
    print("case: 4 threads x 3 calls, tile shapes 422/412 flipping", flush=True)
    try:
        seq = [run(i) for i in inps]
        stop = threading.Event()

        def flip():
            k = 0
            while not stop.is_set():
                for s in (1, 2):
                    lib.fmf_set_mfma_tiles(s, 412 if k % 2 else 422)
                k += 1

        got = [[None] * 3 for _ in inps]
        errors = []

        def worker(i):
            try:
                for r in range(3):
                    got[i][r] = run(inps[i])
            except Exception as e:  # reported below
                errors.append(e)
        flipper = threading.Thread(target=flip)
        flipper.start()
        threads = [threading.Thread(target=worker, args=(i,)) for i in range(len(inps))]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=120)
        stop.set()
        flipper.join(timeout=30)
        assert not errors, errors
        for i, rows in enumerate(got):
            for r in rows:
                assert r is not None and np.array_equal(r, seq[i]), i
    finally:
        for s, shape in before.items():
            lib.fmf_set_mfma_tiles(s, shape)
