"""32x32x16 MFMA tile shapes of the bf16x3 kernel (fmf_corr_mfma32.hip).

Same bars as the 16x16x32 shapes (tests/test_gpu_parity.py): max |dcc| <= 1e-4
against the reference build's golden vectors and the oracle, <= 2e-6 against
the FP64 oracle on the small workload, the same argmax at injected events.
Every tile shape is run on one case; negative moveouts (windows that start
before the data in the first block) are covered at steps 1 and 2 for both
tile families, and the batched launch with a 32x32 shape equals its single
calls bit for bit.  The interleaved-tile shape (1422: the two tiles of a wave
32 lags apart, sharing data fragments) computes every output from the same
products in the same order as 422, so its outputs equal 422's bit for bit.
"""
import numpy as np
import pytest

from tests.golden import cases

pytestmark = pytest.mark.gpu

TOL = 1e-4
SHAPES32 = (412, 413, 421, 422, 241, 121, 141, 1422)


def _hip():
    import fast_matched_filter_amd as fmf
    assert fmf.HIP_LOADED, "libfmf_hip.so must load on an MI355X"
    return fmf


def _set_shape(step, shape):
    from fast_matched_filter_amd import _lib
    _lib.check(_lib.require().fmf_set_mfma_tiles(step, shape), "fmf_set_mfma_tiles")


def _get_shape(step):
    from fast_matched_filter_amd import _lib
    return int(_lib.require().fmf_get_mfma_tiles(step))


@pytest.fixture
def shape32(request):
    """Sets a 32x32 shape for steps 1 and 2, restores the previous ones."""
    _hip()
    shape = getattr(request, "param", 422)
    old = {s: _get_shape(s) for s in (1, 2)}
    for s in (1, 2):
        _set_shape(s, shape)
    yield shape
    for s, v in old.items():
        _set_shape(s, v)


def _run(inp, **kw):
    fmf = _hip()
    print("case: T=%s step=%s kw=%s shape=%s" % (inp["templates"].shape, inp["step"], kw,
                                                _get_shape(1)), flush=True)
    return fmf.matched_filter(inp["templates"], inp["moveouts"], inp["weights"],
                              inp["data"], inp["step"], check_zeros=False, **kw)


@pytest.fixture(scope="module")
def inputs():
    return {name: build() for name, build in cases.CASES.items()}


@pytest.mark.parametrize("shape32", SHAPES32, indirect=True)
def test_every_shape_short_sum(golden, inputs, shape32):
    for name in ("noisy", "selfdet"):
        ref = golden[name + "__cpu_sum__full"]
        out = _run(inputs[name], arch="hip_mfma")
        assert np.abs(out - ref).max() <= TOL, name


def test_full_sum(golden, inputs, shape32):
    for name in ("selfdet", "noisy"):
        ref = golden[name + "__precise_full_sum__full"]
        out = _run(inputs[name], arch="hip_mfma", normalize="full")
        assert np.abs(out - ref).max() <= TOL, name


def test_nosum(golden, inputs, shape32):
    ref = golden["selfdet__cpu_nosum__full"]
    out = _run(inputs["selfdet"], arch="hip_mfma", network_sum=False)
    assert out.shape == ref.shape
    assert np.abs(out - ref).max() <= TOL


def test_small_workload_vs_fp64(golden, inputs, shape32):
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
    print("mfma32 vs fp64: max %.3g rms %.3g" % (err.max(), np.sqrt((err ** 2).mean())))
    assert err.max() <= 2e-6


def test_self_detection(inputs, shape32):
    inp = inputs["selfdet"]
    out = _run(inp, arch="hip_mfma")
    C = inp["templates"].shape[2]
    for t, lag in inp["events"]:
        wsum = float(np.sum(inp["weights"][t])) * C
        assert abs(out[t, lag // inp["step"]] - wsum) <= 1e-5 * wsum


def _negative_moveouts(step, seed):
    """Moveouts down to -70 samples (windows before the data in the first
    block) and up to +500, N = 300, 9000 lags."""
    rng = np.random.default_rng(seed)
    T, S, C, N = 4, 3, 2, 300
    L = 9000 * step + N + 600
    data = rng.standard_normal((S, C, L)).astype(np.float32)
    templates = rng.standard_normal((T, S, C, N)).astype(np.float32)
    moveouts = rng.integers(-70, 500, size=(T, S, C)).astype(np.int32)
    moveouts[0, 0, 0] = -70
    weights = np.ones((T, S, C), np.float32)
    return dict(templates=templates, moveouts=moveouts, weights=weights, data=data, step=step)


@pytest.mark.parametrize("step", [1, 2])
@pytest.mark.parametrize("family", ["16x16", "32x32"])
def test_negative_moveouts(step, family, oracle_built):
    inp = _negative_moveouts(step, 40 + step)
    old = _get_shape(step)
    _set_shape(step, 43 if family == "16x16" else 422)
    try:
        out = _run(inp, arch="hip_mfma")
    finally:
        _set_shape(step, old)
    ref = oracle_built.run(inp["templates"], inp["moveouts"], inp["weights"], inp["data"], step)
    assert out.shape == ref.shape
    assert np.abs(out - ref).max() <= TOL
    assert np.count_nonzero(ref) > 0.9 * ref.size


def test_batched_rows_equal_single_calls(shape32):
    import fast_matched_filter_amd as fmf
    from fast_matched_filter_amd.synthetic import make_workload
    for step in (1, 2):
        w = make_workload("batched", n_templates=6, n_stations=5, n_samples_data=30_000,
                          stations=(2, 5), max_moveout=400, seed=60 + step)
        print("batched: step=%d lengths=%s" % (step, sorted(set(w["lengths"].tolist()))),
              flush=True)
        out = fmf.matched_filter_batched(w["templates"], w["lengths"], w["moveouts"],
                                         w["weights"], w["data"], step)
        for t in range(len(w["lengths"])):
            n = int(w["lengths"][t])
            one = fmf.matched_filter(w["templates"][t:t + 1, ..., :n], w["moveouts"][t:t + 1],
                                     w["weights"][t:t + 1], w["data"], step, arch="hip_mfma",
                                     check_zeros=False)
            k = one.shape[1]
            assert np.array_equal(out[t, :k], one[0]), t
            assert not np.any(out[t, k:])


def _shape_outputs(shape, fn):
    old = {s: _get_shape(s) for s in (1, 2)}
    try:
        for s in (1, 2):
            _set_shape(s, shape)
        return fn()
    finally:
        for s, v in old.items():
            _set_shape(s, v)


@pytest.mark.parametrize("n", [7, 64, 300, 801, 1313])
def test_interleaved_tiles_equal_plain(n):
    """Shape 1422 against 422 at step 1, bit for bit: template lengths from
    the shortest to the longest the tiling takes (n16 from 4 to 84, both
    loop tails, the next channel's loads issued before the MFMA loop (n16 <
    8) and spread over its first 8 steps), negative moveouts, sum and
    per-channel output, 'short' and 'full' normalization, weight-0
    channels."""
    rng = np.random.default_rng(900 + n)
    T, S, C = 3, 4, 3
    L = 20_000 + n
    data = rng.standard_normal((S, C, L)).astype(np.float32)
    data[1, 2, 3000:5000] = 0.0
    templates = rng.standard_normal((T, S, C, n)).astype(np.float32)
    moveouts = rng.integers(-40, 700, size=(T, S, C)).astype(np.int32)
    weights = rng.uniform(0.1, 1.0, size=(T, S, C)).astype(np.float32)
    weights[1, 2, :] = 0.0
    inp = dict(templates=templates, moveouts=moveouts, weights=weights, data=data, step=1)
    for kw in (dict(), dict(normalize="full"), dict(network_sum=False)):
        a = _shape_outputs(422, lambda: _run(inp, arch="hip_mfma", strict_kernel=True, **kw))
        b = _shape_outputs(1422, lambda: _run(inp, arch="hip_mfma", strict_kernel=True, **kw))
        assert a.shape == b.shape
        assert np.array_equal(a, b), (n, kw, float(np.abs(a - b).max()))
        assert np.count_nonzero(b) > 0.5 * b.size


def test_interleaved_tiles_batched_equal_plain():
    import fast_matched_filter_amd as fmf
    from fast_matched_filter_amd.synthetic import make_workload
    w = make_workload("batched", n_templates=8, n_stations=6, n_samples_data=40_000,
                      stations=(2, 6), max_moveout=600, seed=77)
    print("batched 1422 vs 422: lengths=%s" % sorted(set(w["lengths"].tolist())), flush=True)
    run = lambda: fmf.matched_filter_batched(w["templates"], w["lengths"], w["moveouts"],
                                             w["weights"], w["data"], 1)
    a = _shape_outputs(422, run)
    b = _shape_outputs(1422, run)
    assert np.array_equal(a, b)
