"""Detection post-processing on the GPU (fmf_detect.hip) against the oracle
(oracle/detect_ref.py, pinned by tests/test_detect_oracle.py).  Bars (integer
and order-statistics work): median, MAD and threshold bit-identical to
numpy's float32 results; detection lists identical (template, lag, value,
order); the total count reported past the capacity; the lag ranges equal the
reference rule (oracle/fp64.py::lag_range); the one-call matched filter +
detection equals detection on the cc the matched filter returns, and finds
the injected events."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _det():
    import fast_matched_filter_amd as fmf
    assert fmf.HIP_LOADED, "libfmf_hip.so must load on an MI355X"
    from fast_matched_filter_amd import detect
    return detect


def _rows(seed, T, n, kind):
    rng = np.random.default_rng(seed)
    if kind == "noise":
        x = rng.standard_normal((T, n)).astype(np.float32) * 0.05
        for t in range(T):  # a few peaks per row
            for j in rng.integers(0, n, 4):
                x[t, j] = 1.0 + rng.random()
        return x
    if kind == "quantized":
        return (rng.integers(-40, 41, (T, n)) / 64.0).astype(np.float32)
    if kind == "plateau":
        x = rng.standard_normal((T, n)).astype(np.float32) * 0.05
        x[:, n // 4:n // 4 + 9] = 2.0
        x[:, n // 2] = 2.0
        if n // 2 + 3 < n:
            x[:, n // 2 + 3] = 2.0
        return x
    if kind == "ties":
        # 60 % of a row at one value: the select's first group holds more
        # keys than the candidate buffer (n/8), for the median (the tied
        # value) and the MAD (|x - median| = 0), so the last pass rereads
        # the row
        x = rng.standard_normal((T, n)).astype(np.float32) * 0.05
        x[rng.random((T, n)) < 0.6] = np.float32(0.125)
        return x
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["noise", "quantized", "plateau", "ties"])
@pytest.mark.parametrize("n", [1, 2, 1001, 65536, 200_003, 300_001, 700_000])
def test_median_mad_exact(kind, n):
    """Rows of 2^18 lags and more take the sampled bracket (rows with a
    shorter range, and the tied rows, whose bracket overflows, fall back to
    the select over the whole row); shorter calls the three-pass select."""
    import torch
    from oracle import detect_ref as ref
    det = _det()
    T = 5
    x = _rows(n, T, n, kind)
    rng = np.random.default_rng(n + 7)
    ranges = np.array([[0, n], [n // 3, n], [0, max(1, n - n // 5)], [n // 2, n // 2], [1, n]],
                      np.int32)
    ranges[:, 1] = np.maximum(ranges[:, 1], ranges[:, 0])
    ranges = np.minimum(ranges, n)
    print("median/mad: kind=%s n=%d" % (kind, n), flush=True)
    for rg in (None, ranges):
        med, mad = det.median_mad(torch.from_numpy(x).cuda(), None if rg is None else rg)
        med, mad = med.cpu().numpy(), mad.cpu().numpy()
        for t in range(T):
            lo, hi = (0, n) if rg is None else rg[t]
            rm, ra = ref.median_mad(x[t, lo:hi])
            assert med[t] == rm and mad[t] == ra, (t, lo, hi, med[t], rm, mad[t], ra)


@pytest.mark.parametrize("kind", ["noise", "quantized", "plateau"])
@pytest.mark.parametrize("s", [0, 1, 63, 64, 200, 5000])
def test_detections_equal_oracle(kind, s):
    import torch
    from oracle import detect_ref as ref
    det = _det()
    T, n = 4, 120_000
    x = _rows(s + 11, T, n, kind)
    ranges = np.array([[0, n], [777, n - 1000], [5, 6000], [0, n - 1]], np.int32)
    print("detect: kind=%s s=%d" % (kind, s), flush=True)
    for rg in (None, ranges):
        got = det.detect(torch.from_numpy(x).cuda(), n_mad=4.0, min_separation=s, ranges=rg)
        t, lag, val, med, mad, thr = ref.detect(x, 4.0, s, rg)
        assert np.array_equal(got["median"], med) and np.array_equal(got["mad"], mad)
        assert np.array_equal(got["threshold"], thr)
        assert np.array_equal(got["template"], t)
        assert np.array_equal(got["lag"], lag)
        assert np.array_equal(got["value"], val)


def test_capacity_reports_total_and_keeps_order():
    import torch
    from fast_matched_filter_amd import _lib
    from oracle import detect_ref as ref
    det = _det()
    T, n = 3, 50_000
    x = _rows(5, T, n, "quantized")
    cc = torch.from_numpy(x).cuda()
    med, mad = det.median_mad(cc)
    t, lag, val, *_ = ref.detect(x, 0.5, 3)
    assert lag.size > 40
    cap = 17
    lib = _lib.require()
    ws = torch.empty(int(lib.fmf_detect_workspace_bytes(T, n)), dtype=torch.uint8, device="cuda")
    dt = torch.full((cap,), -1, dtype=torch.int32, device="cuda")
    dl = torch.full((cap,), -1, dtype=torch.int32, device="cuda")
    dv = torch.zeros(cap, dtype=torch.float32, device="cuda")
    nd = torch.zeros(1, dtype=torch.int64, device="cuda")
    rc = lib.fmf_detect_device(cc.data_ptr(), T, n, 0, med.data_ptr(), mad.data_ptr(), 0.5, 3,
                               dt.data_ptr(), dl.data_ptr(), dv.data_ptr(), cap, nd.data_ptr(),
                               ws.data_ptr(), ws.numel(), torch.cuda.current_stream().cuda_stream)
    _lib.check(rc, "fmf_detect_device")
    assert int(nd.item()) == lag.size
    assert np.array_equal(dt.cpu().numpy(), t[:cap]) and np.array_equal(dl.cpu().numpy(), lag[:cap])
    full = det.detect(cc, n_mad=0.5, min_separation=3, capacity=cap)  # retried with room
    assert np.array_equal(full["lag"], lag) and np.array_equal(full["template"], t)


def test_lag_ranges_follow_reference_rule():
    from oracle import fp64
    det = _det()
    rng = np.random.default_rng(3)
    T, SC, N, L = 6, 7, 300, 50_000
    mo = rng.integers(-120, 900, (T, SC)).astype(np.int32)
    mo[2] = np.abs(mo[2])
    mo[3, :] = 40_000  # leaves few lags
    for step in (1, 2, 3):
        n_corr = (L - N) // step + 1
        got = det.lag_ranges(mo, N, L, step, n_corr).cpu().numpy()
        for t in range(T):
            start, stop = fp64.lag_range(mo[t], step, N, L)
            # lags start, start + step, ... < stop are output indices
            # start / step ... ceil(stop / step) - 1
            hi = min(n_corr, -(-stop // step)) if stop > 0 else 0
            lo = min(start // step, hi)
            assert tuple(got[t]) == (lo, hi), (step, t, got[t], start, stop)


def test_matched_filter_detect_finds_injected_events():
    import fast_matched_filter_amd as fmf
    from fast_matched_filter_amd.synthetic import make_workload
    from oracle import detect_ref as ref, fp64
    det = _det()
    w = make_workload("small")
    print("end to end: small workload", flush=True)
    out = det.matched_filter_detect(w["templates"], w["moveouts"], w["weights"], w["data"],
                                    w["step"], n_mad=8.0)
    cc = fmf.matched_filter(w["templates"], w["moveouts"], w["weights"], w["data"], w["step"],
                            arch="hip_mfma", check_zeros=False)
    T, n_corr = cc.shape
    N = w["templates"].shape[-1]
    L = w["data"].shape[-1]
    mo = np.asarray(w["moveouts"]).reshape(T, -1)
    rg = np.zeros((T, 2), np.int32)
    for t in range(T):
        start, stop = fp64.lag_range(mo[t], w["step"], N, L)
        rg[t] = (start // w["step"], min(n_corr, (stop + w["step"] - 1) // w["step"]))
    t, lag, val, med, mad, thr = ref.detect(cc, 8.0, N, rg)
    assert np.array_equal(out["template"], t) and np.array_equal(out["lag"], lag)
    assert np.array_equal(out["value"], val)
    found = set(zip(out["template"].tolist(), out["lag"].tolist()))
    for te, le in w["events"]:
        assert (te, le) in found, (te, le)


# The sampled bracket of fmf_detect.hip: 16384 keys from 4096 chunks of 4 lags
# at c * (n - 4) // 4095, bracket ends 256 sample ranks either side of the
# lower middle rank's sample position.
SAMPLE, CHUNK, HALF = 16384, 4, 256


def _bracket_edge_row(n, upper_from_sample):
    """An even row of n = 2^18 lags whose lower middle value is the largest
    key of the sampled bracket [a, b] and whose upper middle value is the
    smallest key above b: the sample holds n_s_low values up to b and the
    rest above it (n_s_low = sample rank of b + 1, with that rank
    floor(k / n * 16384) + 256 for k = n/2 - 1); the other lags hold
    n/2 - n_s_low values below every sampled one, and the rest above b — the
    smallest value above b sits on an unsampled lag or on a sampled one."""
    chunks = SAMPLE // CHUNK
    pos = np.concatenate([c * (n - CHUNK) // (chunks - 1) + np.arange(CHUNK)
                          for c in range(chunks)])
    assert np.unique(pos).size == SAMPLE
    others = np.setdiff1d(np.arange(n), pos)
    k = n // 2 - 1
    n_s_low = int(k / n * SAMPLE) + HALF + 1
    b = 1000.0 + n_s_low - 1
    rng = np.random.default_rng(11 + int(upper_from_sample))
    row = np.empty(n, dtype=np.float32)
    n_low = n // 2 - n_s_low
    s_low = 1000.0 + np.arange(n_s_low)
    s_high = (b + 1.0 if upper_from_sample else 20000.0) + np.arange(SAMPLE - n_s_low)
    o_high = (30000.0 if upper_from_sample else b + 1.0) + np.arange(others.size - n_low)
    row[pos] = rng.permutation(np.concatenate([s_low, s_high])).astype(np.float32)
    row[others] = rng.permutation(np.concatenate([-1.0 - np.arange(n_low), o_high])).astype(
        np.float32)
    return row, b


def test_median_upper_middle_just_above_the_bracket():
    """The sampled-bracket median where rank k + 1 of an even row is the
    smallest key above the bracket (key_hi from the counting pass's minimum
    above b): median and MAD equal numpy's bit for bit."""
    import torch
    from fast_matched_filter_amd import detect
    from oracle import detect_ref as ref
    n = 1 << 18
    (r0, b), (r1, _) = _bracket_edge_row(n, False), _bracket_edge_row(n, True)
    rows = np.stack([r0, r1])
    s = np.sort(rows, axis=1)
    assert np.all(s[:, n // 2 - 1] == b) and np.all(s[:, n // 2] == b + 1.0)
    print("case: bracket edge rows, 2 x 2^18 lags", flush=True)
    med, mad = detect.median_mad(torch.from_numpy(rows).cuda(),
                                 np.array([[0, n], [0, n]], dtype=np.int32))
    med, mad = med.cpu().numpy(), mad.cpu().numpy()
    for t in range(2):
        rm, ra = ref.median_mad(rows[t])
        assert rm == np.float32(b + 0.5)
        assert med[t] == rm and mad[t] == ra, (t, med[t], rm, mad[t], ra)
