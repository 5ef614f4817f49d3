"""Pins the detection oracle (oracle/detect_ref.py) on the CPU:
the vectorised definition equals a pure-Python loop (median by sorting,
window test lag by lag) on rows with ties, plateaus, even and odd counts and
sub-ranges; numpy.median agrees with sorting; and two detections of a row are
always more than s lags apart."""
import numpy as np
import pytest

from oracle import detect_ref as ref


def _row(rng, n, kind):
    if kind == "noise":
        return rng.standard_normal(n).astype(np.float32)
    if kind == "quantized":  # many exact ties
        return (rng.integers(-6, 7, n) / 4.0).astype(np.float32)
    if kind == "plateau":
        x = (rng.standard_normal(n) * 0.1).astype(np.float32)
        x[n // 3:n // 3 + 5] = 3.0
        x[2 * n // 3] = 3.0
        return x
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["noise", "quantized", "plateau"])
@pytest.mark.parametrize("n", [1, 2, 37, 200, 201])
@pytest.mark.parametrize("s", [0, 1, 7, 90])
def test_vectorised_equals_bruteforce(kind, n, s):
    rng = np.random.default_rng(n * 31 + s)
    x = _row(rng, n, kind)
    for lo, hi in ((0, n), (n // 5, n - n // 7)):
        if hi <= lo:
            continue
        t, lag, val, med, mad, thr = ref.detect(x[None, :], 1.5, s, [(lo, hi)])
        want, bmed, bmad, bthr = ref.detect_bruteforce(x, 1.5, s, lo, hi)
        assert lag.tolist() == want
        assert np.all(t == 0) and np.array_equal(val, x[lag])
        assert med[0] == np.float32(bmed) and mad[0] == np.float32(bmad)
        assert thr[0] == np.float32(bthr)
        if lag.size > 1:
            assert np.min(np.diff(lag)) > s


def test_median_of_even_count_is_mean_of_middle_pair():
    x = np.array([4.0, 1.0, 3.0, 2.0], np.float32)
    med, mad = ref.median_mad(x)
    assert med == np.float32(2.5) and mad == np.float32(1.0)
