"""Pins the FP64 spot-check functions that the full-size tests and the bench
use in place of a whole-output oracle run (CPU only):

* ``oracle.fp64.cc_at`` (direct dot products at chosen lags) and
* ``bench.fp64_cc`` (the bench's own FP64 check at injected events)

both equal ``oracle.fp64.matched_filter`` — itself pinned against
pure-Python loops, self-detection, Cauchy-Schwarz and the compiled reference
(tests/test_oracle.py) — at every valid lag of small inputs with negative
moveouts, weight-0 channels, zero gaps, a DC offset, steps 1 to 3, both
normalizations and a template length below the padded one.  Planted
mistakes (a moveout off by one, a dropped channel weight, a template read
past its own length) make the same comparisons fail.
"""
import math

import numpy as np
import pytest

from oracle import fp64


def _case(seed, N=23, n_pad=None, L=700):
    rng = np.random.Generator(np.random.PCG64(seed))
    T, S, C = 2, 3, 2
    n_pad = n_pad or N
    tp = np.zeros((T, S, C, n_pad), dtype=np.float32)
    tp[..., :N] = rng.standard_normal((T, S, C, N)).astype(np.float32)
    if n_pad > N:  # junk past the template length: must never be read
        tp[..., N:] = rng.standard_normal((T, S, C, n_pad - N)).astype(np.float32) * 50
    d = rng.standard_normal((S, C, L)).astype(np.float32)
    d[1, 0, 200:320] = 0.0          # a gap longer than the template
    d[2, 1] += np.float32(2.5)      # DC offset ('full' removes it)
    mo = rng.integers(-15, 40, size=(T, S, C)).astype(np.int32)
    mo[0, 0, 0] = -15
    w = rng.uniform(0.05, 0.4, size=(T, S, C)).astype(np.float32)
    w[1, 2, 0] = 0.0
    d[2, 0, :] = np.where(np.arange(L) % 7 == 0, np.float32(np.nan), d[2, 0])
    w[:, 2, 0] = 0.0                # NaN channel carries weight 0 everywhere
    return tp, mo, w, d


def _valid_lags(mo_t, step, N, L):
    start, stop = fp64.lag_range(mo_t, step, N, L)
    return np.arange(start, stop, step)


@pytest.mark.parametrize("normalize", ["short", "full"])
@pytest.mark.parametrize("step", [1, 2, 3])
def test_cc_at_equals_full_oracle(normalize, step):
    tp, mo, w, d = _case(31 + step)
    full = fp64.matched_filter(tp, mo, w, d, step, normalize)
    L, N = d.shape[-1], tp.shape[-1]
    for t in range(tp.shape[0]):
        lags = _valid_lags(mo[t], step, N, L)
        assert lags.size > 100
        got = fp64.cc_at(tp, mo, w, d, t, lags, normalize=normalize)
        assert np.all(np.isfinite(got))
        np.testing.assert_allclose(got, full[t, lags // step], rtol=0, atol=1e-12)
        # the gap makes some windows of one channel exactly zero: that
        # channel then contributes exactly nothing
        assert np.abs(got).max() <= float(np.abs(w[t]).sum()) + 1e-12


@pytest.mark.parametrize("normalize", ["short", "full"])
def test_cc_at_template_length_override(normalize):
    """n_tp < padded length: only the first n_tp taps count (the padded tail
    holds large junk values here)."""
    n = 19
    tp, mo, w, d = _case(40, N=n, n_pad=32)
    ref = fp64.matched_filter(tp[..., :n], mo, w, d, 1, normalize)
    L = d.shape[-1]
    for t in range(2):
        lags = _valid_lags(mo[t], 1, n, L)
        got = fp64.cc_at(tp, mo, w, d, t, lags, normalize=normalize, n_tp=n)
        np.testing.assert_allclose(got, ref[t, lags], rtol=0, atol=1e-12)


def test_cc_at_brute_force_lags():
    """A few lags straight from the definition in pure Python."""
    tp, mo, w, d = _case(41, N=9, L=120)
    T, S, C, N = tp.shape
    for t in range(T):
        lags = _valid_lags(mo[t], 1, N, d.shape[-1])[::11]
        got = fp64.cc_at(tp, mo, w, d, t, lags)
        for k, lag in enumerate(lags):
            tot = 0.0
            for s in range(S):
                for c in range(C):
                    if w[t, s, c] == 0:
                        continue
                    win = [float(d[s, c, lag + mo[t, s, c] + j]) for j in range(N)]
                    tv = [float(x) for x in tp[t, s, c]]
                    num = sum(a * b for a, b in zip(tv, win))
                    den = sum(a * a for a in tv) * sum(b * b for b in win)
                    tot += float(w[t, s, c]) * (num / math.sqrt(den) if den > 1e-6 else 0.0)
            assert abs(got[k] - tot) < 1e-12


def test_cc_at_planted_mistakes_are_caught():
    tp, mo, w, d = _case(42)
    full = fp64.matched_filter(tp, mo, w, d, 1)
    lags = _valid_lags(mo[0], 1, tp.shape[-1], d.shape[-1])[:-2]
    ok = full[0, lags]
    shifted = fp64.cc_at(tp, mo + 1, w, d, 0, lags)
    assert np.abs(shifted - ok).max() > 1e-2
    w2 = w.copy()
    w2[0, 0, 1] = 0.0
    dropped = fp64.cc_at(tp, mo, w2, d, 0, lags)
    assert np.abs(dropped - ok).max() > 1e-2


def _bench():
    import importlib.util
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(root, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.mark.parametrize("step", [1, 2])
def test_bench_fp64_check_equals_full_oracle(step):
    """The bench's FP64 check (sample lags = output index * step) at every
    valid output index, with a template length below the padded one."""
    bench = _bench()
    n = 21
    tp, mo, w, d = _case(50 + step, N=n, n_pad=30)
    ref = fp64.matched_filter(tp[..., :n], mo, w, d, step)
    L = d.shape[-1]
    for t in range(2):
        outs = _valid_lags(mo[t], step, n, L) // step
        got = bench.fp64_cc(tp, mo, w, d, t, outs, step, n)
        np.testing.assert_allclose(got, ref[t, outs], rtol=0, atol=1e-12)
    # planted: reading the padded tail (length 30 instead of 21) is caught
    outs = _valid_lags(mo[0], step, 30, L) // step
    bad = bench.fp64_cc(tp, mo, w, d, 0, outs, step, 30)
    assert np.abs(bad - ref[0, outs]).max() > 1e-2


def test_bench_lag_counts_match_oracle_lag_range():
    """bench.lag_counts (the MAC count of the metric) against the oracle's
    lag rule, with negative moveouts and per-template lengths."""
    bench = _bench()
    rng = np.random.Generator(np.random.PCG64(60))
    T, L = 40, 5000
    mo = rng.integers(-300, 900, size=(T, 3, 2)).astype(np.int32)
    lengths = rng.choice([50, 120, 400], size=T)
    for step in (1, 2, 3):
        n_corr = (L - int(lengths.min())) // step + 1
        cnt, olo, ohi = bench.lag_counts(mo, step, lengths, L, n_corr)
        for t in range(T):
            start, stop = fp64.lag_range(mo[t], step, int(lengths[t]), L)
            lags = np.arange(start, stop, step)
            assert cnt[t] == lags.size
            if lags.size:
                assert olo[t] == lags[0] // step and ohi[t] == lags[-1] // step + 1


def test_bench_input_digest_of_small_workload():
    """The bench's recorded SHA-256 of the generated inputs: regenerating
    the small workload here gives the recorded digest, and changing one
    sample changes it."""
    import bench
    from fast_matched_filter_amd.synthetic import make_workload
    w = make_workload("small")
    assert bench.input_sha256(w) == bench.INPUT_SHA256["small"]
    w["data"].reshape(-1)[12345] += np.float32(1.0)
    assert bench.input_sha256(w) != bench.INPUT_SHA256["small"]
