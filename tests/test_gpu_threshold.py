"""The reference's stability threshold on the MFMA kernels
(matched_filter.c:17, 622-625: a channel's cc is 0 where
sum(T^2) * sum(d^2) is not above 1e-6).

The bf16x3 kernel skips the per-window test on every channel whose largest
1/sqrt(window energy) stays below the template's bound, i.e. where no window
can fail it.  This data has a quiet segment (samples scaled by 1e-8: the
product falls to ~1e-13) on one channel, so that channel runs the tested
epilogue while the others skip it, and a quiet stretch shorter than the
template on another, whose windows all pass.  Against direct FP64 dot
products (oracle.fp64.cc_at, pinned in tests/test_oracle_spotcheck.py) at
every lag over the quiet segments and at random lags: per channel and network
sum within 1e-4 (the 'hip_mfma' bar), and exactly 0 per channel where the
reference zeroes it.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    T, S, C, N, L = 2, 2, 2, 64, 30000
    data = rng.standard_normal((S, C, L)).astype(np.float32)
    data[1, 0, 8000:14000] *= np.float32(1e-8)    # longer than the template
    data[0, 1, 20000:20030] *= np.float32(1e-8)   # shorter: every window passes
    templates = rng.standard_normal((T, S, C, N)).astype(np.float32)
    moveouts = rng.integers(0, 50, size=(T, S, C)).astype(np.int32)
    weights = np.full((T, S, C), 0.25, dtype=np.float32)
    return templates, moveouts, weights, data


@pytest.mark.parametrize("step", [1, 2])
def test_quiet_segment_zeroed_like_reference(step):
    import fast_matched_filter_amd as fmf
    from fast_matched_filter_amd import _lib
    from oracle import fp64
    tp, mo, wt, data = _case(11 + step)
    T, S, C, N = tp.shape
    L = data.shape[-1]
    print("case: quiet segment, step %d, hip_mfma" % step, flush=True)
    per = fmf.matched_filter(tp, mo, wt, data, step, arch="hip_mfma", network_sum=False,
                             check_zeros=False)
    assert _lib.last_kernel() == "bf16x3"
    summ = fmf.matched_filter(tp, mo, wt, data, step, arch="hip_mfma", check_zeros=False)
    n_corr = per.shape[1]
    rng = np.random.Generator(np.random.PCG64(5))
    zeroed = 0
    for t in range(T):
        start, stop = fp64.lag_range(mo[t].reshape(-1), step, N, L)
        quiet = np.arange(7800, 14100) // step
        near = np.arange(19900, 20100) // step
        outs = np.unique(np.concatenate([quiet, near, rng.integers(0, n_corr, 300)]))
        outs = outs[(outs * step >= start) & (outs * step < stop)]
        lags = outs * step
        exact_sum = fp64.cc_at(tp, mo, wt, data, t, lags)
        assert np.abs(summ[t, outs] - exact_sum).max() <= 1e-4
        for s in range(S):
            for c in range(C):
                one = np.zeros_like(wt)
                one[t, s, c] = 1.0
                exact = fp64.cc_at(tp, mo, one, data, t, lags)
                got = per[t, outs, s, c]
                assert np.abs(got - exact).max() <= 1e-4, (t, s, c)
                assert np.array_equal(got[exact == 0.0], np.zeros(int((exact == 0.0).sum())))
                zeroed += int((exact == 0.0).sum())
    assert zeroed > 1000  # the quiet segment's windows were exercised
