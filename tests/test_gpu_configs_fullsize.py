"""The two largest workloads at their full size, through properties that do
not need a CPU run of the whole output.

Precision workload: 500 templates, 20 stations x 3 components, N = 800,
one day at 100 Hz (8 640 000 samples, 5 % gaps), bf16x3 ('hip_mfma') through
the host API.  The output holds 500 x 8 639 201 values, so rows past
template 248 start beyond element 2^31:
  * FP64 (oracle.fp64.cc_at, pinned in tests/test_oracle_spotcheck.py) at
    every injected event and 32 random lags of templates spread over the
    index range, the last one included: <= 2e-6;
  * the lag of the largest cc within +-32 lags of each of those events
    equals the FP64 one;
  * every lag past the reference's last lag is exactly 0, and the rows of
    the sampled templates are non-zero inside their valid lags.

Batched workload: 2000 templates over 40 stations x 3 components and the
whole day, 8-30 active stations per template, lengths in {400, 600, 800,
1200}, moveouts U[0, 3000), steps 1 and 2, one launch:
  * sampled rows (the first and last of every length group, and the last
    template) equal the single-length bf16x3 call on that template alone bit
    for bit, and are zero past their valid lags;
  * the same rows agree with the oracle's C restatement of the reference
    'cpu' path (pinned bit-identical to the compiled reference) on the lags
    of the first 200 000 samples: <= 1e-4;
  * FP64 at the injected events of the sampled rows: <= 2e-6.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_FP64 = 2e-6
TOL_REF = 1e-4


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def day500():
    from fast_matched_filter_amd.synthetic import make_workload
    _say("generating the 500-template day workload")
    return make_workload("day500")


@pytest.fixture(scope="module")
def day500_out(day500):
    import fast_matched_filter_amd as fmf
    w = day500
    _say("case: day500 hip_mfma host call, T=500 step 1")
    out = fmf.matched_filter(w["templates"], w["moveouts"], w["weights"], w["data"], 1,
                             arch="hip_mfma", check_zeros=False)
    _say("day500 output %s" % (out.shape,))
    return out


def _spread(T, k=8):
    return sorted(set(np.linspace(0, T - 1, k).astype(int).tolist()))


def test_day500_output_past_2_31(day500_out):
    T, n_corr = day500_out.shape
    assert T == 500 and (T - 1) * n_corr > 2 ** 31
    assert day500_out.dtype == np.float32


def test_day500_fp64_at_events(day500, day500_out):
    from oracle import fp64
    w, out = day500, day500_out
    T, n_corr = out.shape
    N, L = w["templates"].shape[-1], w["data"].shape[-1]
    rng = np.random.default_rng(500)
    for t in _spread(T):
        ev = np.asarray([lag for tt, lag in w["events"] if tt == t], dtype=np.int64)
        stop = 1 + L - N - int(np.max(w["moveouts"][t]))
        lags = np.unique(np.concatenate([ev, rng.integers(0, stop, 32)]))
        exact = fp64.cc_at(w["templates"], w["moveouts"], w["weights"], w["data"], t, lags)
        err = np.abs(out[t, lags] - exact).max()
        _say("day500 t=%d max |dcc| %.3g over %d lags" % (t, err, lags.size))
        assert err <= TOL_FP64


def test_day500_argmax_at_events(day500, day500_out):
    from oracle import fp64
    w, out = day500, day500_out
    T = out.shape[0]
    n_ev = 0
    for t in _spread(T):
        for tt, lag in w["events"]:
            if tt != t:
                continue
            win = np.arange(lag - 32, lag + 33)
            exact = fp64.cc_at(w["templates"], w["moveouts"], w["weights"], w["data"], t, win)
            assert int(np.argmax(out[t, win])) == int(np.argmax(exact)), (t, lag)
            n_ev += 1
    assert n_ev == 20 * len(_spread(T))


def test_day500_zero_tail(day500, day500_out):
    w, out = day500, day500_out
    N, L = w["templates"].shape[-1], w["data"].shape[-1]
    for t in _spread(out.shape[0]):
        stop = 1 + L - N - int(np.max(w["moveouts"][t]))
        assert not out[t, stop:].any(), t
        assert np.count_nonzero(out[t, :stop]) > 0.9 * stop


@pytest.fixture(scope="module")
def batched():
    from fast_matched_filter_amd.synthetic import make_workload
    _say("generating the 2000-template batched workload")
    return make_workload("batched")


def _sampled_rows(lengths):
    rows = set()
    for n in np.unique(lengths):
        r = np.nonzero(lengths == n)[0]
        rows.update([int(r[0]), int(r[-1])])
    rows.add(len(lengths) - 1)
    return sorted(rows)


@pytest.fixture(scope="module", params=[1, 2])
def batched_run(request, batched):
    """(step, device output, device copies of the inputs)."""
    import torch
    import fast_matched_filter_amd as fmf
    w = batched
    step = request.param
    _say("case: batched T=2000 S=40 step %d, one launch" % step)
    out = fmf.matched_filter_batched(w["templates"], w["lengths"], w["moveouts"], w["weights"],
                                     w["data"], step, return_device=True)
    torch.cuda.synchronize()
    _say("batched step %d output %s" % (step, tuple(out.shape)))
    yield step, out
    del out
    torch.cuda.empty_cache()


@pytest.fixture(scope="module")
def batched_data_dev(batched):
    import torch
    return torch.from_numpy(batched["data"].reshape(-1)).cuda()


def test_batched_rows_equal_single_length_calls(batched, batched_run, batched_data_dev):
    import torch
    from fast_matched_filter_amd import _lib
    from fast_matched_filter_amd.device import matched_filter_device
    w = batched
    step, out = batched_run
    T, S, C, Nmax = w["templates"].shape
    L = w["data"].shape[-1]
    lengths = w["lengths"]
    assert out.shape == (T, (L - int(lengths.min())) // step + 1)
    for t in _sampled_rows(lengths):
        n = int(lengths[t])
        tp = np.ascontiguousarray(w["templates"][t:t + 1, :, :, :n])
        sst = np.sum(tp ** 2, axis=-1).astype(np.float32)
        k = (L - n) // step + 1
        single = torch.empty((1, k), dtype=torch.float32, device="cuda")
        dv = lambda a: torch.from_numpy(np.ascontiguousarray(a).reshape(-1)).cuda()
        _say("single-length call t=%d N=%d step %d" % (t, n, step))
        matched_filter_device(dv(tp), dv(sst), dv(w["moveouts"][t:t + 1]), dv(w["weights"][t:t + 1]),
                              batched_data_dev, step, n, L, 1, S, C, k, single,
                              kernel=_lib.KERNEL_BF16X3)
        row = out[t]
        assert torch.equal(row[:k], single[0]), "template %d (N=%d)" % (t, n)
        assert not bool(row[k:].any())
        stop = 1 + L - n - int(np.max(w["moveouts"][t]))
        last = (stop - 1) // step
        assert not bool(row[last + 1:].any())
        assert int(torch.count_nonzero(row[:last + 1])) > 0.9 * (last + 1)


def test_batched_rows_match_reference_restatement(batched, batched_run, oracle_built):
    w = batched
    step, out = batched_run
    n_s = 200_000
    for t in _sampled_rows(w["lengths"]):
        n = int(w["lengths"][t])
        st = np.nonzero((w["weights"][t] != 0).any(axis=-1))[0]
        _say("oracle t=%d N=%d on %d stations, first %d samples" % (t, n, st.size, n_s))
        ref = oracle_built.run(np.ascontiguousarray(w["templates"][t:t + 1, st, :, :n]),
                               np.ascontiguousarray(w["moveouts"][t:t + 1, st]),
                               np.ascontiguousarray(w["weights"][t:t + 1, st]),
                               np.ascontiguousarray(w["data"][st, :, :n_s]), step, lib="oracle")[0]
        # lags valid in the slice (moveouts are >= 0: the range starts at 0)
        stop = 1 + n_s - n - int(np.max(w["moveouts"][t][st]))
        m = (stop - 1) // step + 1
        got = out[t, :m].cpu().numpy()
        assert np.count_nonzero(ref[:m]) > 0.9 * m
        assert np.abs(got - ref[:m]).max() <= TOL_REF


def test_batched_fp64_at_events(batched, batched_run):
    from oracle import fp64
    w = batched
    step, out = batched_run
    for t in _sampled_rows(w["lengths"]):
        n = int(w["lengths"][t])
        outs = np.asarray(sorted({lag // step for tt, lag in w["events"] if tt == t}),
                          dtype=np.int64)
        got = out[t, outs].cpu().numpy() if outs.size else np.zeros(0)
        exact = fp64.cc_at(w["templates"], w["moveouts"], w["weights"], w["data"], t,
                           outs * step, n_tp=n)
        err = np.abs(got - exact).max() if outs.size else 0.0
        _say("batched step %d t=%d N=%d max |dcc| %.3g at %d events" % (step, t, n, err, outs.size))
        assert err <= TOL_FP64
