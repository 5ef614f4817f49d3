"""The reference's self-detection harness, test_matched_filter
(fast_matched_filter.py:365-569), through this package on the GPU: templates
cut from random data give, at the lag they were cut from, a network cc of
n_stations * n_components (every channel's cc is 1) for 'short'
normalization; for 'full' each channel gives sqrt(sum (T - mean T)^2 /
sum T^2), since the reference divides by the template's own sum of squares,
not the demeaned one (matched_filter.c:629-657, matched_filter.cu:116-140);
and the largest value of the row is there for every arch, variable_precise
included."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cut_lags(templates, moveouts, data, step):
    """Lag (output index) each template was cut at: where the first channel's
    template equals the data exactly."""
    lags = []
    for t in range(templates.shape[0]):
        tp = templates[t, 0, 0]
        n = int(np.count_nonzero(tp)) or tp.size
        d = data[0, 0]
        w = np.lib.stride_tricks.sliding_window_view(d, n)
        hit = np.nonzero((w == tp[:n]).all(axis=1))[0]
        assert hit.size >= 1
        lags.append((int(hit[0]) - int(moveouts[t, 0, 0])) // step)
    return lags


@pytest.mark.parametrize("arch,normalize,step", [
    ("cpu", "short", 1), ("gpu", "full", 2), ("precise", "full", 1), ("hip", "short", 1),
    ("hip_mfma", "short", 1), ("hip_mfma", "full", 2)])
def test_self_detection(arch, normalize, step):
    import fast_matched_filter_amd as fmf
    S, C = 2, 3
    print("case: test_matched_filter arch=%s normalize=%s step=%d" % (arch, normalize, step),
          flush=True)
    tp, mo, data, st, cc, run_time = fmf.test_matched_filter(
        n_templates=3, n_stations=S, n_components=C, template_duration=2, data_duration=600,
        sampling_rate=100, step=step, arch=arch, normalize=normalize, seed=5,
        check_zeros=False)
    assert st == step and run_time >= 0.0
    assert cc.shape == (3, (60000 - 200) // step + 1)
    for t, lag in enumerate(_cut_lags(tp, mo, data, step)):
        x = tp[t].astype(np.float64)
        if normalize == "full":
            dm = x - x.mean(axis=-1, keepdims=True)
            expect = float(np.sqrt((dm * dm).sum(-1) / (x * x).sum(-1)).sum())
        else:
            expect = float(S * C)
        assert abs(cc[t, lag] - expect) <= 1e-4 * S * C
        assert int(np.argmax(cc[t])) == lag


def test_self_detection_variable_precise():
    import fast_matched_filter_amd as fmf
    print("case: test_matched_filter arch=variable_precise", flush=True)
    out = fmf.test_matched_filter(n_templates=2, n_stations=3, n_components=1,
                                  template_duration=2, data_duration=300, sampling_rate=100,
                                  arch="variable_precise", seed=6, check_zeros=False)
    tp, mo, data, st, cc, lengths, run_time = out
    assert lengths.shape == mo.shape and lengths.max() == tp.shape[-1]
    for t, lag in enumerate(_cut_lags(tp, mo, data, st)):
        assert int(np.argmax(cc[t])) == lag
