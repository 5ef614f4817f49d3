"""ORACLE — test infrastructure only.

FP64 NumPy restatement of the matched-filter mathematics
(fast_matched_filter/src/matched_filter.c:416-455 network sum, :611-627 short
normalization, :629-657 full normalization), evaluated exactly in float64:

    cc_ch[i] = sum_k T[k] (d[i+mo+k] - m) / sqrt(sum_k T[k]^2 * sum_k (d[i+mo+k] - m)^2)
    cc_sum[i] = sum_{ch: w != 0} w_ch * cc_ch[i]

with m = 0 ('short') or the window mean ('full'), cc_ch = 0 when the
denominator product is not above 1e-6 (matched_filter.c:17,624), the lag range
of matched_filter.c:53-79 and zeros elsewhere.

Parity pins (tests/test_oracle.py): brute-force loops on tiny inputs,
self-detection (cc == sum of weights where the template was cut from the data),
Cauchy-Schwarz |cc_ch| <= 1, scale invariance, and agreement with the C
restatement and the compiled reference to float32 rounding.
"""
import numpy as np

THRESHOLD = 1e-6


def lag_range(mo_t, step, n_tp, n_data):
    """matched_filter.c:53-76 — returns (start, stop) sample lags, stop
    exclusive; (0, 0) when the template does not fit."""
    lo = min(0, int(np.min(mo_t)))
    hi = max(0, int(np.max(mo_t)))
    start = int(np.ceil(-lo / step)) * step if lo < 0 else 0
    stop = 1 + n_data - n_tp - hi
    if stop <= 0:
        return 0, 0
    return start, stop


def _window_sums(x, n):
    c = np.concatenate([[0.0], np.cumsum(x, dtype=np.float64)])
    return c[n:] - c[:-n]


def _corr_valid(d, t):
    """sum_k t[k] d[p+k] for every p with p+len(t) <= len(d), in float64."""
    n = len(t)
    m = len(d) - n + 1
    if m <= 0:
        return np.zeros(0)
    if n * m <= 4_000_000:
        w = np.lib.stride_tricks.sliding_window_view(d, n)
        return w @ t
    size = 1 << int(np.ceil(np.log2(len(d) + n)))
    f = np.fft.rfft(d, size) * np.conj(np.fft.rfft(t, size))
    return np.fft.irfft(f, size)[:m]


def channel_cc(t, d, normalize="short"):
    """Per-position normalized correlation of one channel (float64), for every
    start p of a full window: returns (cc[p], den[p])."""
    t = np.asarray(t, dtype=np.float64)
    d = np.asarray(d, dtype=np.float64)
    n = len(t)
    num = _corr_valid(d, t)
    ss = _window_sums(d * d, n)
    sst = float(np.dot(t, t))
    if normalize == "full":
        s1 = _window_sums(d, n)
        num = num - s1 / n * t.sum()
        ss = ss - s1 * s1 / n
        ss = np.maximum(ss, 0.0)
    den = sst * ss
    cc = np.zeros_like(num)
    ok = den > THRESHOLD
    cc[ok] = num[ok] / np.sqrt(den[ok])
    return cc, den


def matched_filter(templates, moveouts, weights, data, step, normalize="short",
                   network_sum=True):
    """Full FP64 oracle over every lag.  Inputs in the reference's 4D/3D
    layouts: templates (T,S,C,N), moveouts/weights (T,S,C), data (S,C,L).
    Output (T, n_corr) or (T, n_corr, S, C) like the reference wrapper."""
    templates = np.asarray(templates)
    T, S, C, N = templates.shape
    L = data.shape[-1]
    n_corr = (L - N) // step + 1
    data = np.asarray(data, dtype=np.float64).reshape(S * C, L)
    tp = templates.reshape(T, S * C, N)
    mo = np.asarray(moveouts).reshape(T, S * C)
    w = np.asarray(weights, dtype=np.float64).reshape(T, S * C)
    if network_sum:
        out = np.zeros((T, n_corr))
    else:
        out = np.zeros((T, n_corr, S * C))
    for t in range(T):
        start, stop = lag_range(mo[t], step, N, L)
        lags = np.arange(start, stop, step)
        if lags.size == 0:
            continue
        o = lags // step
        for ch in range(S * C):
            if w[t, ch] == 0.0:
                continue
            cc, _ = channel_cc(tp[t, ch], data[ch], normalize)
            vals = cc[lags + mo[t, ch]]
            if network_sum:
                out[t, o] += w[t, ch] * vals
            else:
                out[t, o, ch] = vals
    if not network_sum:
        out = out.reshape(T, n_corr, S, C)
    return out


def cc_at(templates, moveouts, weights, data, t, lags, normalize="short",
          n_tp=None):
    """FP64 network cc of template t at the given sample lags (direct dot
    products; for spot checks at full size).  ``n_tp`` overrides the template
    length (templates zero-padded beyond it)."""
    T, S, C, N = templates.shape
    N = N if n_tp is None else int(n_tp)
    L = data.shape[-1]
    lags = np.asarray(lags, dtype=np.int64)
    res = np.zeros(lags.size)
    tp = templates.reshape(T, S * C, -1)[t, :, :N].astype(np.float64)
    mo = np.asarray(moveouts).reshape(T, S * C)[t]
    w = np.asarray(weights, dtype=np.float64).reshape(T, S * C)[t]
    d2 = data.reshape(S * C, L)
    idx = np.arange(N)
    for ch in range(S * C):
        if w[ch] == 0.0:
            continue
        win = d2[ch][(lags + mo[ch])[:, None] + idx[None, :]].astype(np.float64)
        if normalize == "full":
            win = win - win.mean(axis=1, keepdims=True)
        num = win @ tp[ch]
        den = float(tp[ch] @ tp[ch]) * np.einsum("ij,ij->i", win, win)
        cc = np.where(den > THRESHOLD, num / np.sqrt(np.maximum(den, 1e-300)), 0.0)
        res += w[ch] * cc
    return res
