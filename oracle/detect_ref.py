"""ORACLE — test infrastructure only.

NumPy restatement of the detection post-processing (no reference
counterpart: the reference returns cc rows, fast_matched_filter.py:318-362):

* median and MAD per row with ``numpy.median`` in float32 over the row's valid
  lags [lo, hi);
* threshold = median + n_mad * MAD in float32;
* lag j is a detection when x[j] > threshold and j is the first argmax of x
  over [max(lo, j - s), min(hi - 1, j + s)].

``detect_bruteforce`` is the pure-Python loop of that definition (pins the
vectorised ``detect`` on small inputs in tests/test_detect_oracle.py).
"""
import numpy as np


def median_mad(x):
    x = np.asarray(x, dtype=np.float32)
    if x.size == 0:
        return np.float32(0), np.float32(0)
    med = np.float32(np.median(x))
    mad = np.float32(np.median(np.abs(x - med)))
    return med, mad


def threshold(med, mad, n_mad):
    return np.float32(np.float32(med) + np.float32(n_mad) * np.float32(mad))


def _first_max_flags(x, lo, hi, s, thr):
    """Vectorised: flags of lags in [lo, hi) that pass the definition."""
    v = np.asarray(x[lo:hi], dtype=np.float32)
    n = v.size
    flags = np.zeros(n, dtype=bool)
    cand = np.nonzero(v > thr)[0]
    for j in cand:
        a = max(0, j - s)
        b = min(n - 1, j + s)
        w = v[a:b + 1]
        flags[j] = (a + int(np.argmax(w))) == j
    return flags


def detect(cc, n_mad, s, ranges=None):
    """Returns (template, lag, value, median, mad, threshold) arrays."""
    cc = np.asarray(cc, dtype=np.float32)
    T, n_corr = cc.shape
    tt, ll, vv, meds, mads, thrs = [], [], [], [], [], []
    for t in range(T):
        lo, hi = (0, n_corr) if ranges is None else (int(ranges[t][0]), int(ranges[t][1]))
        med, mad = median_mad(cc[t, lo:hi])
        thr = threshold(med, mad, n_mad)
        f = _first_max_flags(cc[t], lo, hi, s, thr)
        idx = np.nonzero(f)[0] + lo
        tt.append(np.full(idx.size, t, np.int32))
        ll.append(idx.astype(np.int32))
        vv.append(cc[t, idx])
        meds.append(med); mads.append(mad); thrs.append(thr)
    return (np.concatenate(tt), np.concatenate(ll), np.concatenate(vv),
            np.array(meds, np.float32), np.array(mads, np.float32), np.array(thrs, np.float32))


def detect_bruteforce(row, n_mad, s, lo, hi):
    """Pure-Python loops over one row (median by sorting)."""
    x = [float(np.float32(v)) for v in row[lo:hi]]
    n = len(x)
    srt = sorted(x)
    med = srt[n // 2] if n % 2 else float(np.float32((np.float32(srt[n // 2 - 1]) + np.float32(srt[n // 2])) / np.float32(2)))
    dev = sorted(float(np.float32(abs(np.float32(v) - np.float32(med)))) for v in x)
    mad = dev[n // 2] if n % 2 else float(np.float32((np.float32(dev[n // 2 - 1]) + np.float32(dev[n // 2])) / np.float32(2)))
    thr = float(np.float32(np.float32(med) + np.float32(n_mad) * np.float32(mad)))
    out = []
    for j in range(n):
        if not x[j] > thr:
            continue
        ok = True
        for i in range(max(0, j - s), min(n - 1, j + s) + 1):
            if (i < j and not x[j] > x[i]) or (i > j and x[i] > x[j]):
                ok = False
                break
        if ok:
            out.append(j + lo)
    return out, med, mad, thr
