"""Seeded synthetic seismic workloads for the matched-filter engine.

Every workload is a pure function of its name, its sizes and a seed (NumPy
PCG64 streams), so any host regenerates byte-identical inputs.  The reference
ships only an unseeded generator (fast_matched_filter/fast_matched_filter.py:
365-569, ``test_matched_filter``); these follow its idea (templates embedded in
noise so that the injected times are known detections) with fixed seeds.

Workloads (T templates, S stations, C components, N template samples,
L data samples):

* ``small``   — T=10, S=4, C=3, N=200, L=100 000; N(0,1) noise; Hann-tapered
  N(0,1) templates injected 3x each at amplitude 0.5; moveouts U[0,300) per
  (template, station); weights 1/12; step 1.
* ``day``     — T=100, S=20, C=3, N=800, L=8 640 000 (one day at 100 Hz);
  N(0,1) noise with 5 % of 6000-sample windows zeroed per channel (gaps);
  20 injections per template at SNR log-uniform in [0.2, 2]; moveouts
  U[0,2000); weights 1/60; step 1.
* ``day500``  — ``day`` with T=500.
* ``batched`` — T=2000 over S=40, C=3, L=8 640 000; each template uses a
  random subset of 8-30 stations (weight 0 elsewhere) and a length drawn from
  {400, 600, 800, 1200}; moveouts U[0,3000); templates zero-padded to 1200.

Sizes can be overridden (e.g. ``n_samples_data=200_000``) to build reduced
versions of the same workload for tests.
"""
import numpy as np

__all__ = ["make_workload", "WORKLOADS"]

WORKLOADS = {
    "small": dict(n_templates=10, n_stations=4, n_components=3,
                  n_samples_template=200, n_samples_data=100_000,
                  n_events=3, snr=(0.5, 0.5), max_moveout=300,
                  gap_fraction=0.0, seed=1),
    "day": dict(n_templates=100, n_stations=20, n_components=3,
                n_samples_template=800, n_samples_data=8_640_000,
                n_events=20, snr=(0.2, 2.0), max_moveout=2000,
                gap_fraction=0.05, seed=2),
    "day500": dict(n_templates=500, n_stations=20, n_components=3,
                   n_samples_template=800, n_samples_data=8_640_000,
                   n_events=20, snr=(0.2, 2.0), max_moveout=2000,
                   gap_fraction=0.05, seed=3),
    "batched": dict(n_templates=2000, n_stations=40, n_components=3,
                    n_samples_template=(400, 600, 800, 1200),
                    n_samples_data=8_640_000, n_events=4, snr=(0.2, 2.0),
                    max_moveout=3000, gap_fraction=0.0, stations=(8, 30),
                    seed=4),
}

GAP_WINDOW = 6000


def _hann(n):
    if n == 1:
        return np.ones(1)
    return 0.5 - 0.5 * np.cos(2.0 * np.pi * np.arange(n) / (n - 1))


def make_workload(name, **overrides):
    """Build a workload dict with keys

    templates (T,S,C,Nmax) f32, moveouts (T,S,C) i32, weights (T,S,C) f32,
    data (S,C,L) f32, step, lengths (T,) i32 (per-template N),
    events list of (template, sample lag) injected, name, params.
    """
    p = dict(WORKLOADS[name])
    p.update(overrides)
    rng = np.random.Generator(np.random.PCG64(p["seed"]))
    T, S, C = p["n_templates"], p["n_stations"], p["n_components"]
    L = p["n_samples_data"]
    if isinstance(p["n_samples_template"], (tuple, list)):
        lengths = rng.choice(np.asarray(p["n_samples_template"]), size=T).astype(np.int32)
    else:
        lengths = np.full(T, p["n_samples_template"], dtype=np.int32)
    n_max = int(lengths.max())

    data = rng.standard_normal((S, C, L), dtype=np.float32)
    if p["gap_fraction"] > 0:
        n_win = L // GAP_WINDOW
        for s in range(S):
            for c in range(C):
                gaps = rng.random(n_win) < p["gap_fraction"]
                for g in np.nonzero(gaps)[0]:
                    data[s, c, g * GAP_WINDOW:(g + 1) * GAP_WINDOW] = 0.0

    templates = np.zeros((T, S, C, n_max), dtype=np.float32)
    for t in range(T):
        n = int(lengths[t])
        tp = rng.standard_normal((S, C, n)) * _hann(n)
        tp /= np.sqrt(np.mean(tp * tp, axis=-1, keepdims=True))
        templates[t, :, :, :n] = tp.astype(np.float32)

    mo_station = rng.integers(0, p["max_moveout"], size=(T, S), dtype=np.int32)
    moveouts = np.repeat(mo_station[:, :, None], C, axis=2).astype(np.int32)

    if "stations" in p:
        lo, hi = p["stations"]
        weights = np.zeros((T, S, C), dtype=np.float32)
        for t in range(T):
            k = int(rng.integers(lo, hi + 1))
            st = rng.choice(S, size=k, replace=False)
            weights[t, st, :] = 1.0 / (k * C)
    else:
        weights = np.full((T, S, C), 1.0 / (S * C), dtype=np.float32)

    events = []
    lo_snr, hi_snr = p["snr"]
    for t in range(T):
        n = int(lengths[t])
        span = n + int(moveouts[t].max())
        limit = L - span - 1
        placed = []
        tries = 0
        while len(placed) < p["n_events"] and tries < 10_000:
            tries += 1
            lag = int(rng.integers(0, limit))
            if all(abs(lag - q) > span for q in placed):
                placed.append(lag)
        for lag in placed:
            if lo_snr == hi_snr:
                amp = lo_snr
            else:
                amp = float(np.exp(rng.uniform(np.log(lo_snr), np.log(hi_snr))))
            for s in range(S):
                for c in range(C):
                    if weights[t, s, c] == 0.0:
                        continue
                    at = lag + int(moveouts[t, s, c])
                    data[s, c, at:at + n] += np.float32(amp) * templates[t, s, c, :n]
            events.append((t, lag))
    return dict(templates=templates, moveouts=moveouts, weights=weights,
                data=data, step=int(p.get("step", 1)), lengths=lengths,
                events=events, name=name, params=p)
