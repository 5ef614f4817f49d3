"""Deterministic inputs of the golden-vector cases.

Each builder returns dict(templates, moveouts, weights, data, step[, events]).
`reference_data/` holds the reference's own test data files
(fast_matched_filter/tests/data/noisy_*.npy, used by its tests/test_python.py:
26-32); everything else is generated from fixed seeds.
"""
import hashlib
import os

import numpy as np

from fast_matched_filter_amd.synthetic import make_workload

HERE = os.path.dirname(os.path.abspath(__file__))


def input_digest(inp):
    h = hashlib.sha256()
    for k in ("templates", "moveouts", "weights", "data"):
        h.update(np.ascontiguousarray(inp[k]).tobytes())
    h.update(str(inp["step"]).encode())
    return h.hexdigest()


def small():
    w = make_workload("small")
    return dict(templates=w["templates"], moveouts=w["moveouts"],
                weights=w["weights"], data=w["data"], step=1,
                events=w["events"])


def noisy():
    d = os.path.join(HERE, "reference_data")
    data = np.load(os.path.join(d, "noisy_data.npy"))
    templates = np.load(os.path.join(d, "noisy_template.npy"))
    pads = np.load(os.path.join(d, "noisy_pads.npy"))
    # tests/test_python.py:38-40 — weights 1/n_stations, shape (T, S)
    weights = np.ones((templates.shape[0], data.shape[0])) / data.shape[0]
    return dict(templates=templates, moveouts=pads, weights=weights, data=data,
                step=1)


def edge():
    """Negative moveouts, step 3, a weight-0 channel with the largest
    moveout, a zero gap, a DC offset and N = 37 (not a multiple of 4)."""
    rng = np.random.Generator(np.random.PCG64(11))
    T, S, C, N, L = 3, 3, 2, 37, 3000
    data = rng.standard_normal((S, C, L)).astype(np.float32)
    data[1, 0, 1000:1400] = 0.0
    data[2, 1] += np.float32(3.0)
    templates = rng.standard_normal((T, S, C, N)).astype(np.float32)
    moveouts = rng.integers(0, 120, size=(T, S, C)).astype(np.int32)
    moveouts[0, 0, 0] = -20
    moveouts[0, 1, 1] = -7
    moveouts[2, 2, 0] = 900
    weights = rng.uniform(0.05, 0.3, size=(T, S, C)).astype(np.float32)
    weights[2, 2, 0] = 0.0
    return dict(templates=templates, moveouts=moveouts, weights=weights,
                data=data, step=3)


def selfdet():
    """Templates cut from the data (step 2): cc = sum of weights there."""
    rng = np.random.Generator(np.random.PCG64(12))
    T, S, C, N, L = 2, 2, 3, 64, 5000
    data = rng.standard_normal((S, C, L)).astype(np.float32)
    moveouts = rng.integers(0, 200, size=(T, S)).astype(np.int32)
    cut = [1000, 3400]
    templates = np.zeros((T, S, C, N), dtype=np.float32)
    for t in range(T):
        for s in range(S):
            a = cut[t] + moveouts[t, s]
            templates[t, s] = data[s, :, a:a + N]
    weights = np.full((T, S), 1.0 / S, dtype=np.float32)
    return dict(templates=templates, moveouts=moveouts, weights=weights,
                data=data, step=2, events=[(0, cut[0]), (1, cut[1])])


def varlen():
    """Per-channel template lengths for arch 'variable_precise'."""
    rng = np.random.Generator(np.random.PCG64(13))
    T, S, C, L = 2, 3, 2, 4000
    lengths = rng.integers(30, 61, size=(T, S, C)).astype(np.int32)
    n_max = int(lengths.max())
    templates = np.zeros((T, S, C, n_max), dtype=np.float32)
    for t in range(T):
        for s in range(S):
            for c in range(C):
                templates[t, s, c, :lengths[t, s, c]] = rng.standard_normal(lengths[t, s, c])
    data = rng.standard_normal((S, C, L)).astype(np.float32)
    moveouts = rng.integers(0, 100, size=(T, S, C)).astype(np.int32)
    weights = rng.uniform(0.1, 1.0, size=(T, S, C)).astype(np.float32)
    weights[1, 0, 1] = 0.0
    return dict(templates=templates, moveouts=moveouts, weights=weights,
                data=data, step=1, lengths=lengths)


CASES = dict(small=small, noisy=noisy, edge=edge, selfdet=selfdet, varlen=varlen)

_ALL = dict(
    cpu_sum=dict(kind="cpu"),
    cpu_nosum=dict(kind="cpu", network_sum=False),
    precise_short_sum=dict(kind="precise"),
    precise_full_sum=dict(kind="precise", normalize="full"),
    precise_short_nosum=dict(kind="precise", network_sum=False),
    precise_full_nosum=dict(kind="precise", normalize="full", network_sum=False),
)

VARIANTS = dict(
    small=dict(cpu_sum=_ALL["cpu_sum"], precise_full_sum=_ALL["precise_full_sum"]),
    noisy=dict(cpu_sum=_ALL["cpu_sum"], precise_short_sum=_ALL["precise_short_sum"],
               precise_full_sum=_ALL["precise_full_sum"]),
    edge=_ALL,
    selfdet=_ALL,
    varlen=dict(varprec_short=dict(kind="variable_precise", normalize="short"),
                varprec_full=dict(kind="variable_precise", normalize="full")),
)
