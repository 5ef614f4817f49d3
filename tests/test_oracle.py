"""Pins the oracle before anything is checked against it (CPU only).

* The C restatement reproduces the compiled reference's float32 bits on every
  golden case (the golden vectors were written by the reference's own
  matched_filter.c, compiled with its flags).
* The FP64 oracle agrees with brute-force loops on tiny inputs, with the
  self-detection identity, Cauchy-Schwarz, scale invariance and the
  DC-offset invariance of 'full', and with the reference to float32 rounding.
* A planted mistake (moveouts off by one sample) is caught by the same checks.
"""
import hashlib
import math

import numpy as np
import pytest

from oracle import fp64
from tests.golden import cases


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def inputs():
    return {name: build() for name, build in cases.CASES.items()}


def test_inputs_regenerate_identically(golden, inputs):
    for name, inp in inputs.items():
        assert cases.input_digest(inp) == str(golden[name + "__input_sha"]), name


def _variants():
    for name, vs in cases.VARIANTS.items():
        for key, opts in vs.items():
            yield name, key, opts


@pytest.mark.parametrize("name,key,opts", list(_variants()))
def test_restatement_bit_identical_to_reference(golden, inputs, oracle_built,
                                                name, key, opts):
    inp = inputs[name]
    opts = dict(opts)
    if opts.get("kind") == "variable_precise":
        opts["n_samples_template"] = inp["lengths"]
    out = oracle_built.run(inp["templates"], inp["moveouts"], inp["weights"],
                           inp["data"], inp["step"], lib="oracle", **opts)
    tag = name + "__" + key
    assert tuple(out.shape) == tuple(golden[tag + "__shape"])
    assert _sha(out) == str(golden[tag + "__sha"]), tag


def _brute(templates, moveouts, weights, data, step, normalize, network_sum):
    """Pure-Python loops straight from the definition (tiny inputs only)."""
    T, S, C, N = templates.shape
    L = data.shape[-1]
    n_corr = (L - N) // step + 1
    out = np.zeros((T, n_corr) if network_sum else (T, n_corr, S, C))
    for t in range(T):
        lo = min(0, int(moveouts[t].min()))
        hi = max(0, int(moveouts[t].max()))
        start = math.ceil(-lo / step) * step if lo < 0 else 0
        for i in range(start, 1 + L - N - hi, step):
            for s in range(S):
                for c in range(C):
                    if weights[t, s, c] == 0:
                        continue
                    win = [float(data[s, c, i + moveouts[t, s, c] + k]) for k in range(N)]
                    tp = [float(x) for x in templates[t, s, c]]
                    m = sum(win) / N if normalize == "full" else 0.0
                    num = sum(a * (b - m) for a, b in zip(tp, win))
                    den = sum(a * a for a in tp) * sum((b - m) ** 2 for b in win)
                    cc = num / math.sqrt(den) if den > 1e-6 else 0.0
                    if network_sum:
                        out[t, i // step] += weights[t, s, c] * cc
                    else:
                        out[t, i // step, s, c] = cc
    return out


@pytest.mark.parametrize("normalize", ["short", "full"])
@pytest.mark.parametrize("network_sum", [True, False])
@pytest.mark.parametrize("step", [1, 2])
def test_fp64_matches_brute_force(normalize, network_sum, step):
    rng = np.random.Generator(np.random.PCG64(5))
    T, S, C, N, L = 2, 2, 2, 5, 40
    tp = rng.standard_normal((T, S, C, N))
    d = rng.standard_normal((S, C, L)) + 0.3
    mo = rng.integers(-3, 6, size=(T, S, C))
    w = rng.uniform(0.1, 1.0, size=(T, S, C))
    w[1, 0, 1] = 0.0
    d[1, 1, 10:20] = 0.0
    a = fp64.matched_filter(tp, mo, w, d, step, normalize, network_sum)
    b = _brute(tp, mo, w, d, step, normalize, network_sum)
    assert np.abs(a - b).max() < 1e-12


def test_fp64_self_detection_and_bounds(inputs):
    inp = inputs["selfdet"]
    T, S, C, N = inp["templates"].shape
    mo = np.repeat(inp["moveouts"][:, :, None], C, axis=2)
    w = np.repeat(inp["weights"][:, :, None], C, axis=2)
    cc = fp64.matched_filter(inp["templates"], mo, w, inp["data"], inp["step"],
                             network_sum=False)
    assert np.abs(cc).max() <= 1.0 + 1e-12
    for t, lag in inp["events"]:
        np.testing.assert_allclose(cc[t, lag // inp["step"]], 1.0, atol=1e-12)


def test_fp64_invariances(inputs):
    inp = inputs["edge"]
    args = (inp["templates"], inp["moveouts"], inp["weights"])
    a = fp64.matched_filter(*args, inp["data"].astype(np.float64), 3)
    b = fp64.matched_filter(inp["templates"] * 0.25, inp["moveouts"], inp["weights"],
                            inp["data"].astype(np.float64) * 8.0, 3)
    assert np.abs(a - b).max() < 1e-12
    # full normalization does not see a constant offset of the data
    # (consequences_nonzero.pdf eq. (1)-(3): short is biased, full is not)
    d = inp["data"].astype(np.float64)
    f1 = fp64.matched_filter(*args, d, 3, normalize="full")
    f2 = fp64.matched_filter(*args, d + 5.0, 3, normalize="full")
    ok = (np.abs(f1) > 0) & (np.abs(f2) > 0)
    assert np.abs(f1 - f2)[ok].max() < 1e-9


@pytest.mark.parametrize("name,key,normalize", [
    ("edge", "cpu_sum", "short"), ("selfdet", "cpu_sum", "short"),
    ("noisy", "cpu_sum", "short"), ("edge", "precise_full_sum", "full"),
    ("noisy", "precise_full_sum", "full")])
def test_fp64_vs_reference(golden, inputs, name, key, normalize):
    inp = inputs[name]
    from oracle.cpu import prepare
    p = prepare(inp["templates"], inp["moveouts"], inp["weights"], inp["data"], inp["step"])
    tp = p["templates"].reshape(p["T"], p["S"], p["C"], -1)
    mo = p["moveouts"].reshape(p["T"], p["S"], p["C"])
    w = p["weights"].reshape(p["T"], p["S"], p["C"])
    d = p["data"].reshape(p["S"], p["C"], -1)
    f = fp64.matched_filter(tp, mo, w, d, inp["step"], normalize)
    ref = golden[name + "__" + key + "__full"]
    assert np.abs(f - ref).max() < 2e-5


def test_planted_moveout_error_is_caught(golden, inputs):
    """The same comparison fails for an oracle whose moveouts are off by one."""
    inp = inputs["edge"]
    f = fp64.matched_filter(inp["templates"], inp["moveouts"] + 1, inp["weights"],
                            inp["data"], 3)
    assert np.abs(f - golden["edge__cpu_sum__full"]).max() > 1e-2


def test_neumaier_prefix_exact_on_integers(oracle_built):
    """matched_filter.c:680-705 on integer-valued data: every prefix is an
    exact integer, so the running sum must equal the integer cumsum."""
    rng = np.random.Generator(np.random.PCG64(6))
    x = rng.integers(-300, 300, size=(3, 50_000)).astype(np.float32)
    x[1, 1000:3000] = 0
    P = oracle_built.cumsum_square_data(x)
    exact = np.concatenate([np.zeros((3, 1), dtype=np.int64),
                            np.cumsum(x.astype(np.int64) ** 2, axis=1)], axis=1)
    np.testing.assert_array_equal(P, exact.astype(np.float64))
    # constant over the zero gap
    assert np.all(P[1, 1001:3001] == P[1, 1000])


def test_neumaier_prefix_vs_fsum(oracle_built):
    rng = np.random.Generator(np.random.PCG64(7))
    x = rng.standard_normal((1, 20_000)).astype(np.float32) * 1e3
    P = oracle_built.cumsum_square_data(x)[0]
    sq = [float(v) * float(v) for v in x[0]]
    for j in (1, 17, 999, 12_345, 20_000):
        assert abs(P[j] - math.fsum(sq[:j])) <= 1e-15 * P[j] * 4
