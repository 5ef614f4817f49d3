"""Pins the oracle's restatement of the wrapper's input normalisation
(oracle.cpu.prepare, fast_matched_filter.py:180-274) without retyping it:

* the reference's own format equivalences (tests/test_python.py:135-153:
  ``test_single``, ``test_format``) on the C restatement: the single-template
  and the traces forms give the outputs of the (T, S, C, N) form;
* the (T, S, C, N) form with (T, S) float moveouts and weights against the
  FP64 oracle fed the raw 4D arrays (moveouts truncated to int32 and repeated
  over components by numpy, outside prepare), which never goes through
  prepare;
* planted mistakes in prepare (components repeated in the wrong order, the
  sum of squares taken over the wrong axis, moveouts rounded instead of
  truncated) make those comparisons fail.
"""
import numpy as np
import pytest

from oracle import cpu, fp64

T, S, C, N, L = 3, 4, 2, 24, 600


def _inputs(seed=21):
    rng = np.random.Generator(np.random.PCG64(seed))
    tp = rng.standard_normal((T, S, C, N)).astype(np.float32)
    d = rng.standard_normal((S, C, L)).astype(np.float32)
    mo = (rng.integers(0, 40, size=(T, S)) + 0.7).astype(np.float32)  # float pads, truncated
    w = rng.uniform(0.1, 0.3, size=(T, S)).astype(np.float32)
    w[1, 2] = 0.0
    return tp, mo, w, d


def _forms(tp, mo, w, d):
    traces = (tp.reshape(T, S * C, N), np.repeat(mo, C, axis=1), np.repeat(w, C, axis=1),
              d.reshape(S * C, L))
    single = (tp[0], mo[0], w[0], d)
    single_traces = (tp[0].reshape(S * C, N), np.repeat(mo[0], C), np.repeat(w[0], C),
                     d.reshape(S * C, L))
    return {"4d": (tp, mo, w, d), "traces": traces, "single": single,
            "single_traces": single_traces}


@pytest.mark.parametrize("step", [1, 2])
def test_formats_agree_on_the_restatement(oracle_built, step):
    f = _forms(*_inputs())
    full = cpu.run(*f["4d"], step)
    assert full.shape == (T, (L - N) // step + 1) and np.abs(full).max() > 0.1
    np.testing.assert_array_equal(cpu.run(*f["traces"], step), full)
    np.testing.assert_array_equal(cpu.run(*f["single"], step).reshape(-1), full[0])
    np.testing.assert_array_equal(cpu.run(*f["single_traces"], step).reshape(-1), full[0])


@pytest.mark.parametrize("normalize,kind", [("short", "cpu"), ("full", "precise")])
def test_4d_form_matches_fp64_on_raw_arrays(oracle_built, normalize, kind):
    tp, mo, w, d = _inputs()
    got = cpu.run(tp, mo, w, d, 1, kind=kind, normalize=normalize)
    mo3 = np.repeat(mo.astype(np.int32)[:, :, None], C, axis=2)
    w3 = np.repeat(w[:, :, None], C, axis=2)
    exact = fp64.matched_filter(tp, mo3, w3, d, 1, normalize)
    assert np.abs(got - exact).max() < 2e-5


def test_planted_prepare_mistakes_are_caught(oracle_built, monkeypatch):
    tp, mo, w, d = _inputs()
    mo3 = np.repeat(mo.astype(np.int32)[:, :, None], C, axis=2)
    w3 = np.repeat(w[:, :, None], C, axis=2)
    exact = fp64.matched_filter(tp, mo3, w3, d, 1)
    good = cpu.prepare

    def tiled(*a, **k):  # components repeated as tiles (np.tile) instead of per entry
        p = good(*a, **k)
        if np.asarray(a[1]).size * p["C"] == p["T"] * p["S"] * p["C"]:
            p["moveouts"] = np.ascontiguousarray(
                np.tile(np.asarray(a[1]).reshape(p["T"], p["S"]), (1, p["C"])).reshape(-1),
                dtype=np.int32)
        return p

    def sst_axis(*a, **k):  # sum of squares over stations instead of samples
        p = good(*a, **k)
        t4 = p["templates"].reshape(p["T"], p["S"], p["C"], -1)
        p["sst"] = np.ascontiguousarray(
            np.repeat(np.sum(t4 ** 2, axis=1)[:, None], p["S"], axis=1).reshape(-1),
            dtype=np.float32)
        return p

    def rounded(*a, **k):  # moveouts rounded instead of truncated
        a = list(a)
        a[1] = np.round(np.asarray(a[1]))
        return good(*a, **k)

    for bad in (tiled, sst_axis, rounded):
        monkeypatch.setattr(cpu, "prepare", bad)
        got = cpu.run(tp, mo, w, d, 1)
        assert np.abs(got - exact).max() > 1e-3, bad.__name__
    monkeypatch.setattr(cpu, "prepare", good)
    assert np.abs(cpu.run(tp, mo, w, d, 1) - exact).max() < 2e-5
