"""Batched launch (templates of different lengths in one MFMA launch).

Pinned by (no reference counterpart; the reference takes one length per call,
matched_filter.c:24):
  * every row equals the single-length bf16x3 call on that template alone,
    bit for bit, and is zero outside its own valid lags (also with 13
    distinct lengths, more than one window-energy launch holds);
  * per length group, max |dcc| <= 1e-4 against the oracle's restatement of
    the reference 'cpu' arch;
  * the persistent launch equals the one-block-per-item launch;
  * a HIP graph of the call replays to the eager result (several replays,
    with eager work between them).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[1, 2])
def workload(request):
    from fast_matched_filter_amd.synthetic import make_workload
    w = make_workload("batched", n_templates=10, n_stations=6,
                      n_samples_data=40_000, stations=(2, 5), max_moveout=600,
                      seed=21 + request.param)
    w["step"] = request.param
    return w


def _batched(w, **kw):
    import fast_matched_filter_amd as fmf
    print("batched: T=%d step=%d lengths=%s" % (len(w["lengths"]), w["step"],
                                                sorted(set(w["lengths"].tolist()))), flush=True)
    return fmf.matched_filter_batched(w["templates"], w["lengths"], w["moveouts"],
                                      w["weights"], w["data"], w["step"], **kw)


def test_rows_equal_single_length_calls(workload):
    import fast_matched_filter_amd as fmf
    w = workload
    out = _batched(w)
    step = w["step"]
    L = w["data"].shape[-1]
    assert out.shape == (len(w["lengths"]), (L - int(w["lengths"].min())) // step + 1)
    for t, n in enumerate(w["lengths"]):
        single = fmf.matched_filter(w["templates"][t:t + 1, :, :, :n], w["moveouts"][t:t + 1],
                                    w["weights"][t:t + 1], w["data"], step,
                                    arch="hip_mfma", check_zeros=False)[0]
        k = single.size
        assert np.array_equal(out[t, :k], single), "template %d (N=%d)" % (t, n)
        assert not np.any(out[t, k:])


# 13 distinct lengths: the window-energy pass builds the tables of the lengths
# up to 1280 in launches of at most 8 (here 8 + 4) and gives 1281 a launch of
# its own; every table must hold the bits of a single-length call's table.
MANY_LENGTHS = (96, 200, 304, 400, 512, 600, 704, 800, 904, 1000, 1104, 1200, 1281)


@pytest.mark.parametrize("step", [1, 2])
def test_many_lengths_rows_equal_single_length_calls(step):
    import fast_matched_filter_amd as fmf
    from fast_matched_filter_amd.synthetic import make_workload
    w = make_workload("batched", n_templates=30, n_stations=4, n_samples_data=30_000,
                      stations=(2, 4), max_moveout=500, n_samples_template=MANY_LENGTHS,
                      seed=41)
    w["step"] = step
    assert set(w["lengths"].tolist()) == set(MANY_LENGTHS)
    out = _batched(w)
    for t, n in enumerate(w["lengths"]):
        single = fmf.matched_filter(w["templates"][t:t + 1, :, :, :n], w["moveouts"][t:t + 1],
                                    w["weights"][t:t + 1], w["data"], step,
                                    arch="hip_mfma", check_zeros=False)[0]
        k = single.size
        assert np.array_equal(out[t, :k], single), "template %d (N=%d)" % (t, n)
        assert not np.any(out[t, k:])


def test_rows_match_reference_per_length_group(workload, oracle_built):
    w = workload
    out = _batched(w)
    for n in sorted(set(w["lengths"].tolist())):
        rows = np.nonzero(w["lengths"] == n)[0]
        ref = oracle_built.run(w["templates"][rows][:, :, :, :n], w["moveouts"][rows],
                               w["weights"][rows], w["data"], w["step"], lib="oracle")
        assert np.abs(out[rows, :ref.shape[1]] - ref).max() <= 1e-4


def test_persistent_launch_equals_grid_launch(workload):
    a = _batched(workload)
    b = _batched(workload, persistent_blocks=24)
    assert np.array_equal(a, b)


def test_graph_replay_matches_eager(workload):
    import torch
    from fast_matched_filter_amd.device import matched_filter_batched_device
    from fast_matched_filter_amd.fast_matched_filter import prepare_inputs
    w = workload
    p = prepare_inputs(w["templates"], w["moveouts"], w["weights"], w["data"], w["step"])
    dev = torch.device("cuda:0")
    to = lambda a: torch.from_numpy(a).to(dev)
    args = [to(p["templates"]), to(p["sst"]), to(p["moveouts"]), to(p["weights"]),
            to(p["data"]), to(w["lengths"].astype(np.int32))]
    distinct = np.unique(w["lengths"])
    n_corr = (p["L"] - int(w["lengths"].min())) // w["step"] + 1
    sizes = (w["step"], p["N"], p["L"], p["T"], p["S"], p["C"], n_corr)
    eager = torch.empty((p["T"], n_corr), device=dev)
    matched_filter_batched_device(*args, distinct, *sizes, eager)
    torch.cuda.synchronize()
    from fast_matched_filter_amd.device import batched_workspace_bytes
    ws = torch.empty(batched_workspace_bytes(distinct, p["N"], p["L"], p["T"], p["S"], p["C"]),
                     dtype=torch.uint8, device=dev)
    out = torch.zeros((p["T"], n_corr), device=dev)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):  # warm up on the capture stream
        matched_filter_batched_device(*args, distinct, *sizes, out, workspace=ws, stream=s)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        matched_filter_batched_device(*args, distinct, *sizes, out, workspace=ws,
                                      stream=torch.cuda.current_stream())
    for _ in range(3):
        out.fill_(7.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager)


def test_autotuner_picks_a_candidate(workload):
    import torch
    from fast_matched_filter_amd import autotune
    import fast_matched_filter_amd as fmf
    from fast_matched_filter_amd import _lib
    w = workload
    shape_before = int(_lib.require().fmf_get_mfma_tiles(w["step"]))
    try:
        best, times = autotune.tune(lambda: _batched(w, return_device=True), w["step"],
                                    int(w["lengths"].max()), 18, repeats=1, use_cache=False)
        assert best in autotune.CANDIDATES and len(times) == len(autotune.CANDIDATES)
        torch.cuda.synchronize()
        # results do not depend on the tile shape beyond FP32 summation order
        a = _batched(w)
        autotune.apply(w["step"], 24)
        b = _batched(w)
        assert np.abs(a - b).max() <= 1e-5
    finally:  # the tile shape is process-wide: leave it as it was
        autotune.apply(w["step"], shape_before)
