"""Tile-shape autotuner for the bf16x3 MFMA kernel.

The kernel's shape trades the reuse of each template fragment against
registers and occupancy.  Two tile families: 16x16x32 MFMA tiles of 256 lags
(2, 4 or 8 per wave; coded 10 * tiles + waves per SIMD: 24, 42, 43, 81) and
32x32x16 MFMA tiles of 1024 lags (1, 2 or 4 per wave in blocks of 2 or 4
waves; coded 100 * waves per block + 10 * tiles + waves per SIMD: 412, 413,
421, 422, 241; one-wave blocks 121, 141; 1422: 422 with interleaved tiles that
share data reads).  The best choice depends on the template length, the
number of active channels and the step.  ``tune`` times each candidate on the caller's own device-resident
problem (or a slice of it) with HIP events, keeps the fastest and records it
in a JSON cache keyed by (step, length bucket, channels, device name), so
later processes start from the tuned shape without timing again.
"""
import json
import os
import sys
import time

from . import _lib

CANDIDATES = (43, 412, 413, 421, 422, 241, 1422)


def _cache_path():
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(base, "fast_matched_filter_amd", "mfma_tiles.json")


def _load_cache():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save_cache(cache):
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp%d" % os.getpid()
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # read-only home: tuning still applies to this process


def key(step, n_samples_template, n_active_channels):
    import torch
    name = torch.cuda.get_device_name() if torch.cuda.is_available() else "none"
    bucket = 1 << max(0, int(n_samples_template - 1).bit_length())
    # v3: the candidate set includes the interleaved-tile shape
    return "v3/step%d/N%d/ch%d/%s" % (step, bucket, n_active_channels, name)


def apply(step, tiles):
    _lib.check(_lib.require().fmf_set_mfma_tiles(int(step), int(tiles)), "fmf_set_mfma_tiles")


def tune(run, step, n_samples_template, n_active_channels, repeats=2,
         candidates=CANDIDATES, use_cache=True, verbose=False):
    """Pick the tiles-per-wave for ``run()`` (a callable that launches the
    MFMA kernel on the current stream) and apply it.  Returns
    (best_tiles, {tiles: ms})."""
    import torch
    k = key(step, n_samples_template, n_active_channels)
    cache = _load_cache() if use_cache else {}
    if k in cache:
        apply(step, cache[k]["tiles"])
        return cache[k]["tiles"], cache[k].get("ms", {})
    times = {}
    for g in candidates:
        apply(step, g)
        run()  # warm (code object load, LDS attribute)
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(repeats):
            run()
        stop.record()
        torch.cuda.synchronize()
        times[g] = start.elapsed_time(stop) / repeats
        if verbose:
            print("autotune %s tiles=%d: %.3f ms" % (k, g, times[g]), file=sys.stderr, flush=True)
    best = min(times, key=times.get)
    apply(step, best)
    cache[k] = {"tiles": best, "ms": {str(g): t for g, t in times.items()},
                "when": time.strftime("%Y-%m-%dT%H:%M:%S")}
    if use_cache:
        _save_cache(cache)
    return best, times
