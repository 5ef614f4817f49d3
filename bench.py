"""Matched-filter throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload batched]
                    [--kernel fp32|bf16x3] [--no-cpu-baseline]

Default workload (``batched``): 2000 templates over a 40-station,
3-component, one-day 100 Hz array (40, 3, 8 640 000); each template uses 8-30
stations (weight 0 elsewhere) and a length drawn from {400, 600, 800, 1200};
moveouts U[0, 3000); step 1; all templates in one bf16x3 MFMA launch.  A step
is one pass of the hot path (window energies per distinct length, template
preparation, network correlation of every template over the whole day) with
all inputs already resident in HBM, through libfmf_hip.so's device entry
points.  Prints one JSON line:

* template*channel*lag MACs/s of the whole step (eager launches bracketed by
  device synchronisation), the dominant kernel's roofline position (its
  duration from HIP events bracketing each of its launches on the stream it
  runs on, over the timed steps) and its HBM traffic per launch (FETCH_SIZE
  and WRITE_SIZE from rocprofv3 runs of this script on the same workload,
  scaled by those of a 1 GiB device copy);
* the max and RMS |dcc| against FP64 dot products at the injected events and
  random lags of sampled templates of every length group, and the lag-argmax
  agreement with FP64 at those events;
* the plain FP32 kernel on the same workload (one launch per length group)
  and the bf16x3 / FP32 throughput ratio;
* HIP-graph replays of the step, and detection (median/MAD, peaks) on its
  output;
* beside the batched headline, the single-length day (T=100, N=800, 20
  stations): its exact scan, and the learned screener on it with its event
  recall and its recall of the exact scan's detection list;
* the oracle's C restatement of the reference CPU path (OpenMP, every host
  core) on a bounded sample: one template of each length over part of the
  day, with the CPU model and thread count.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "template·channel·lag MACs/s, % of FP32/MFMA peak; max |Δcc| vs FP64 oracle"
PEAK_TFLOPS = {"fp32": 157.3, "bf16x3": 2500.0}  # MI355X published dense peaks
KERNELS = {"fp32": 0, "bf16x3": 1}
THRESHOLD = 1e-6  # the reference's stability threshold on sum(T^2) * sum(d^2)

# SHA-256 of the generated inputs (templates, moveouts, weights, data,
# lengths; in that order, raw bytes) of the default workloads, computed where
# the fixed seeds were chosen; the bench regenerates them and compares.
INPUT_SHA256 = {
    "batched": "aee5c3aa77e967edd558419682e16ab43a7e8a659a27b3b9b94c84b3a83b04df",
    "day": "aaa6d6b714228e89226837eca73ac2f391152b448482524ddd411264c15b3272",
    "day500": "a0a126800809e127c1bbe9a84fefac3c1ad702f579cf2524babd8ef9c835ba64",
    "small": "fa8a1a5a141079a3eca80dd882c6c6175ab8752c0372182a22998c2b04dc0087",
}


def input_sha256(w):
    import hashlib
    h = hashlib.sha256()
    for k in ("templates", "moveouts", "weights", "data", "lengths"):
        h.update(memoryview(np.ascontiguousarray(w[k])).cast("B"))
    return h.hexdigest()


# HBM bytes per correlation launch recorded with rocprofv3 --pmc FETCH_SIZE
# and WRITE_SIZE on these exact configurations (workload, kernel, tile shape,
# step), each scaled by what the counters read for a device copy of 1 GiB
# (FETCH_SIZE x2.00, WRITE_SIZE x1.00); summaries and commands in
# profiles/hbm_traffic_*.md.  The line measures the same counters in its own
# run (measure_traffic) and reports these figures beside that measurement, or
# in its place when the profiler is missing or --no-traffic is given; other
# configurations then report null.
MEASURED_TRAFFIC = {
    ("day", "bf16x3", 1422, 1): 8.237e10 + 3.464e9,
    ("batched", "bf16x3", 1422, 1): 5.093e12 + 6.997e10,
}


def lag_counts(moveouts, step, N, L, n_corr):
    """Computed lags per template: lag range rule of the reference
    (matched_filter.c:53-79), output index clipped to n_corr.  N may be one
    length or one per template.  Returns (count, first output, end output)."""
    N = np.asarray(N, dtype=np.int64)
    mo = moveouts.reshape(moveouts.shape[0], -1)
    lo = np.minimum(mo.min(axis=1), 0).astype(np.int64)
    hi = np.maximum(mo.max(axis=1), 0).astype(np.int64)
    start = np.where(lo < 0, (-lo + step - 1) // step * step, 0)
    stop = 1 + L - N - hi
    olo = start // step
    ohi = np.where(stop > 0, (stop + step - 1) // step, 0)
    ohi = np.minimum(ohi, n_corr)
    return np.maximum(ohi - olo, 0), olo, ohi


def fp64_cc(templates, moveouts, weights, data, t, outs, step, N):
    """FP64 network cc of template t at output indices `outs` (sample lag
    o * step), template length N (taps past N are ignored), short
    normalization with the reference threshold (matched_filter.c:611-627,
    :416-455), from direct dot products."""
    T = templates.shape[0]
    L = data.shape[-1]
    SC = data.size // L
    d = data.reshape(SC, L)
    tp = templates.reshape(T, SC, -1)[t, :, :N].astype(np.float64)
    mo = moveouts.reshape(T, SC)[t].astype(np.int64)
    wt = weights.reshape(T, SC)[t].astype(np.float64)
    outs = np.asarray(outs, dtype=np.int64)
    idx = np.arange(N)
    res = np.zeros(outs.size)
    for ch in range(SC):
        if wt[ch] == 0:
            continue
        win = d[ch][(outs * step + mo[ch])[:, None] + idx[None, :]].astype(np.float64)
        num = win @ tp[ch]
        den = float(tp[ch] @ tp[ch]) * np.einsum("ij,ij->i", win, win)
        res += wt[ch] * np.where(den > THRESHOLD, num / np.sqrt(np.maximum(den, 1e-300)), 0.0)
    return res


def spot_templates(lengths, per_group=2):
    """Templates of the FP64 check: the first and last of every length group
    (and the last template, whose output row lies furthest into memory)."""
    sel = set()
    for n in np.unique(lengths):
        rows = np.nonzero(lengths == n)[0]
        sel.update(int(r) for r in rows[:1])
        sel.update(int(r) for r in rows[-(per_group - 1):] if per_group > 1)
    sel.add(len(lengths) - 1)
    return sorted(sel)


def fp64_spot_check(w, out_dev, step, lengths, rng, n_random=24, half=32):
    """Max and RMS |dcc| against FP64 at the injected events and random lags
    of the spot templates, and the lag-argmax agreement: at each injected
    event, the lag of the largest cc within +-`half` output lags, GPU
    against FP64."""
    import torch
    T = w["templates"].shape[0]
    L = w["data"].shape[-1]
    n_corr = out_dev.shape[1]
    _, olo, ohi = lag_counts(w["moveouts"], step, lengths, L, n_corr)
    args = (w["templates"], w["moveouts"], w["weights"], w["data"])
    errs = []
    agree = n_ev = 0
    for t in spot_templates(lengths):
        N = int(lengths[t])
        ev = [lag // step for (tt, lag) in w["events"] if tt == t]
        outs = np.unique(np.concatenate([np.asarray(ev, dtype=np.int64),
                                         rng.integers(olo[t], ohi[t], size=n_random)]))
        outs = outs[(outs >= olo[t]) & (outs < ohi[t])]
        got = out_dev[t, torch.as_tensor(outs, device=out_dev.device)].cpu().numpy()
        errs.append(np.abs(got - fp64_cc(*args, t, outs, step, N)))
        for e in ev:
            win = np.arange(max(olo[t], e - half), min(ohi[t], e + half + 1))
            if win.size == 0:
                continue
            g = out_dev[t, torch.as_tensor(win, device=out_dev.device)].cpu().numpy()
            n_ev += 1
            agree += int(np.argmax(g) == np.argmax(fp64_cc(*args, t, win, step, N)))
    e = np.concatenate(errs)
    return float(e.max()), float(np.sqrt(np.mean(e * e))), int(e.size), agree, n_ev


def cpu_model():
    """Model name of the host CPU with its family/model numbers and widest
    vector extension (a virtualised host may report only the brand)."""
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if not line.strip():
                    break  # first processor only
                k, _, v = line.partition(":")
                info[k.strip()] = v.strip()
    except OSError:
        return "unknown"
    flags = info.get("flags", "").split()
    isa = "avx512f" if "avx512f" in flags else "avx2" if "avx2" in flags else "sse"
    return "%s (family %s model %s, %s, %d CPUs usable)" % (
        info.get("model name", "unknown"), info.get("cpu family", "?"), info.get("model", "?"),
        isa, len(os.sched_getaffinity(0)))


def cpu_baseline(w, lengths, step, n_samples):
    """The oracle's C restatement of the reference 'cpu' path
    (matched_filter.c:22-96, 416-455, 611-627, 660-705; OpenMP over lags, every
    host core) on a bounded sample: for each length group its first template,
    restricted to its active stations, over the first `n_samples` samples of
    the day.  Returns the line's cpu_baseline dict and, per sampled template,
    (t, its cpu output row)."""
    from oracle import cpu
    cores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    threads = int(os.environ["OMP_NUM_THREADS"])
    tot_macs = tot_s = 0.0
    groups, rows = {}, []
    for n in np.unique(lengths):
        t = int(np.nonzero(lengths == n)[0][0])
        st = np.nonzero((w["weights"][t] != 0).any(axis=-1))[0]
        tp = np.ascontiguousarray(w["templates"][t:t + 1, st, :, :n])
        mo = np.ascontiguousarray(w["moveouts"][t:t + 1, st])
        wt = np.ascontiguousarray(w["weights"][t:t + 1, st])
        d = np.ascontiguousarray(w["data"][st, :, :n_samples])
        t0 = time.perf_counter()
        out = cpu.run(tp, mo, wt, d, step, lib="oracle")
        dt = time.perf_counter() - t0
        n_corr = out.shape[1]
        lags, _, _ = lag_counts(mo, step, int(n), n_samples, n_corr)
        macs = float(lags[0] * (wt != 0).sum() * int(n))
        groups[str(int(n))] = {"template": t, "channels": int((wt != 0).sum()),
                               "lags": int(lags[0]), "MAC_per_s": macs / dt, "s": dt}
        tot_macs += macs
        tot_s += dt
        rows.append((t, out[0]))
    res = {"value": tot_macs / tot_s, "unit": "MAC/s", "cores": threads, "kind": "port",
           "cpu_model": cpu_model(),
           "sample": "one template per length group (%s), its active channels, the first "
                     "%d samples of the day (%.2e MACs, %.1f s); oracle/libfmf_oracle.so, "
                     "the C restatement of the reference CPU path built with the "
                     "reference's flags (-O3 -fopenmp -ftree-vectorize, -march=x86-64-v3)"
                     % ("/".join(groups), n_samples, tot_macs, tot_s),
           "per_length": groups}
    return res, rows


def detection_leg(lib, out, ranges, w, T, n_corr, sep, step, n_mad=8.0):
    """Median/MAD and detection on the cc rows of the last step, kept on the
    device; kernel time from events on the launch stream.  Algorithmic bytes:
    the two order statistics and the detection test each read every valid
    lag once (12 B per template-lag)."""
    import torch
    from fast_matched_filter_amd import _lib, detect as fdet
    dev = out.device
    rg = ranges
    ws = torch.empty(int(lib.fmf_detect_workspace_bytes(T, n_corr)), dtype=torch.uint8, device=dev)
    cap = 1 << 20
    det_t = torch.empty(cap, dtype=torch.int32, device=dev)
    det_l = torch.empty(cap, dtype=torch.int32, device=dev)
    det_v = torch.empty(cap, dtype=torch.float32, device=dev)
    n_det = torch.zeros(1, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream(dev)

    def run():
        med, mad = fdet.median_mad(out, rg, ws)
        e1.record(stream)
        _lib.check(lib.fmf_detect_device(
            out.data_ptr(), T, n_corr, rg.data_ptr(), med.data_ptr(), mad.data_ptr(),
            float(n_mad), int(sep), det_t.data_ptr(), det_l.data_ptr(), det_v.data_ptr(), cap,
            n_det.data_ptr(), ws.data_ptr(), ws.numel(), stream.cuda_stream), "fmf_detect_device")

    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    run()
    torch.cuda.synchronize()
    e0.record(stream)
    run()
    e2.record(stream)
    torch.cuda.synchronize()
    n = int(n_det.item())
    lags = rg[:, 1].long() - rg[:, 0].long()
    n_valid = float(lags.clamp(min=0).sum().item())
    t_ms = e0.elapsed_time(e2)
    found = set(zip(det_t[:min(n, cap)].cpu().tolist(), det_l[:min(n, cap)].cpu().tolist()))
    # an event at lag l lies between outputs l // step and ceil(l / step); the
    # detection is whichever of the two holds the larger cc.  At step > 1 only
    # events on the lag grid are detectable here: the synthetic templates are
    # white noise, whose cc one sample off the event is at noise level.
    events = [(int(t), int(l)) for t, l in w.get("events", [])]
    hits = [(t, l) for t, l in events
            if (t, l // step) in found or (t, -(-l // step)) in found]
    on_grid = [e for e in events if e[1] % step == 0]
    del ws
    return {"n_mad": n_mad, "min_separation": int(sep), "ms": t_ms,
            "median_mad_ms": e0.elapsed_time(e1), "detect_ms": e1.elapsed_time(e2),
            "n_detections": n, "injected_events_detected": "%d/%d" % (len(hits), len(events)),
            "injected_events_on_lag_grid_detected": "%d/%d" % (
                sum(1 for e in hits if e[1] % step == 0), len(on_grid)),
            "algorithmic_GB_per_s": 12.0 * n_valid / (t_ms * 1e-3) / 1e9}


def _pmc_kib(rocprof, counter, child_args, out_dir, pattern, timeout):
    """KiB per dispatch of one counter for the largest dispatch whose kernel
    name contains `pattern`, from a rocprofv3 run of this script (a child
    process; the profiler's own library starts with it)."""
    import csv
    import subprocess
    d = os.path.join(out_dir, counter)
    cmd = [rocprof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run", "--",
           sys.executable, os.path.abspath(__file__)] + child_args
    env = dict(os.environ, TMPDIR="/tmp")
    r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, timeout=timeout)
    path = None
    for root, _, files in os.walk(d):
        for f in files:
            if f.endswith("counter_collection.csv"):
                path = os.path.join(root, f)
    if path is None:
        raise RuntimeError("rocprofv3 %s wrote no counters (exit %d): %s"
                           % (counter, r.returncode, r.stdout[-300:].decode(errors="replace")))
    best = 0.0
    with open(path) as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] == counter and pattern in row["Kernel_Name"]:
                best = max(best, float(row["Counter_Value"]))
    if best <= 0.0:
        raise RuntimeError("no %s dispatch of a kernel named *%s*" % (counter, pattern))
    return best


def measure_traffic(child_args, kernel_pattern, timeout=150):
    """HBM bytes of one launch of the correlation kernel, measured in this run:
    FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of that launch in two
    rocprofv3 --pmc runs of this script on the same workload, each scaled by
    what the same counter reads for a device copy of 1 GiB (two more runs),
    the known-size kernel that calibrates the counters' units.  A run that
    fails or outlasts `timeout` seconds (a child takes about 10 s for the
    batched workload) ends the leg with an error; the line then carries the
    recorded figure."""
    import shutil
    import tempfile
    rocprof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(rocprof):
        raise RuntimeError("rocprofv3 not found")
    out_dir = tempfile.mkdtemp(prefix="fmf_traffic_", dir="/tmp")
    copy_bytes = 1 << 30
    copy_args = ["--copy-calibration", str(copy_bytes >> 20), "--steps", "2"]
    fetch = _pmc_kib(rocprof, "FETCH_SIZE", child_args, out_dir, kernel_pattern, timeout)
    write = _pmc_kib(rocprof, "WRITE_SIZE", child_args, out_dir, kernel_pattern, timeout)
    fetch_c = _pmc_kib(rocprof, "FETCH_SIZE", copy_args, out_dir + "/copy", "copyBuffer", timeout)
    write_c = _pmc_kib(rocprof, "WRITE_SIZE", copy_args, out_dir + "/copy", "copyBuffer", timeout)
    sf = copy_bytes / (fetch_c * 1024.0)
    sw = copy_bytes / (write_c * 1024.0)
    shutil.rmtree(out_dir, ignore_errors=True)
    return {"bytes": fetch * 1024.0 * sf + write * 1024.0 * sw,
            "fetch_KiB": fetch, "write_KiB": write,
            "fetch_scale": sf, "write_scale": sw,
            "source": "rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE of one launch in this run, "
                      "scaled by the counters of a 1 GiB device copy"}


def optional_leg(name, fn):
    """Run one of the line's optional legs (detection, FP32 kernel, FP64 spot
    check, screener, graph replays, CPU baseline): an error there is reported
    in the line instead of ending the run after the headline figure is
    measured."""
    try:
        return fn()
    except Exception as e:  # the line still carries the measured value
        print("%s leg failed: %s: %s" % (name, type(e).__name__, e), file=sys.stderr, flush=True)
        return {"error": "%s: %s" % (type(e).__name__, e)}


def screened_leg(w, d_tp, d_sst, d_mo, d_w, d_data, out, T, SC, N, L, exact_ms, reps=3):
    """The learned screener on the same day (committed weights): wall time of
    the screened scan (coarse scan, features, network, gate, exact scan of the
    kept items, MAD of the sampled tiles) against the exact step, and two
    recalls against the exact scan in `out`:

    * event recall: injected events that the exact scan detects (peak cc
      within +-64 lags above 6 MAD of its row) and the screened output also
      detects above 6 of its own MAD estimate;
    * detection recall: the exact scan's detection list (cc above 6 MAD and
      the first maximum within +-N lags, ``fmf_detect``) against the list the
      same detection finds in the screened output with the same thresholds,
      a detection counted when the screened list holds one of that template
      within +-8 lags."""
    import torch
    from fast_matched_filter_amd import screener
    from fast_matched_filter_amd.detect import median_mad
    from fast_matched_filter_amd.screener_train import (
        DETECT_MAD, _event_windows, _window_max, detection_recall)
    path = os.path.join(os.path.dirname(screener.__file__), "screener_weights.npz")
    wts = np.load(path)
    dev = out.device
    net = screener.Network(dev, seed=0)
    net.load(wts["w16"])
    gate = float(wts["gate"])
    day = dict(templates=d_tp.view(T, SC, N), moveouts=d_mo.view(T, SC), weights=d_w.view(T, SC),
               data=d_data.view(SC, L), sst=d_sst.view(T, SC),
               events=[(int(t), int(l)) for t, l in w["events"]], n_tp=N)
    times, info = [], None
    for rep in range(reps + 1):  # the first scan carries one-time set-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cs, computed, mads_est, _, info = screener.screened_scan(
            day, net, gate, seed=rep, factor=int(wts["coarse"]))
        torch.cuda.synchronize()
        if rep:
            times.append(time.perf_counter() - t0)
    _, mads = median_mad(out)
    evs = _event_windows(out, mads, day["events"], info["tile_out"])
    me = mads_est.cpu().numpy()
    mx = _window_max(cs, [(t, lag) for t, lag, _ in evs]).cpu().numpy() if evs else []
    hit = sum(1 for (t, _, _), v in zip(evs, mx) if v > DETECT_MAD * me[t])
    det = detection_recall(out, cs, mads_est, day["moveouts"], N, L, 1, sep=N,
                           events=day["events"])
    det9 = detection_recall(out, cs, mads_est, day["moveouts"], N, L, 1, sep=N,
                            events=day["events"], n_mad=9.0)
    # kept items sum in the exact scan's order (32x32 tile shapes): same bits
    mask = computed.repeat_interleave(info["tile_out"], dim=1)[:, :cs.shape[1]]
    same = bool(torch.equal(cs[mask], out[mask]))
    del mask
    ms = 1e3 * float(np.mean(times))
    return {"weights": "fast_matched_filter_amd/screener_weights.npz", "reps": reps,
            "ms_per_step": ms, "exact_ms_per_step": exact_ms, "speedup_vs_exact": exact_ms / ms,
            "kept_fraction": info["kept_fraction"], "sampled_fraction": info["sampled_fraction"],
            "fallback": info["fallback"],
            "events_detected_by_exact": len(evs), "event_recall": hit / max(len(evs), 1),
            "exact_detections": det["n_exact"], "screened_detections": det["n_screened"],
            "detection_recall_vs_exact": det["recall"],
            "detection_recall_of_event_detections": det["recall_events"],
            "at_9_mad": {"exact_detections": det9["n_exact"],
                         "exact_detections_on_events": det9["n_exact_on_events"],
                         "detection_recall_vs_exact": det9["recall"]},
            "kept_tiles_equal_exact": same,
            "stage_ms": {k: 1e3 * v for k, v in info["times"].items()},
            "mad_rel_err_max": float(np.abs(me / mads.cpu().numpy() - 1.0).max())}


def day_leg(lib, steps=3, screened=True):
    """The single-length day beside the batched headline: workload ``day``
    (T=100 templates of N=800 over 20 stations x 3 components, L=8 640 000,
    step 1) through the single-call device entry point with the library's
    default tile shape, timed as the headline is (eager steps between device
    synchronisations, the correlation launch from its HIP events), then the
    learned screener on the same day against that exact scan.  Its buffers
    are released before it returns."""
    import ctypes as ct
    import torch
    from fast_matched_filter_amd import _lib
    from fast_matched_filter_amd.device import matched_filter_device, workspace_bytes
    from fast_matched_filter_amd.synthetic import make_workload
    w = make_workload("day")
    digest = input_sha256(w)
    tp = w["templates"]
    T, S, C, N = tp.shape
    SC, L, step = S * C, w["data"].shape[-1], w["step"]
    n_corr = (L - N) // step + 1
    dev = torch.device("cuda:0")
    sst = np.sum(tp.astype(np.float32) ** 2, axis=-1).astype(np.float32)
    d_tp = torch.from_numpy(tp.reshape(-1)).to(dev)
    d_sst = torch.from_numpy(sst.reshape(-1)).to(dev)
    d_mo = torch.from_numpy(w["moveouts"].reshape(-1)).to(dev)
    d_w = torch.from_numpy(w["weights"].reshape(-1)).to(dev)
    d_data = torch.from_numpy(w["data"].reshape(-1)).to(dev)
    out = torch.empty((T, n_corr), dtype=torch.float32, device=dev)
    ws = torch.empty(workspace_bytes(N, L, T, S, C, 0, KERNELS["bf16x3"], 0),
                     dtype=torch.uint8, device=dev)

    def step_fn():
        matched_filter_device(d_tp, d_sst, d_mo, d_w, d_data, step, N, L, T, S, C, n_corr,
                              out, normalize=0, network_sum=True, kernel=KERNELS["bf16x3"],
                              workspace=ws, stream=torch.cuda.current_stream(dev))
    step_fn()
    torch.cuda.synchronize()
    lib.fmf_kernel_timer_enable(1)
    t0 = time.perf_counter()
    for _ in range(steps):
        step_fn()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / steps
    lib.fmf_kernel_timer_enable(0)
    tot, cnt = ct.c_double(), ct.c_int()
    _lib.check(lib.fmf_kernel_timer_read(0, ct.byref(tot), ct.byref(cnt)), "timer")
    kern_ms = tot.value / max(cnt.value, 1)
    _lib.check(lib.fmf_kernel_timer_read(1, ct.byref(tot), ct.byref(cnt)), "timer")
    lags, _, _ = lag_counts(w["moveouts"], step, N, L, n_corr)
    active = (w["weights"].reshape(T, -1) != 0).sum(axis=1)
    macs = float((lags * active * N).sum())
    achieved = 2.0 * macs / (kern_ms * 1e-3) / 1e12
    r = {"workload": "day: T=%d templates, S=%d stations, C=%d components, N=%d, L=%d samples, "
                     "step=%d, short normalization, network sum" % (T, S, C, N, L, step),
         "input_sha256_matches_recorded": digest == INPUT_SHA256["day"],
         "mfma_tile_shape": int(lib.fmf_get_mfma_tiles(step)),
         "steps": steps, "ms_per_step": el * 1e3, "value": macs / el, "unit": "MAC/s",
         "kernel_ms": kern_ms, "frac": achieved / PEAK_TFLOPS["bf16x3"]}
    print("day: %.2f ms per step, %.3g MAC/s, correlation launch %.2f ms (frac %.3f)"
          % (r["ms_per_step"], r["value"], kern_ms, r["frac"]), file=sys.stderr, flush=True)
    if screened:
        r["screened"] = optional_leg("screened", lambda: screened_leg(
            w, d_tp, d_sst, d_mo, d_w, d_data, out, T, SC, N, L, el * 1e3))
        s = r["screened"]
        if "error" not in s:
            print("screened scan: %.2f ms vs exact %.2f ms (%.2fx), event recall %.4f, "
                  "detection recall %.4f" % (s["ms_per_step"], s["exact_ms_per_step"],
                                             s["speedup_vs_exact"], s["event_recall"],
                                             s["detection_recall_vs_exact"]),
                  file=sys.stderr, flush=True)
    del out, ws, d_tp, d_sst, d_mo, d_w, d_data
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="batched")
    ap.add_argument("--kernel", default="bf16x3", choices=sorted(KERNELS))
    ap.add_argument("--fp32-steps", type=int, default=1,
                    help="also time the plain FP32 kernel on the same workload (0 = skip)")
    ap.add_argument("--templates", type=int, default=None, help="override T")
    ap.add_argument("--samples", type=int, default=None, help="override L")
    ap.add_argument("--lengths", default=None,
                    help="override the template lengths (comma-separated; one value gives "
                         "one length group of the workload on its own)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-samples", type=int, default=4_320_000,
                    help="samples of the day in the CPU baseline's sample")
    ap.add_argument("--no-screened", action="store_true",
                    help="skip the learned screener's leg (step-1 single-length workloads)")
    ap.add_argument("--no-detection", action="store_true")
    ap.add_argument("--no-day", action="store_true",
                    help="batched workloads: skip the single-length day leg (T=100 day, "
                         "exact and screened scans)")
    ap.add_argument("--step", type=int, default=0, help="override the workload's step")
    ap.add_argument("--autotune", action="store_true",
                    help="time the bf16x3 tile shapes on this workload first and keep the best")
    ap.add_argument("--graph-steps", type=int, default=3,
                    help="HIP-graph replays of the captured step timed after the eager steps")
    ap.add_argument("--persistent-blocks", type=int, default=0,
                    help="batched workloads: persistent grid of about this many blocks (0 = one per item)")
    ap.add_argument("--mfma-tiles", type=int, default=0,
                    help="bf16x3 tile shape: 24, 42, 43, 81 (16x16 tiles) or 212, 412, 413, 421, "
                         "422, 241, 121, 141, 1422 (32x32 tiles); 0 = library choice")
    ap.add_argument("--host-calls", type=int, default=0,
                    help="also time this many calls of the host-array API on the workload "
                         "(inputs and output in host memory, PCIe copies included)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the HBM traffic measurement (rocprofv3 runs of this script "
                         "on the same workload); the line then carries the recorded figure")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (steps times), "
                         "the known-size kernel that scales FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box of this pool has one MI355X, "
              "and template sharding across GPUs is a design note only. Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(args.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4, "copies": args.steps}))
        return
    from fast_matched_filter_amd import _lib
    from fast_matched_filter_amd.device import matched_filter_device, workspace_bytes
    from fast_matched_filter_amd.synthetic import make_workload

    over = {}
    if args.templates:
        over["n_templates"] = args.templates
    if args.samples:
        over["n_samples_data"] = args.samples
    if args.lengths:
        vals = tuple(int(v) for v in args.lengths.split(","))
        over["n_samples_template"] = vals if len(vals) > 1 else vals[0]
    t_gen = time.perf_counter()
    w = make_workload(args.workload, **over)
    print("workload %s generated in %.1f s" % (args.workload, time.perf_counter() - t_gen),
          file=sys.stderr, flush=True)
    digest = input_sha256(w)
    expected = INPUT_SHA256.get(args.workload) if not over else None
    if expected and digest != expected:
        print("warning: generated inputs differ from the recorded SHA-256 (%s)" % digest,
              file=sys.stderr, flush=True)
    tp = w["templates"]
    T, S, C, N = tp.shape
    SC = S * C
    L = w["data"].shape[-1]
    step = args.step or w["step"]
    lengths = np.asarray(w["lengths"], dtype=np.int32)
    distinct = np.unique(lengths)
    batched = bool(np.any(lengths != N))
    n_corr = (L - int(lengths.min())) // step + 1
    kernel = KERNELS[args.kernel]
    if batched and args.kernel != "bf16x3":
        print("templates of several lengths run on the batched bf16x3 launch", file=sys.stderr)
        sys.exit(2)

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    # sum of squares over each template's own length (zero taps add nothing)
    sst = np.sum(tp.astype(np.float32) ** 2, axis=-1).astype(np.float32)
    d_tp = torch.from_numpy(tp.reshape(-1)).to(dev)
    d_sst = torch.from_numpy(sst.reshape(-1)).to(dev)
    d_mo = torch.from_numpy(w["moveouts"].reshape(-1)).to(dev)
    d_w = torch.from_numpy(w["weights"].reshape(-1)).to(dev)
    d_data = torch.from_numpy(w["data"].reshape(-1)).to(dev)
    out = torch.empty((T, n_corr), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)
    lib = _lib.require()
    if args.mfma_tiles:
        _lib.check(lib.fmf_set_mfma_tiles(step, args.mfma_tiles), "fmf_set_mfma_tiles")
    if batched:
        from fast_matched_filter_amd.device import (batched_workspace_bytes,
                                                    matched_filter_batched_device)
        d_len = torch.from_numpy(lengths).to(dev)
        ws = torch.empty(batched_workspace_bytes(distinct, N, L, T, S, C),
                         dtype=torch.uint8, device=dev)

        def step_fn(st=stream):
            matched_filter_batched_device(d_tp, d_sst, d_mo, d_w, d_data, d_len, distinct,
                                          step, N, L, T, S, C, n_corr, out,
                                          persistent_blocks=args.persistent_blocks,
                                          workspace=ws, stream=st)
    else:
        ws = torch.empty(workspace_bytes(N, L, T, S, C, 0, kernel, 0), dtype=torch.uint8, device=dev)

        def step_fn(st=stream):
            matched_filter_device(d_tp, d_sst, d_mo, d_w, d_data, step, N, L, T, S, C,
                                  n_corr, out, normalize=0, network_sum=True,
                                  kernel=kernel, workspace=ws, stream=st)

    tuned = None
    if args.autotune and args.kernel == "bf16x3" and step in (1, 2):
        from fast_matched_filter_amd import autotune
        n_act = int(np.round((w["weights"].reshape(T, -1) != 0).sum(axis=1).mean()))
        tuned = autotune.tune(step_fn, step, N, n_act, repeats=1, use_cache=False, verbose=True)
        print("autotuned tiles per wave: %s" % (tuned,), file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        step_fn()
    torch.cuda.synchronize()
    print("warmup done", file=sys.stderr, flush=True)
    lib.fmf_kernel_timer_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step_fn()
        if args.steps >= 10 and (i + 1) % 10 == 0:
            print("step %d queued" % (i + 1), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    lib.fmf_kernel_timer_enable(0)
    import ctypes as ct
    tot, cnt = ct.c_double(), ct.c_int()
    _lib.check(lib.fmf_kernel_timer_read(0, ct.byref(tot), ct.byref(cnt)), "timer")
    corr_ms, corr_n = tot.value, cnt.value
    _lib.check(lib.fmf_kernel_timer_read(1, ct.byref(tot), ct.byref(cnt)), "timer")
    energy_ms, energy_n = tot.value, cnt.value
    print("timed %d steps in %.3f s" % (args.steps, elapsed), file=sys.stderr, flush=True)

    lags, olo, ohi = lag_counts(w["moveouts"], step, lengths, L, n_corr)
    active = (w["weights"].reshape(T, -1) != 0).sum(axis=1)
    macs_t = lags * active * lengths.astype(np.int64)
    macs = float(macs_t.sum())
    ms_step = elapsed * 1e3 / args.steps
    value = macs / (elapsed / args.steps)

    # The step captured in a HIP graph (kernel timer off) and replayed: the
    # launch form of a fixed workload repeated over many days.
    def graph_leg():
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(stream)
        with torch.cuda.stream(side):  # warm-up on the capture stream
            step_fn(side)
        stream.wait_stream(side)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            step_fn(torch.cuda.current_stream(dev))
        g.replay()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.graph_steps):
            g.replay()
        torch.cuda.synchronize()
        g_ms = (time.perf_counter() - t1) * 1e3 / args.graph_steps
        del g
        print("graph replays: %.3f ms per step (eager %.3f)" % (g_ms, ms_step),
              file=sys.stderr, flush=True)
        return {"ms_per_step": g_ms, "eager_ms_per_step": ms_step,
                "value": macs / (g_ms * 1e-3), "steps": args.graph_steps}
    graph_info = optional_leg("graph", graph_leg) if args.graph_steps > 0 else None

    kern_ms = corr_ms / max(corr_n, 1)
    achieved = 2.0 * macs / (kern_ms * 1e-3) / 1e12
    peak = PEAK_TFLOPS[args.kernel]
    # normalization pass: the data read once and every distinct length's
    # table written once (one pass serves all the lengths)
    energy_bytes = float(SC * 4.0 * L + sum(SC * 4.0 * (L - int(n) + 1) for n in distinct))

    # FP64 check on the bf16x3 output of the last step
    rng = np.random.Generator(np.random.PCG64(99))
    spot = optional_leg("fp64 spot check",
                        lambda: fp64_spot_check(w, out, step, lengths, rng))
    if isinstance(spot, dict):
        dmax = drms = None
        argmax_agree = spot["error"]
    else:
        dmax, drms, dn, n_agree, n_ev = spot
        argmax_agree = "%d/%d" % (n_agree, n_ev)
        print("fp64 spot check: max %.3g rms %.3g over %d lags; argmax agrees at %d/%d events"
              % (dmax, drms, dn, n_agree, n_ev), file=sys.stderr, flush=True)

    detection = None
    if not args.no_detection:
        def det_leg():
            rg = torch.from_numpy(np.stack([olo, ohi], axis=1).astype(np.int32)).to(dev)
            return detection_leg(lib, out, rg, w, T, n_corr, int(lengths.min()), step)
        detection = optional_leg("detection", det_leg)

    gpu_rows = {}
    if not args.no_cpu_baseline:  # rows of the CPU sample, kept before `out` is reused
        for n in distinct:
            t = int(np.nonzero(lengths == n)[0][0])
            gpu_rows[t] = out[t, :min(n_corr, 1 << 20)].cpu().numpy()

    screened = None
    if not args.no_screened and args.kernel == "bf16x3" and step == 1 and not batched:
        screened = optional_leg("screened", lambda: screened_leg(
            w, d_tp, d_sst, d_mo, d_w, d_data, out, T, SC, N, L, ms_step))
        if "error" not in screened:
            print("screened scan: %.2f ms vs exact %.2f ms (%.2fx), event recall %.4f, "
                  "detection recall %.4f" % (screened["ms_per_step"], screened["exact_ms_per_step"],
                                             screened["speedup_vs_exact"], screened["event_recall"],
                                             screened["detection_recall_vs_exact"]),
                  file=sys.stderr, flush=True)

    fp32 = None
    if args.kernel == "bf16x3" and args.fp32_steps > 0:
        def fp32_leg():
            """The plain FP32 kernel on the same workload: one launch per
            length group (each group's templates cut to their own length),
            the output rows of a group in a buffer of their own."""
            groups = []
            for n in distinct:
                rows = np.nonzero(lengths == n)[0]
                sel = torch.from_numpy(rows).to(dev)
                g_tp = d_tp.view(T, SC, N)[sel][:, :, :int(n)].contiguous().view(-1)
                g_sst = d_sst.view(T, SC)[sel].contiguous().view(-1)
                g_mo = d_mo.view(T, SC)[sel].contiguous().view(-1)
                g_w = d_w.view(T, SC)[sel].contiguous().view(-1)
                g_n_corr = (L - int(n)) // step + 1
                groups.append((int(n), rows.size, g_tp, g_sst, g_mo, g_w, g_n_corr))
            g_out = torch.empty(max(g[1] * g[6] for g in groups), dtype=torch.float32, device=dev)
            g_ws = torch.empty(max(workspace_bytes(g[0], L, g[1], S, C, 0, 0, 0) for g in groups),
                               dtype=torch.uint8, device=dev)

            def fp32_step():
                for n, tg, g_tp, g_sst, g_mo, g_w, g_n_corr in groups:
                    matched_filter_device(g_tp, g_sst, g_mo, g_w, d_data, step, n, L, tg, S, C,
                                          g_n_corr, g_out[:tg * g_n_corr], normalize=0,
                                          network_sum=True, kernel=KERNELS["fp32"],
                                          workspace=g_ws, stream=stream)
            # warm-up: the smallest group's first template
            n0, _, g_tp0, g_sst0, g_mo0, g_w0, nc0 = groups[0]
            matched_filter_device(g_tp0[:SC * n0], g_sst0[:SC], g_mo0[:SC], g_w0[:SC], d_data,
                                  step, n0, L, 1, S, C, nc0, g_out[:nc0], kernel=KERNELS["fp32"],
                                  workspace=g_ws, stream=stream)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for _ in range(args.fp32_steps):
                fp32_step()
            torch.cuda.synchronize()
            f_el = (time.perf_counter() - t1) / args.fp32_steps
            del g_out, g_ws, groups
            r = {"value": macs / f_el, "unit": "MAC/s", "ms_per_step": f_el * 1e3,
                 "steps": args.fp32_steps, "launches_per_step": len(distinct),
                 "achieved_tflops_step": 2.0 * macs / f_el / 1e12,
                 "bf16x3_over_fp32": value / (macs / f_el)}
            print("fp32 kernel: %.3g MAC/s (bf16x3 %.2fx)" % (r["value"], r["bf16x3_over_fp32"]),
                  file=sys.stderr, flush=True)
            return r
        fp32 = optional_leg("fp32 kernel", fp32_leg)

    cpu = None
    if not args.no_cpu_baseline:
        def cpu_leg():
            r, cpu_rows = cpu_baseline(w, lengths, step, min(args.cpu_samples, L))
            dd = 0.0
            for t, row in cpu_rows:
                g = gpu_rows[t]
                # lags valid in both calls: the CPU sample ends earlier
                m = min(row.size, g.size, int(ohi[t]))
                stop = int(np.nonzero(row[:m])[0].max()) + 1 if np.any(row[:m]) else 0
                dd = max(dd, float(np.abs(g[int(olo[t]):stop] - row[int(olo[t]):stop]).max()))
            r["max_abs_dcc_vs_gpu"] = dd
            print("cpu baseline: %.3g MAC/s on %d threads (%s)" % (r["value"], r["cores"],
                                                                  r["cpu_model"]),
                  file=sys.stderr, flush=True)
            return r
        cpu = optional_leg("cpu baseline", cpu_leg)

    day = None
    if batched and not args.no_day and args.kernel == "bf16x3" and not over:
        t_day = time.perf_counter()
        day = optional_leg("day", lambda: day_leg(lib, screened=not args.no_screened))
        print("day leg: %.1f s" % (time.perf_counter() - t_day), file=sys.stderr, flush=True)

    host = None
    if args.host_calls > 0:
        def host_leg():
            """The host-array entry points on the same workload, after the
            device buffers above are released: matched_filter(arch='hip_mfma')
            (one library call: uploads, template blocks whose outputs are
            copied back while the next block computes) or, for templates of
            several lengths, matched_filter_batched (uploads, one launch,
            download).  The first call is timed too: the library is already
            initialised by the steps above."""
            import fast_matched_filter_amd as fmf
            nonlocal out, ws
            out = ws = None
            torch.cuda.empty_cache()
            arch = "hip_mfma" if args.kernel == "bf16x3" else "hip"
            if batched:
                run = lambda: fmf.matched_filter_batched(
                    w["templates"], lengths, w["moveouts"], w["weights"], w["data"], step)
                entry = "matched_filter_batched"
            else:
                run = lambda: fmf.matched_filter(
                    w["templates"], w["moveouts"], w["weights"], w["data"], step,
                    arch=arch, check_zeros=False)
                entry = "matched_filter(arch=%r)" % arch
            times = []
            for _ in range(args.host_calls):
                t1 = time.perf_counter()
                cc = run()
                times.append(time.perf_counter() - t1)
                del cc
            h = min(times)
            r = {"entry": entry, "calls": args.host_calls, "ms_per_call": h * 1e3,
                 "value": macs / h, "unit": "MAC/s",
                 "output_GB": T * n_corr * 4 / 1e9,
                 "input_GB": (w["data"].nbytes + w["templates"].nbytes) / 1e9,
                 "over_device_step": h * 1e3 / ms_step}
            print("host call: %.1f ms (%.3g MAC/s, %.1f GB out; device step %.1f ms)"
                  % (h * 1e3, r["value"], r["output_GB"], ms_step), file=sys.stderr, flush=True)
            return r
        host = optional_leg("host call", host_leg)

    shape = int(lib.fmf_get_mfma_tiles(step)) if args.kernel == "bf16x3" else 0
    recorded = MEASURED_TRAFFIC.get((args.workload if not over else None, args.kernel, shape, step))
    traffic = None
    if not args.no_traffic:
        # the profiled child allocates its own copy of the workload beside this
        # process's (about 100 GB for the batched workload: both fit in HBM)
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        child = ["--workload", args.workload, "--step", str(step), "--kernel", args.kernel,
                 "--steps", "1", "--warmup", "0", "--no-cpu-baseline", "--fp32-steps", "0",
                 "--graph-steps", "0", "--no-detection", "--no-screened", "--no-traffic",
                 "--no-day"]
        if shape:
            child += ["--mfma-tiles", str(shape)]
        for flag, v in (("--templates", args.templates), ("--samples", args.samples),
                        ("--lengths", args.lengths)):
            if v:
                child += [flag, str(v)]
        pattern = "corr_fp32" if args.kernel == "fp32" else "corr_mfma"
        t_tr = time.perf_counter()
        traffic = optional_leg("traffic", lambda: measure_traffic(child, pattern))
        print("traffic leg: %.1f s" % (time.perf_counter() - t_tr), file=sys.stderr, flush=True)
    measured = traffic.get("bytes") if traffic and "bytes" in traffic else None
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "MAC/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.kernel == "fp32" else "bf16x3->f32",
        "data": "synthetic",
        "config": {"workload": "%s: T=%d templates, S=%d stations, C=%d components, "
                               "N=%s, L=%d samples, step=%d, short normalization, network sum%s"
                               % (args.workload, T, S, C,
                                  "/".join(str(v) for v in distinct), L, step,
                                  ", 8-30 active stations per template, one batched launch"
                                  if batched else ""),
                   "autotuned_tiles": tuned[0] if tuned else None,
                   "kernel": args.kernel, "macs_per_step": macs,
                   "macs_per_length": {str(int(n)): float(macs_t[lengths == n].sum())
                                       for n in distinct},
                   "mfma_tile_shape": shape or None,
                   "input_sha256": digest,
                   "input_sha256_matches_recorded": (digest == expected) if expected else None},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak,
                     "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": measured if measured is not None else recorded,
                     "traffic_measured": traffic,
                     "traffic_recorded": recorded,
                     "kernel": "corr_%s" % args.kernel, "kernel_ms": kern_ms,
                     "launches": corr_n,
                     # three bf16 products per useful MAC: the useful-work
                     # ceiling of the split form is a third of the dense peak
                     "frac_of_bf16x3_ceiling": (3.0 * achieved / peak
                                                if args.kernel == "bf16x3" else None)},
        "hip_graph": graph_info,
        "normalization_pass": {"kernel": "window_energy", "ms": energy_ms / max(energy_n, 1),
                               "tables": int(len(distinct)),
                               "algorithmic_GB_per_s": energy_bytes / (energy_ms / max(energy_n, 1) * 1e-3) / 1e9
                               if energy_n else None},
        "max_abs_dcc_fp64": dmax,
        "rms_abs_dcc_fp64": drms,
        "argmax_agreement_fp64": argmax_agree,
        "fp64_checked_templates": spot_templates(lengths),
        "detection": detection,
        "fp32_kernel": fp32,
        "screened": screened,
        "day": day,
        "cpu_baseline": cpu,
        "host_call": host,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
