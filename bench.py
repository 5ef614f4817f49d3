"""Matched-filter throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload day]
                    [--kernel fp32|bf16x3] [--no-cpu-baseline]

A step is one pass of the hot path (window energies + template preparation +
network correlation of every template over the whole day) with all inputs
already resident in HBM, through libfmf_hip.so's device entry point.  Prints
one JSON line: template*channel*lag MACs/s of the whole step, the dominant
kernel's roofline position (its duration taken from HIP events bracketing
each of its launches on the stream it runs on), the max |dcc| against FP64
dot products at sampled lags, and the reference CPU path timed on a bounded
sample on this host's cores.
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

# HBM bytes per correlation launch measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE on this exact configuration (workload, kernel, tile shape),
# each scaled by what the counter reads for a device copy of 1 GiB (FETCH_SIZE
# x2.00, WRITE_SIZE x1.00); summary and commands in
# profiles/hbm_traffic_corr_day.md.  Other configurations report null.
MEASURED_TRAFFIC = {
    ("day", "bf16x3", 422, 1, None, None): 4.11e10 + 3.46e9,
}


def lag_counts(moveouts, step, N, L, n_corr):
    """Computed lags per template: lag range rule of the reference
    (matched_filter.c:53-79), output index clipped to n_corr.  N may be one
    length or one per template."""
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


def fp64_spot_check(w, out_dev, step, lengths, rng, n_templates=4, n_random=48, half=32):
    """Max and RMS |dcc| against FP64 dot products at the injected events and
    random lags of a few templates (short normalization), and the lag-argmax
    agreement: at each injected event of those templates, the lag of the
    largest cc within +-`half` output lags, GPU against FP64."""
    import torch
    tp = w["templates"]
    T, S, C, _ = tp.shape
    mo = w["moveouts"].reshape(T, S * C)
    wt = w["weights"].reshape(T, S * C).astype(np.float64)
    d = w["data"].reshape(S * C, -1)
    L = d.shape[1]
    n_corr = out_dev.shape[1]
    _, olo, ohi = lag_counts(w["moveouts"], step, lengths, L, n_corr)
    def fp64_cc(t, outs):
        N = int(lengths[t])
        idx = np.arange(N)
        exp = np.zeros(outs.size)
        for ch in range(S * C):
            if wt[t, ch] == 0:
                continue
            win = d[ch][(outs * step + mo[t, ch])[:, None] + idx[None, :]].astype(np.float64)
            tk = tp.reshape(T, S * C, -1)[t, ch, :N].astype(np.float64)
            num = win @ tk
            den = (tk @ tk) * np.einsum("ij,ij->i", win, win)
            exp += wt[t, ch] * np.where(den > 1e-6, num / np.sqrt(np.maximum(den, 1e-300)), 0.0)
        return exp

    errs = []
    agree = n_ev = 0
    for t in range(min(n_templates, T)):
        ev = [lag // step for (tt, lag) in w["events"] if tt == t]
        outs = np.unique(np.concatenate([
            np.asarray(ev, dtype=np.int64),
            rng.integers(olo[t], ohi[t], size=n_random)]))
        outs = outs[(outs >= olo[t]) & (outs < ohi[t])]
        got = out_dev[t, torch.as_tensor(outs, device=out_dev.device)].cpu().numpy()
        errs.append(np.abs(got - fp64_cc(t, outs)))
        for e in ev:
            win = np.arange(max(olo[t], e - half), min(ohi[t], e + half + 1))
            if win.size == 0:
                continue
            g = out_dev[t, torch.as_tensor(win, device=out_dev.device)].cpu().numpy()
            n_ev += 1
            agree += int(np.argmax(g) == np.argmax(fp64_cc(t, win)))
    e = np.concatenate(errs)
    return float(e.max()), float(np.sqrt(np.mean(e * e))), int(e.size), agree, n_ev


def cpu_baseline(w, N, step):
    """Reference CPU path (compiled reference, else the oracle restatement)
    on one template of the workload over the whole trace (the reference
    takes one template length per call: template 0 at its own length)."""
    from oracle import cpu
    cores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    cores = int(os.environ["OMP_NUM_THREADS"])
    if cpu.reference_available(portable=True):
        lib, kind = "ref_portable", "reference"
    else:
        lib, kind = "oracle", "port"
    tsel = [0]
    tp = w["templates"][tsel][..., :N]
    mo = w["moveouts"][tsel]
    wt = w["weights"][tsel]
    t0 = time.perf_counter()
    out = cpu.run(tp, mo, wt, w["data"], step, lib=lib)
    dt = time.perf_counter() - t0
    L = w["data"].shape[-1]
    n_corr = (L - N) // step + 1
    lags, _, _ = lag_counts(mo, step, N, L, n_corr)
    active = (wt.reshape(len(tsel), -1) != 0).sum(axis=1)
    macs = float((lags * active).sum() * N)
    return dict(value=macs / dt, unit="MAC/s", cores=cores, kind=kind,
                sample="1 template x %d channels x %d lags x N=%d (%.2e MACs, %.1f s), %s"
                % (int(active[0]), int(lags[0]), N, macs, dt,
                   "oracle/_ref reference build (-O3 -fopenmp -ftree-vectorize -march=x86-64-v3)"
                   if kind == "reference" else "oracle C restatement")), out


def detection_leg(lib, out, d_mo, w, T, n_corr, N, L, step, n_mad=8.0):
    """Median/MAD and detection on the cc rows of the last step, kept on the
    device; kernel time from events on the launch stream.  Algorithmic bytes:
    the two order statistics and the detection test each read every valid
    lag once (12 B per template-lag)."""
    import torch
    from fast_matched_filter_amd import _lib, detect as fdet
    dev = out.device
    rg = fdet.lag_ranges(d_mo.reshape(T, -1), N, L, step, n_corr)
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
            float(n_mad), int(N), det_t.data_ptr(), det_l.data_ptr(), det_v.data_ptr(), cap,
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
    # detection is whichever of the two holds the larger cc
    events = [(int(t), int(l)) for t, l in w.get("events", [])]
    hit = sum(1 for t, l in events
              if (t, l // step) in found or (t, -(-l // step)) in found)
    return {"n_mad": n_mad, "min_separation": int(N), "ms": t_ms,
            "median_mad_ms": e0.elapsed_time(e1), "detect_ms": e1.elapsed_time(e2),
            "n_detections": n, "injected_events_detected": "%d/%d" % (hit, len(events)),
            "algorithmic_GB_per_s": 12.0 * n_valid / (t_ms * 1e-3) / 1e9}


def optional_leg(name, fn):
    """Run one of the line's optional legs (detection, FP32 kernel, FP64 spot
    check, screener, CPU baseline): an error there is reported in the line
    instead of ending the run after the headline figure is measured."""
    try:
        return fn()
    except Exception as e:  # the line still carries the measured value
        print("%s leg failed: %s: %s" % (name, type(e).__name__, e), file=sys.stderr, flush=True)
        return {"error": "%s: %s" % (type(e).__name__, e)}


def screened_leg(w, d_tp, d_sst, d_mo, d_w, d_data, out, T, SC, N, L, exact_ms, reps=3):
    """The learned screener on the same day (committed weights): wall time of
    the screened scan (coarse scan, features, network, gate, exact scan of the
    kept items, MAD of the sampled tiles) against the exact step, and the
    recall of the injected events that the exact scan in `out` detects (peak
    cc within +-64 lags above 6 MAD of its row), as the screener's own
    evaluation counts them."""
    import torch
    from fast_matched_filter_amd import screener
    from fast_matched_filter_amd.detect import median_mad
    from fast_matched_filter_amd.screener_train import DETECT_MAD, _event_windows, _window_max
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
            "kept_tiles_equal_exact": same,
            "stage_ms": {k: 1e3 * v for k, v in info["times"].items()},
            "mad_rel_err_max": float(np.abs(me / mads.cpu().numpy() - 1.0).max())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="day")
    ap.add_argument("--kernel", default="bf16x3", choices=sorted(KERNELS))
    ap.add_argument("--fp32-steps", type=int, default=2,
                    help="also time the plain FP32 kernel on the same workload (0 = skip)")
    ap.add_argument("--templates", type=int, default=None, help="override T")
    ap.add_argument("--samples", type=int, default=None, help="override L")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-screened", action="store_true",
                    help="skip the learned screener's leg (step-1 single-length workloads)")
    ap.add_argument("--step", type=int, default=0, help="override the workload's step")
    ap.add_argument("--autotune", action="store_true",
                    help="time the bf16x3 tile shapes on this workload first and keep the best")
    ap.add_argument("--graph", type=int, default=-1, choices=(-1, 0, 1),
                    help="1: time HIP-graph replays of one captured step (value from them); "
                         "0: eager calls; -1 (default): graph for batched workloads")
    ap.add_argument("--persistent-blocks", type=int, default=0,
                    help="batched workloads: persistent grid of about this many blocks (0 = one per item)")
    ap.add_argument("--mfma-tiles", type=int, default=0,
                    help="bf16x3 tile shape: 24, 42, 43, 81 (16x16 tiles) or 212, 412, 413, 421, "
                         "422, 241, 242 (32x32 tiles); 0 = library choice")
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
    t_gen = time.perf_counter()
    w = make_workload(args.workload, **over)
    print("workload %s generated in %.1f s" % (args.workload, time.perf_counter() - t_gen),
          file=sys.stderr, flush=True)
    tp = w["templates"]
    T, S, C, N = tp.shape
    L = w["data"].shape[-1]
    step = args.step or w["step"]
    lengths = np.asarray(w["lengths"], dtype=np.int32)
    batched = bool(np.any(lengths != N))
    n_corr = (L - int(lengths.min())) // step + 1
    kernel = KERNELS[args.kernel]
    if batched and args.kernel != "bf16x3":
        print("templates of several lengths run on the batched bf16x3 launch", file=sys.stderr)
        sys.exit(2)

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
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
        distinct = np.unique(lengths)
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
    for _ in range(args.steps):
        step_fn()
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

    lags, _, _ = lag_counts(w["moveouts"], step, lengths, L, n_corr)
    active = (w["weights"].reshape(T, -1) != 0).sum(axis=1)
    macs = float((lags * active * lengths.astype(np.int64)).sum())
    ms_step = elapsed * 1e3 / args.steps
    value = macs / (elapsed / args.steps)

    # The step captured in a HIP graph (kernel timer off) and replayed: the
    # launch form of the batched workload.  The kernel roofline above comes
    # from the eager steps' HIP events.
    graph_info = None
    if args.graph == 1 or (args.graph == -1 and batched):
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
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g.replay()
        torch.cuda.synchronize()
        g_el = time.perf_counter() - t0
        graph_info = {"ms_per_step": g_el * 1e3 / args.steps,
                      "eager_ms_per_step": ms_step, "value_from": "graph replays"}
        print("graph replays: %.3f ms per step (eager %.3f)" % (g_el * 1e3 / args.steps, ms_step),
              file=sys.stderr, flush=True)
        ms_step = g_el * 1e3 / args.steps
        value = macs / (g_el / args.steps)
    kern_ms = corr_ms / max(corr_n, 1)
    achieved = 2.0 * macs / (kern_ms * 1e-3) / 1e12
    peak = PEAK_TFLOPS[args.kernel]
    energy_bytes = float(sum(S * C * (4.0 * L + 4.0 * (L - int(n) + 1)) for n in np.unique(lengths)))

    detection = None if batched else optional_leg(
        "detection", lambda: detection_leg(lib, out, d_mo, w, T, n_corr, N, L, step))

    fp32 = None
    if args.kernel == "bf16x3" and args.fp32_steps > 0 and not batched:
        def fp32_step():
            matched_filter_device(d_tp, d_sst, d_mo, d_w, d_data, step, N, L, T, S, C,
                                  n_corr, out, normalize=0, network_sum=True,
                                  kernel=KERNELS["fp32"], workspace=ws, stream=stream)

        def fp32_leg():
            fp32_step()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for _ in range(args.fp32_steps):
                fp32_step()
            torch.cuda.synchronize()
            f_el = (time.perf_counter() - t1) / args.fp32_steps
            r = {"value": macs / f_el, "unit": "MAC/s", "ms_per_step": f_el * 1e3,
                 "steps": args.fp32_steps,
                 "achieved_tflops_step": 2.0 * macs / f_el / 1e12,
                 "bf16x3_over_fp32": (macs / (elapsed / args.steps)) / (macs / f_el)}
            print("fp32 kernel: %.3g MAC/s" % r["value"], file=sys.stderr, flush=True)
            return r
        fp32 = optional_leg("fp32 kernel", fp32_leg)
        step_fn()  # leave the bf16x3 result in `out` for the spot check

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
              % (dmax, drms, dn, n_agree, n_ev),
              file=sys.stderr, flush=True)

    screened = None
    if not args.no_screened and args.kernel == "bf16x3" and step == 1 and not batched:
        screened = optional_leg("screened", lambda: screened_leg(
            w, d_tp, d_sst, d_mo, d_w, d_data, out, T, S * C, N, L, elapsed * 1e3 / args.steps))
        if "error" not in screened:
            print("screened scan: %.2f ms vs exact %.2f ms (%.2fx), event recall %.4f"
                  % (screened["ms_per_step"], screened["exact_ms_per_step"],
                     screened["speedup_vs_exact"], screened["event_recall"]),
                  file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        def cpu_leg():
            r, cpu_out = cpu_baseline(w, int(lengths[0]), step)
            gpu0 = out[0, :cpu_out.shape[1]].cpu().numpy()
            r["max_abs_dcc_vs_gpu"] = float(np.abs(gpu0 - cpu_out[0]).max())
            print("cpu baseline: %s" % r["sample"], file=sys.stderr, flush=True)
            return r
        cpu = optional_leg("cpu baseline", cpu_leg)

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
                                  "/".join(str(v) for v in np.unique(lengths)), L, step,
                                  ", one batched launch" if batched else ""),
                   "autotuned_tiles": tuned[0] if tuned else None,
                   "kernel": args.kernel, "macs_per_step": macs,
                   "mfma_tiles_per_wave": int(lib.fmf_get_mfma_tiles(step)) if args.kernel == "bf16x3" else None},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak,
                     "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": MEASURED_TRAFFIC.get(
                         (args.workload, args.kernel,
                          int(lib.fmf_get_mfma_tiles(step)) if args.kernel == "bf16x3" else 0,
                          step, args.templates, args.samples)),
                     "kernel": "corr_%s" % args.kernel, "kernel_ms": kern_ms,
                     "launches": corr_n,
                     # three bf16 products per useful MAC: the useful-work
                     # ceiling of the split form is a third of the dense peak
                     "frac_of_bf16x3_ceiling": (3.0 * achieved / peak
                                                if args.kernel == "bf16x3" else None)},
        "hip_graph": graph_info,
        "normalization_pass": {"kernel": "window_energy", "ms": energy_ms / max(energy_n, 1),
                               "algorithmic_GB_per_s": energy_bytes / (energy_ms / max(energy_n, 1) * 1e-3) / 1e9
                               if energy_n else None},
        "max_abs_dcc_fp64": dmax,
        "rms_abs_dcc_fp64": drms,
        "argmax_agreement_fp64": argmax_agree,
        "detection": detection,
        "fp32_kernel": fp32,
        "screened": screened,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
