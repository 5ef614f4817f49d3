"""Train, calibrate and evaluate the learned screener on synthetic days.

    python -m fast_matched_filter_amd.screener_train [--days 200] [--eval-days 20]
        [--steps 2000] [--out fast_matched_filter_amd/screener_weights.npz]
        [--report report.json] [--scale 1.0]

1. Training days (seeds 1000 + i, the 'day' workload's shapes, injected events
   at amplitude log-uniform in [0.1, 3]): the exact bf16x3 scan gives cc; the
   target is the decimated peak |cc| in units of the template's cc MAD,
   log-compressed (log1p); crops of 1024 screener positions (half centred on
   injected events) of the coarse-correlation features form the training set.
2. Adam on the HIP forward/backward kernels.
3. Calibration (the last few training days, kept out of the crops): the gate
   is the largest threshold on the prediction whose tiles keep 99.7 % of the
   injected events that the exact scan detects (max cc within +-64 lags above
   6 MAD); the same is done for the coarse |cc| feature alone (the method
   without the network).  The threshold that would keep 99.7 % of all
   above-6-MAD samples is reported too: noise crossings of 6 MAD occur in
   every tile, so it keeps almost every tile.
4. Evaluation on held-out days (seeds 5000 + i): full exact scan vs screened
   scan — recall of detected events, of all above-threshold samples and of
   the exact scan's detection list (the same peak picking on both outputs, at
   6 MAD and at the stricter RECALL_MADS thresholds), fraction of tiles
   computed, wall time of both paths; the same event recall for the
   coarse-only gate.
"""
import argparse
import json
import math
import os
import time

import numpy as np

from . import screener

DETECT_MAD = 6.0
# detection thresholds (multiples of the MAD) at which the recall of the
# exact scan's detection list is also reported
RECALL_MADS = (8.0, 9.0, 10.0)
CROP = 1024


def _day_arrays(day, factor):
    import torch
    from .detect import median_mad
    cc = screener.exact_scan(day)
    T, n_corr = cc.shape
    _, mads = median_mad(cc)
    peak = screener.peak_target(cc)
    target = torch.log1p(peak / mads[:, None].clamp_min(1e-12))
    feat, _, _ = screener.day_features(day, n_corr, factor=factor)
    return cc, mads, target, feat


def _crops(feat, target, events, n_crops, gen, device):
    import torch
    T, F, n_out = feat.shape
    xs, ys = [], []
    ev = np.array(events) if events else np.zeros((0, 2), dtype=int)
    for i in range(n_crops):
        if i % 2 == 0 and len(ev):
            t, lag = ev[int(torch.randint(0, len(ev), (1,), generator=gen))]
            c = lag // screener.DEC + int(torch.randint(-CROP // 2 + 64, CROP // 2 - 64, (1,), generator=gen))
        else:
            t = int(torch.randint(0, T, (1,), generator=gen))
            c = int(torch.randint(CROP // 2, n_out - CROP // 2, (1,), generator=gen))
        lo = int(min(max(c - CROP // 2, 0), n_out - CROP))
        xs.append(feat[t, :, lo:lo + CROP])
        ys.append(target[t, lo:lo + CROP])
    return torch.stack(xs), torch.stack(ys)


def _tile_max(score, tile_dec, n_tiles, margin):
    """(T, n_tiles) max of a (T, n_out) score over each tile widened by margin."""
    import torch
    T, n_out = score.shape
    pad = n_tiles * tile_dec + 2 * margin - n_out
    s = torch.nn.functional.pad(score, (margin, max(pad - margin, 0)), value=-1e30)
    win = s.unfold(1, tile_dec + 2 * margin, tile_dec)[:, :n_tiles]
    return win.max(dim=-1).values


def _detection_tiles(cc, mads, tile_out):
    """(T, n_tiles) number of above-threshold samples per tile."""
    import torch
    T, n_corr = cc.shape
    n_tiles = (n_corr + tile_out - 1) // tile_out
    det = (cc > DETECT_MAD * mads[:, None]).float()
    det = torch.nn.functional.pad(det, (0, n_tiles * tile_out - n_corr))
    return det.view(T, n_tiles, tile_out).sum(-1)


def _window_max(cc, events, half=64):
    """max cc within +-half lags of every event (template, lag), on the device."""
    import torch
    ev = torch.as_tensor(np.asarray(events, dtype=np.int64).reshape(-1, 2), device=cc.device)
    off = torch.arange(-half, half + 1, device=cc.device)
    idx = (ev[:, 1:2] + off[None, :]).clamp(0, cc.shape[1] - 1)
    return cc[ev[:, 0:1].expand_as(idx), idx].max(dim=1).values


def _event_windows(cc, mads, events, tile_out, half=64):
    """Injected events the exact scan detects (max cc within +-half lags of
    the injection above 6 MAD): list of (template, lag, tile)."""
    if not events:
        return []
    mx = _window_max(cc, events, half).cpu().numpy()
    m = mads.cpu().numpy()
    return [(t, lag, lag // tile_out) for (t, lag), v in zip(events, mx)
            if v > DETECT_MAD * m[t]]


def detection_recall(cc, cs, mads_est, moveouts, n_tp, n_data, step, sep, tol=8,
                     events=(), event_half=64, n_mad=DETECT_MAD):
    """Recall of the exact scan's detection list by the screened scan.

    Both lists come from the same peak picking (``detect.detect``: cc above
    0 + n_mad MAD and the first maximum within +-`sep` lags, over each template's
    valid lags): the exact one on the exact cc with the rows' exact MADs, the
    screened one on the screened output with the screened scan's own MAD
    estimates (its median is not estimated; the cc median of noise is ~1e-3
    of a MAD).  An exact detection counts as recalled when the screened list
    holds a detection of the same template within +-`tol` lags.  Returned:
    counts, the recall, and the recall of the exact detections that lie
    within +-`event_half` lags of an injected event (output index
    lag // step)."""
    import torch
    from .detect import detect, lag_ranges, median_mad
    T, n_corr = cc.shape
    rg = lag_ranges(moveouts.reshape(T, -1), n_tp, n_data, step, n_corr)
    zero = torch.zeros(T, dtype=torch.float32, device=cc.device)
    _, mads = median_mad(cc, rg)
    ex = detect(cc, n_mad=n_mad, min_separation=sep, ranges=rg, median=zero, mad=mads)
    sc = detect(cs, n_mad=n_mad, min_separation=sep, ranges=rg, median=zero,
                mad=mads_est.to(torch.float32))
    big = np.int64(n_corr + 4 * tol + 4)
    kx = ex["template"].astype(np.int64) * big + ex["lag"].astype(np.int64)
    ks = np.sort(sc["template"].astype(np.int64) * big + sc["lag"].astype(np.int64))
    if ks.size:
        i = np.clip(np.searchsorted(ks, kx), 0, ks.size - 1)
        j = np.clip(i - 1, 0, ks.size - 1)
        near = np.minimum(np.abs(ks[i] - kx), np.abs(ks[j] - kx))
        hit = near <= tol
    else:
        hit = np.zeros(kx.size, dtype=bool)
    on_event = np.zeros(kx.size, dtype=bool)
    if len(events) and kx.size:
        ke = np.sort(np.asarray([t * big + lag // step for t, lag in events], dtype=np.int64))
        i = np.clip(np.searchsorted(ke, kx), 0, ke.size - 1)
        j = np.clip(i - 1, 0, ke.size - 1)
        on_event = np.minimum(np.abs(ke[i] - kx), np.abs(ke[j] - kx)) <= event_half
    n_ev = int(on_event.sum())
    return {"n_exact": int(kx.size), "n_screened": int(ks.size),
            "recall": float(hit.mean()) if kx.size else 1.0,
            "n_exact_on_events": n_ev,
            "recall_events": float(hit[on_event].mean()) if n_ev else 1.0}


def calibrate_events(tile_scores, ev_lists, recall=0.997):
    """Largest threshold whose tiles keep `recall` of the detected events."""
    s = np.array([ts[t, k] for ts, evs in zip(tile_scores, ev_lists) for t, _, k in evs])
    if s.size == 0:
        return -1e30
    return float(np.quantile(s, 1.0 - recall)) - 1e-6


def calibrate(scores, dets, recall=0.997):
    """Largest threshold keeping `recall` of the detection samples: dets are
    per-tile sample counts, scores per-tile maxima (lists over days)."""
    s = np.concatenate([x.reshape(-1) for x in scores])
    d = np.concatenate([x.reshape(-1) for x in dets])
    m = d > 0
    s, d = s[m], d[m]
    order = np.argsort(s)
    cum = np.cumsum(d[order]) / d.sum()
    k = np.searchsorted(cum, 1.0 - recall, side="right")
    return float(s[order][min(k, len(s) - 1)]) - 1e-6


def main(argv=None):
    import torch
    from .detect import median_mad
    ap = argparse.ArgumentParser()
    ap.add_argument("--days", type=int, default=200)
    ap.add_argument("--eval-days", type=int, default=20)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--crops-per-day", type=int, default=128)
    ap.add_argument("--scale", type=float, default=1.0,
                    help="fraction of the day's length (tests use small days)")
    ap.add_argument("--templates", type=int, default=100)
    ap.add_argument("--coarse", type=int, default=screener.COARSE,
                    help="decimation of the coarse scan (1..16)")
    ap.add_argument("--out", default=None)
    ap.add_argument("--report", default=None)
    ap.add_argument("--weights", default=None,
                    help="evaluate these trained weights (npz written by --out) on the held-out "
                         "days, without training or calibration")
    ap.add_argument("--sample-fraction", type=float, default=0.02,
                    help="fraction of whole tiles per template scanned exactly at random for "
                         "the MAD estimate")
    ap.add_argument("--item-shape", type=int, default=screener.ITEM_SHAPE,
                    help="MFMA tile shape of the screened exact scan (its tile = one item)")
    ap.add_argument("--time-only", type=int, default=0, metavar="REPS",
                    help="with --weights: only time REPS screened scans of held-out day 5000 "
                         "(after one untimed scan) and print the stage times")
    args = ap.parse_args(argv)
    dev = torch.device("cuda:0")
    L = int(8_640_000 * args.scale)
    shapes = dict(n_templates=args.templates, n_data=L)
    gen = torch.Generator().manual_seed(0)
    lb = screener.lib()
    tile_out = lb.fmf_mfma_item_outputs(1)
    tile_dec = tile_out // screener.DEC
    margin = 16

    if args.weights:
        w = np.load(args.weights)
        net = screener.Network(dev, seed=0)
        net.load(w["w16"])
        gate, gate_feat = float(w["gate"]), float(w["gate_coarse_only"])
        if "coarse" in w and int(w["coarse"]) != args.coarse:
            raise SystemExit("weights were trained with --coarse %d" % int(w["coarse"]))
        gate_all = keep_cal = keep_cal_feat = None
        losses = []
        args.days = args.steps = 0
        print("evaluating %s: gate %.4f, coarse-only gate %.4f" % (args.weights, gate, gate_feat),
              flush=True)
        if args.time_only:
            day = screener.synthetic_day(5000, dev, **shapes)
            reps = []
            for rep in range(args.time_only + 1):
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                _, _, _, _, info = screener.screened_scan(
                    day, net, gate, seed=rep, factor=args.coarse,
                    sample_fraction=args.sample_fraction, item_shape=args.item_shape)
                torch.cuda.synchronize()
                row = dict(rep=rep, screened_ms=1e3 * (time.perf_counter() - t1),
                           kept_fraction=info["kept_fraction"],
                           sampled_fraction=info["sampled_fraction"],
                           stage_ms={k: 1e3 * v for k, v in info["times"].items()})
                print(json.dumps(row), flush=True)
                if rep:
                    reps.append(row)
            return dict(weights=args.weights, reps=reps)
    else:
        n_cal = max(1, min(5, args.days // 10))
        X, Y = [], []
        cal_scores, cal_feat_scores, cal_dets = [], [], []
        cal_days = []
        t0 = time.perf_counter()
        for i in range(args.days):
            day = screener.synthetic_day(1000 + i, dev, **shapes)
            cc, mads, target, feat = _day_arrays(day, args.coarse)
            if i >= args.days - n_cal:
                n_tiles = (cc.shape[1] + tile_out - 1) // tile_out
                cal_days.append((feat, _detection_tiles(cc, mads, tile_out).cpu().numpy(),
                                 _event_windows(cc, mads, day["events"], tile_out), n_tiles))
            else:
                x, y = _crops(feat, target, day["events"], args.crops_per_day, gen, dev)
                X.append(x)
                Y.append(y)
            del day, cc
            if (i + 1) % 20 == 0:
                print("training day %d/%d (%.0f s)" % (i + 1, args.days, time.perf_counter() - t0), flush=True)
        X = torch.cat(X)
        Y = torch.cat(Y)
        print("training set: %d crops x %d" % (X.shape[0], CROP), flush=True)

        net = screener.Network(dev, seed=0)
        losses = []
        for step in range(args.steps):
            idx = torch.randint(0, X.shape[0], (args.batch,), generator=gen).to(dev)
            xb, yb = X[idx].contiguous(), Y[idx].contiguous()
            pred, acts = net.forward(xb, keep=True)
            loss, gy = net.loss_and_grad(pred, yb, boost=1.0)
            lr = 2e-3 * (0.5 * (1 + math.cos(math.pi * step / args.steps)))
            net.adam(net.backward(acts, gy), lr=lr)
            if step % 100 == 0 or step == args.steps - 1:
                losses.append((step, float(loss)))
                print("step %d loss %.5f" % (step, float(loss)), flush=True)

        # calibration: network score vs the best single feature (no network)
        cal_events = []
        for feat, dets, evs, n_tiles in cal_days:
            pred = net.infer(feat)  # as in the screened scan
            cal_scores.append(_tile_max(pred, tile_dec, n_tiles, margin).cpu().numpy())
            # the coarse |cc| feature alone (no network)
            cal_feat_scores.append(_tile_max(feat[:, 0].contiguous(), tile_dec, n_tiles, margin).cpu().numpy())
            cal_dets.append(dets)
            cal_events.append(evs)
        # gate on the detected injected events (noise crossings of 6 MAD are
        # spread over every tile and no screener can anticipate them)
        gate = calibrate_events(cal_scores, cal_events)
        gate_feat = calibrate_events(cal_feat_scores, cal_events)
        gate_all = calibrate(cal_scores, cal_dets)
        keep_cal = float(np.mean(np.concatenate([s.reshape(-1) for s in cal_scores]) > gate))
        keep_cal_feat = float(np.mean(np.concatenate([s.reshape(-1) for s in cal_feat_scores]) > gate_feat))
        print("gate %.4f keeps %.3f of tiles; coarse-only gate %.4f keeps %.3f" %
              (gate, keep_cal, gate_feat, keep_cal_feat), flush=True)
        del X, Y, cal_days

    # evaluation on held-out days
    rows = []
    for i in range(args.eval_days):
        day = screener.synthetic_day(5000 + i, dev, **shapes)
        if i == 0:  # one untimed pass of both paths (first-call set-up)
            screener.exact_scan(day)
            screener.screened_scan(day, net, gate, seed=i, factor=args.coarse,
                                   sample_fraction=args.sample_fraction,
                                   item_shape=args.item_shape)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        cc = screener.exact_scan(day)
        torch.cuda.synchronize()
        t_full = time.perf_counter() - t1
        _, mads = median_mad(cc)
        det = cc > DETECT_MAD * mads[:, None]
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        cs, computed, mads_est, pred, info = screener.screened_scan(
            day, net, gate, seed=i, factor=args.coarse, sample_fraction=args.sample_fraction,
            item_shape=args.item_shape)
        torch.cuda.synchronize()
        t_scr = time.perf_counter() - t2
        det_s = cs > DETECT_MAD * mads_est[:, None]
        n_det = int(det.sum())
        hit = int((det & det_s).sum())
        evs = _event_windows(cc, mads, day["events"], tile_out)
        me = mads_est.cpu().numpy()
        mad_rel = np.abs(me / mads.cpu().numpy() - 1.0)
        mx_s = _window_max(cs, [(t, lag) for t, lag, _ in evs]).cpu().numpy() if evs else []
        ev_hit = sum(1 for (t, lag, _), v in zip(evs, mx_s) if v > DETECT_MAD * me[t])
        drec = detection_recall(cc, cs, mads_est, day["moveouts"], day["n_tp"], L, 1,
                                sep=day["n_tp"], events=day["events"])
        # the same at detection thresholds where noise crossings are rarer
        dhi = {k: detection_recall(cc, cs, mads_est, day["moveouts"], day["n_tp"], L, 1,
                                   sep=day["n_tp"], events=day["events"], n_mad=k)
               for k in RECALL_MADS}
        # the coarse-only method (no network) at its own calibrated gate
        feat, _, _ = screener.day_features(day, cc.shape[1], factor=args.coarse)
        n_tiles = (cc.shape[1] + tile_out - 1) // tile_out  # baseline tiles: the default shape's
        fmax = _tile_max(feat[:, 0].contiguous(), tile_dec, n_tiles, margin)
        dt = _detection_tiles(cc, mads, tile_out)
        keep_feat = fmax > gate_feat
        rec_feat = float((dt * keep_feat).sum() / dt.sum().clamp_min(1))
        kf = keep_feat.cpu().numpy()
        ev_feat = sum(1 for t, lag, k in evs if kf[t, k])
        rows.append(dict(day=5000 + i, detections=n_det, recall=hit / max(n_det, 1),
                         mad_rel_err_max=float(mad_rel.max()),
                         mad_rel_err_rms=float(np.sqrt(np.mean(mad_rel ** 2))),
                         events_detected=len(evs), event_recall=ev_hit / max(len(evs), 1),
                         exact_detections=drec["n_exact"], screened_detections=drec["n_screened"],
                         detection_recall_vs_exact=drec["recall"],
                         exact_detections_on_events=drec["n_exact_on_events"],
                         detection_recall_on_events=drec["recall_events"],
                         detection_recall_by_mad={str(k): dict(
                             recall=v["recall"], exact=v["n_exact"],
                             exact_on_events=v["n_exact_on_events"],
                             recall_on_events=v["recall_events"]) for k, v in dhi.items()},
                         coarse_only_event_recall=ev_feat / max(len(evs), 1),
                         kept_fraction=info["kept_fraction"],
                         sampled_fraction=info["sampled_fraction"], fallback=info["fallback"],
                         full_s=t_full, screened_s=t_scr, speedup=t_full / t_scr,
                         stage_s=info["times"],
                         coarse_only_kept_fraction=float(keep_feat.float().mean()),
                         coarse_only_recall=rec_feat))
        print(json.dumps(rows[-1]), flush=True)
        del day, cc, cs
    rep = dict(weights=args.weights, days=args.days, eval_days=args.eval_days, steps=args.steps,
               scale=args.scale, sample_fraction=args.sample_fraction,
               mad_rel_err_max=float(max(r["mad_rel_err_max"] for r in rows)),
               mad_rel_err_rms=float(np.sqrt(np.mean([r["mad_rel_err_rms"] ** 2 for r in rows]))),
               coarse=args.coarse,
               templates=args.templates, gate=gate, gate_coarse_only=gate_feat,
               gate_all_samples=gate_all,
               event_recall=float(sum(r["event_recall"] * r["events_detected"] for r in rows) /
                                  max(1, sum(r["events_detected"] for r in rows))),
               coarse_only_event_recall=float(np.mean([r["coarse_only_event_recall"] for r in rows])),
               detection_recall_vs_exact=float(
                   sum(r["detection_recall_vs_exact"] * r["exact_detections"] for r in rows) /
                   max(1, sum(r["exact_detections"] for r in rows))),
               detection_recall_vs_exact_per_day=[r["detection_recall_vs_exact"] for r in rows],
               detection_recall_by_mad={
                   str(k): dict(
                       recall=float(sum(r["detection_recall_by_mad"][str(k)]["recall"] *
                                        r["detection_recall_by_mad"][str(k)]["exact"]
                                        for r in rows) /
                                    max(1, sum(r["detection_recall_by_mad"][str(k)]["exact"]
                                               for r in rows))),
                       exact_per_day=float(np.mean(
                           [r["detection_recall_by_mad"][str(k)]["exact"] for r in rows])),
                       exact_on_events_per_day=float(np.mean(
                           [r["detection_recall_by_mad"][str(k)]["exact_on_events"]
                            for r in rows])),
                       min_day=float(min(r["detection_recall_by_mad"][str(k)]["recall"]
                                         for r in rows)))
                   for k in RECALL_MADS},
               detection_recall_on_events=float(
                   sum(r["detection_recall_on_events"] * r["exact_detections_on_events"]
                       for r in rows) / max(1, sum(r["exact_detections_on_events"] for r in rows))),
               calibration_keep=keep_cal, calibration_keep_coarse_only=keep_cal_feat,
               losses=losses,
               recall=float(sum(r["recall"] * r["detections"] for r in rows) /
                            max(1, sum(r["detections"] for r in rows))),
               kept_fraction=float(np.mean([r["kept_fraction"] for r in rows])),
               sampled_fraction=float(np.mean([r["sampled_fraction"] for r in rows])),
               speedup=float(np.sum([r["full_s"] for r in rows]) / np.sum([r["screened_s"] for r in rows])),
               coarse_only_kept_fraction=float(np.mean([r["coarse_only_kept_fraction"] for r in rows])),
               coarse_only_recall=float(np.mean([r["coarse_only_recall"] for r in rows])),
               rows=rows)
    print(json.dumps({k: v for k, v in rep.items() if k != "rows"}), flush=True)
    if args.out:
        np.savez(args.out, w16=net.w16.cpu().numpy(), gate=np.float32(gate),
                 gate_coarse_only=np.float32(gate_feat), coarse=np.int32(args.coarse))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(rep, f, indent=1)
    return rep


if __name__ == "__main__":
    main()
